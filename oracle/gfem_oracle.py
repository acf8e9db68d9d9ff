"""CPU oracle: an FP64 NumPy/SciPy restatement of the gfem (EXFEM 0.3.0) solver hot path.

TEST INFRASTRUCTURE ONLY. This module is the checker: only the tests, the smoke check and the
benchmark's CPU-baseline leg may import it. The product package (``exfem_amd``) never imports,
links or calls anything here, and nothing here is ever used as the measured path.

Every function restates the reference algorithm at the cited ``gfem/...`` file:line. The
reference itself (JAX + PyAMG) cannot run in this environment, so the restatement is pinned by
closed forms, invariants and independent routes instead (tests/test_oracle.py):

=====================  ==========================================================  ==========
function               pinned by                                                   status
=====================  ==========================================================  ==========
element tables         dN == central differences of N; partition of unity;         pinned
                       Gauss rules integrate the exact polynomials
element_stiffness      symmetry; 6 rigid-body modes in the null space; exact       pinned
                       constant-strain patch test (tests/test_verification.py:27
                       closed form); parallelepiped B-bar == selective reduced
                       integration closed form; Ke == d(fe)/du by central
                       differences of the independent force routine
element_force          fe(u) == Ke u (linear material) and the patch test;          pinned
                       with fields: free thermal expansion gives fe = 0, a clamped
                       heated block pushes -3 K alpha dT A on a face, eigenstrain
                       eps0 I == thermal strain alpha dT = eps0
ops_matvec /           == scipy CSR (ops_assemble) @ x; scatter == gather order;   pinned
gather_matvec          27-point stencil of a uniform unit-cube mesh
pcg                    textbook CG identity vs scipy.sparse.linalg.spsolve;        pinned
                       exact stopping rule and iteration count on a diagonal
                       system whose CG iterate count is known in closed form
cheb                   error propagator == T_3((theta-lam)/delta)/T_3(theta/delta)  pinned
                       on the eigenpairs of D^-1/2 A D^-1/2 (numpy Chebyshev
                       series); a [rho/10, rho] interval or another degree fails
cheby_precond          M A == I - T_7(...)/T_7(...) on [lmax/30, lmax]; M is SPD    pinned
lmax_power /           == 1.1 |M^k v0| / |M^(k-1) v0| from the eigenvalue powers    pinned
rho_estimate           (k = 25 from the sine start, k = 15 from default_rng(0));
                       k +- 1 or another start fails; -> 1.1 lam_max when separated
vcycle                 2-level hierarchy with exact coarse solve, dense route       pinned
mixed_solve            final FP64 residual <= tol, error vs direct                 pinned
multi_load_cases       each column == its single-RHS pcg                           pinned
traction_vector        total force == t * area, corner/edge/interior closed form   pinned
transformation /       a chain u_a = u_b = 0.5 u_m resolved; cycles broken;        pinned
op_bc_mpc              prescribed slaves dropped; two meshes coupled at coincident
                       interface nodes solve to the merged mesh's displacements;
                       a normal constraint on an axis plane == fixing that dof
spring_element /       force == central differences of the restated energy;        pinned
contact_pair /         stiffness == differences of the force; dg == central
wall_evaluate          differences of the gap; pn = max(lam - k g, 0) branches
contact_select         the nearest clamped projection == brute force over facets   pinned
J2Plasticity.ret       uniaxial-stress closed form E (sy0 + H' eps)/(E + H') with   pinned
                       H' = H + Hk; yield consistency q(s - beta) == sy(alpha)
                       after the return; plastic incompressibility; elastic
                       branch == the linear material
element_force_state /  J2 with an unreachable yield stress == the linear           pinned
element_stiffness_     element routines; Ke == d(fe)/du by central differences
state                  of element_force_state at a plastic state
hyper_pk1              == complex-step dW/dF of hyper_energy (every material)       pinned
hyper_element_force    == complex-step gradient of the element energy              pinned
hyper_element_         symmetric; == central differences of the force; at u = 0 == pinned
stiffness              the linear element_stiffness (B-bar for C3D8) with the
                       material's small-strain mu, lambda
element_cauchy         uniaxial homogeneous strain: sigma_xx = E eps, others 0;      pinned
                       NeoHookean homogeneous stretch == its closed-form Cauchy
                       stress 2 C10 J^-5/3 dev(B) + 2 (J - 1)/D1 I
geo_stiff              u^T K u == sum_g vol tr(H sig H^T) with H = grad u (the      pinned
                       second-order work of the initial stress); rigid translations
                       in the null space
buckling_factors       a clamped column under end compression: first load factor   pinned
                       within 3% of Euler's pi^2 E I / (4 L^2)
=====================  ==========================================================  ==========

The smoothed-aggregation set-up that gfem/linalg.py:103-136 delegates to PyAMG (an optional
dependency, unpinned in the reference's pyproject.toml:16, absent here) is restated from PyAMG's
published algorithm (``sa_*`` below) and pinned by:

=====================  ==========================================================  ==========
sa_standard_aggregation PyAMG's own documented examples (4-vertex 1D Poisson -> two   pinned
                       aggregates; an isolated first vertex stays out); hand-worked
                       chain, 3x3 grid and directed graphs that reach pass 3
sa_fit_candidates      PyAMG's documented examples (constant and constant + linear   pinned
                       candidates, Q and R as listed); rank-deficient aggregate:
                       orthonormal columns, zero columns, T R == B
sa_smooth_prolongator  hand-computed P on the 4-node chain (Gershgorin row sums;     pinned
                       the diagonal would give another first row)
sa_hierarchy           free-floating box: P Bc == B and A_c Bc == 0 at every level;  pinned
                       clamped box: I - P A_c^-1 P^T A is an A-orthogonal projector
                       (fails for a coarse operator formed from T)
=====================  ==========================================================  ==========
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

I3 = np.eye(3)

# ----------------------------------------------------------------------------- element tables
# gfem/elements.py:113-126 (HEX8 node order, 2x2x2 Gauss points at H/sqrt(3) in node order,
# weights 1, B-bar centre derivative at xi=0) and gfem/elements.py:156-179 (TET10).
_H = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
               [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)


def hex8_N(p):
    """gfem/elements.py:120-121."""
    p = np.asarray(p, float)
    return 0.125 * (1 + _H[:, 0] * p[0]) * (1 + _H[:, 1] * p[1]) * (1 + _H[:, 2] * p[2])


def hex8_dN(p):
    """d N_a / d xi_i of hex8_N (the reference tabulates jax.jacfwd, gfem/elements.py:11-15)."""
    p = np.asarray(p, float)
    f = 1 + _H * p[None, :]                                   # (8,3)
    d = np.empty((8, 3))
    d[:, 0] = 0.125 * _H[:, 0] * f[:, 1] * f[:, 2]
    d[:, 1] = 0.125 * _H[:, 1] * f[:, 0] * f[:, 2]
    d[:, 2] = 0.125 * _H[:, 2] * f[:, 0] * f[:, 1]
    return d


_EDGES10 = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]  # gfem/elements.py:167
_dL = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)


def tet10_N(p):
    """gfem/elements.py:160-172: corners L(2L-1), mid-edge 4 Li Lj."""
    p = np.asarray(p, float)
    L = np.array([1 - p[0] - p[1] - p[2], p[0], p[1], p[2]])
    return np.concatenate([L * (2 * L - 1), [4 * L[i] * L[j] for i, j in _EDGES10]])


def tet10_dN(p):
    p = np.asarray(p, float)
    L = np.array([1 - p[0] - p[1] - p[2], p[0], p[1], p[2]])
    d = np.empty((10, 3))
    for a in range(4):
        d[a] = (4 * L[a] - 1) * _dL[a]
    for k, (i, j) in enumerate(_EDGES10):
        d[4 + k] = 4 * (L[j] * _dL[i] + L[i] * _dL[j])
    return d


class ElementType:
    """Tabulated element (gfem/elements.py:18-36)."""

    def __init__(self, name, nnode, gp, gw, Nfun, dNfun, fbar=False, center=None):
        self.name, self.nnode = name, nnode
        self.gp, self.gw = np.asarray(gp, float), np.asarray(gw, float)
        self.ngp = len(self.gw)
        self.Nfun, self.dNfun = Nfun, dNfun
        self.N = np.array([Nfun(p) for p in self.gp])
        self.dN = np.array([dNfun(p) for p in self.gp])
        self.fbar = fbar
        self.dN0 = dNfun(center) if center is not None else None


HEX8 = ElementType("C3D8", 8, _H / np.sqrt(3), np.ones(8), hex8_N, hex8_dN, fbar=True,
                   center=[0, 0, 0])                                    # gfem/elements.py:124-125
HEX8_PLAIN = ElementType("C3D8_FULL", 8, _H / np.sqrt(3), np.ones(8), hex8_N, hex8_dN)  # :126
_a, _b = 0.5854101966249685, 0.1381966011250105
TET10 = ElementType("C3D10", 10, [[_b, _b, _b], [_a, _b, _b], [_b, _a, _b], [_b, _b, _a]],
                    [1 / 24] * 4, tet10_N, tet10_dN)                    # gfem/elements.py:175-179
ELEMENTS = {"C3D8": HEX8, "C3D8_FULL": HEX8_PLAIN, "C3D10": TET10}


# ----------------------------------------------------------------------------- material
def lame(E, nu):
    """gfem/materials.py:18-19 -> (mu, lambda)."""
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def pk1_linear(F, E, nu):
    """gfem/materials.py:66-68: eps = sym(F) - I, sigma = lam tr(eps) I + 2 mu eps. F (...,3,3)."""
    mu, lam = lame(E, nu)
    eps = 0.5 * (F + np.swapaxes(F, -1, -2)) - I3
    tr = np.trace(eps, axis1=-2, axis2=-1)
    return lam * tr[..., None, None] * I3 + 2 * mu * eps


def material_tangent(E, nu):
    """A[i,j,l,m] = dP_ij/dF_lm (jax.jacfwd of the stress, gfem/model.py:99). pk1 is affine in F,
    so evaluating it on unit perturbations is exact."""
    P0 = pk1_linear(I3, E, nu)
    A = np.zeros((3, 3, 3, 3))
    for l in range(3):
        for m in range(3):
            dF = np.zeros((3, 3))
            dF[l, m] = 1.0
            A[:, :, l, m] = pk1_linear(I3 + dF, E, nu) - P0
    return A


# ----------------------------------------------------------------------------- J2 plasticity
class J2Plasticity:
    """gfem/materials.py:228-300: small-strain von Mises plasticity with combined hardening
    (isotropic: the (yield stress, plastic strain) table extended linearly to eps_p = 1e6, or
    Johnson-Cook at T = 0; kinematic: linear Prager, modulus Hk), radial return with a 25-step
    local Newton solve on dgamma. State per Gauss point: [eps_p (9), alpha, backstress (9)]."""
    nstate = 19
    small_strain = True

    def __init__(self, E, nu, table=None, Hk=0.0, johnson_cook=None):
        self.E, self.nu = float(E), float(nu)
        self.mu, self.lam = lame(E, nu)
        self.K = self.lam + 2 * self.mu / 3
        self.Hk = float(Hk)
        self.jc = johnson_cook
        if self.jc is None:
            t = np.asarray(table, dtype=np.float64).reshape(-1, 2)
            if t.shape[0] == 1:                                   # perfect plasticity
                t = np.array([[t[0, 0], 0.0], [t[0, 0], 1e3]])
            sy_tab, ep_tab = t[:, 0], t[:, 1]
            slope = (sy_tab[-1] - sy_tab[-2]) / max(ep_tab[-1] - ep_tab[-2], 1e-30)
            self.ep_tab = np.append(ep_tab, 1e6)
            self.sy_tab = np.append(sy_tab, sy_tab[-1] + slope * (1e6 - ep_tab[-1]))

    def _seg(self, a):
        """Segment jnp.interp evaluates at a: searchsorted(side='right') clipped to [1, n-1]."""
        return int(np.clip(np.searchsorted(self.ep_tab, a, side="right"), 1, len(self.ep_tab) - 1))

    def sy(self, a):
        """gfem/materials.py:259-267 at T = 0 (no thermal softening)."""
        if self.jc is not None:
            c = self.jc
            return c["A"] + c["B"] * (a + 1e-9) ** c["n"]
        xp, fp = self.ep_tab, self.sy_tab
        if a < xp[0]:
            return fp[0]
        if a > xp[-1]:
            return fp[-1]
        i = self._seg(a)
        dx = xp[i] - xp[i - 1]
        return fp[i - 1] + (a - xp[i - 1]) / dx * (fp[i] - fp[i - 1])

    def dsy(self, a):
        """d sy / d alpha (jax.grad of sy, gfem/materials.py:282)."""
        if self.jc is not None:
            c = self.jc
            return c["B"] * c["n"] * (a + 1e-9) ** (c["n"] - 1)
        xp, fp = self.ep_tab, self.sy_tab
        if a < xp[0] or a > xp[-1]:
            return 0.0
        i = self._seg(a)
        return (fp[i] - fp[i - 1]) / (xp[i] - xp[i - 1])

    def ret(self, eps, s):
        """_return, gfem/materials.py:269-297: one Gauss point. eps (3,3), s (19,)."""
        ep = s[:9].reshape(3, 3)
        a0 = s[9]
        beta = s[10:19].reshape(3, 3)
        ee = eps - ep
        tr = np.trace(ee)
        dev = ee - tr / 3 * I3
        s_tr = 2 * self.mu * dev
        xi = s_tr - beta
        q_tr = np.sqrt(1.5) * np.sqrt(np.sum(xi * xi) + 1e-30)
        f_tr = q_tr - self.sy(a0)
        H3 = 3 * self.mu + self.Hk
        dg = 0.0
        for _ in range(25):
            g = q_tr - H3 * dg - self.sy(a0 + dg)
            dg = max(dg + g / (H3 + self.dsy(a0 + dg)), 0.0)
        if not f_tr > 0:
            dg = 0.0
        n = 1.5 * xi / (q_tr + 1e-30)
        s_dev = s_tr - 2 * self.mu * dg * n
        sig = s_dev + self.K * tr * I3
        ep_new = ep + dg * n
        beta_new = beta + (2.0 / 3.0) * self.Hk * dg * n
        return sig, np.concatenate([ep_new.reshape(-1), [a0 + dg], beta_new.reshape(-1)])

    def pk1(self, F, s):
        """gfem/materials.py:299-303: the return on eps = sym(F) - I."""
        return self.ret(0.5 * (F + F.T) - I3, s)

    def tangent(self, F, s, h=1e-8):
        """A[i,j,l,m] = dP_ij/dF_lm (jax.jacfwd of the stress, gfem/model.py:99) by central
        differences in the 9 components of F."""
        A = np.zeros((3, 3, 3, 3))
        for l in range(3):
            for m in range(3):
                dF = np.zeros((3, 3))
                dF[l, m] = h
                A[:, :, l, m] = (self.pk1(F + dF, s)[0] - self.pk1(F - dF, s)[0]) / (2 * h)
        return A


# ----------------------------------------------------------------------------- element kernels
def geom(et, Xe):
    """gfem/model.py:35-38. Xe (ne,nnode,3) -> dNdX (ne,ngp,nnode,3), vol (ne,ngp)."""
    Jac = np.einsum("ean,gai->egni", Xe, et.dN)
    Jinv = np.linalg.inv(Jac)
    dNdX = np.einsum("gai,egij->egaj", et.dN, Jinv)
    return dNdX, et.gw[None, :] * np.linalg.det(Jac)


def defgrad(et, Xe, ue, dNdX):
    """gfem/model.py:40-52 (small-strain B-bar branch :48-51)."""
    F = I3 + np.einsum("eai,egaj->egij", ue, dNdX)
    if et.fbar:
        J0 = np.einsum("ean,ai->eni", Xe, et.dN0)
        b0 = np.einsum("ai,eij->eaj", et.dN0, np.linalg.inv(J0))
        F0 = I3 + np.einsum("eai,eaj->eij", ue, b0)
        tr0 = np.trace(F0, axis1=1, axis2=2) - 3.0
        tr = np.trace(F, axis1=2, axis2=3) - 3.0
        F = F + ((tr0[:, None] - tr) / 3.0)[:, :, None, None] * I3
    return F


def defgrad_jacobian(et, Xe, dNdX):
    """G[e,g,i,j,a,k] = dF_ij/du_ak (jax.jacfwd(Fmap), gfem/model.py:96-98). defgrad is affine
    in ue, so columns are defgrad(unit) - defgrad(0)."""
    ne, nn = Xe.shape[0], et.nnode
    F0 = defgrad(et, Xe, np.zeros((ne, nn, 3)), dNdX)
    G = np.zeros((ne, et.ngp, 3, 3, nn, 3))
    for a in range(nn):
        for k in range(3):
            u = np.zeros((ne, nn, 3))
            u[:, a, k] = 1.0
            G[..., a, k] = defgrad(et, Xe, u, dNdX) - F0
    return G


def field_shift(et, ne, Te=None, T0e=None, alpha=0.0, eig=None):
    """The small-strain field of stress_gp (gfem/model.py:54-64): th I + eig per Gauss point,
    th = alpha (N.Te - N.T0e) with Te, T0e the element's nodal temperatures (ne, nnode) and eig
    (ne, 9) or (ne, 3, 3). (ne, ngp, 3, 3), zeros without fields."""
    out = np.zeros((ne, et.ngp, 3, 3))
    if Te is not None and alpha:
        T0e = np.zeros_like(Te) if T0e is None else T0e
        th = alpha * (np.einsum("ga,ea->eg", et.N, Te) - np.einsum("ga,ea->eg", et.N, T0e))
        out += th[:, :, None, None] * I3
    if eig is not None:
        out += np.asarray(eig, float).reshape(ne, 1, 3, 3)
    return out


def element_force(et, Xe, ue, E, nu, scale=None, Te=None, T0e=None, alpha=0.0, eig=None):
    """Block.force, gfem/model.py:70-81 (linear elastic): P = pk1(F - th I - eig) with the
    thermal strain th = alpha (T - T0) at the Gauss points and the eigenstrain eig of stress_gp
    (:54-64; none by default). B-bar: fe = vjp(defgrad)(P * sc * vol); plain:
    einsum("gij,gaj,g->ai")."""
    dNdX, vol = geom(et, Xe)
    sc = np.ones(len(Xe)) if scale is None else np.asarray(scale, float)
    wv = sc[:, None] * vol
    F = defgrad(et, Xe, ue, dNdX) - field_shift(et, len(Xe), Te, T0e, alpha, eig)
    P = pk1_linear(F, E, nu)
    if et.fbar:
        G = defgrad_jacobian(et, Xe, dNdX)
        return np.einsum("egij,egijak->eak", P * wv[:, :, None, None], G)
    return np.einsum("egij,egaj,eg->eai", P, dNdX, wv)


def element_stiffness(et, Xe, E, nu, scale=None, chunk=2048):
    """Block.stiff, gfem/model.py:83-111 (linear elastic): B-bar K = sum_g wv G^T A G
    (:101), plain K = einsum("gaj,gijkl,gbl,g->aibk") (:110). Returns (ne, 3nn, 3nn)."""
    Xe = np.asarray(Xe, float)
    ne, nn = Xe.shape[0], et.nnode
    A = material_tangent(E, nu)
    sc_all = np.ones(ne) if scale is None else np.asarray(scale, float)
    out = np.empty((ne, 3 * nn, 3 * nn))
    for s in range(0, ne, chunk):
        X = Xe[s:s + chunk]
        dNdX, vol = geom(et, X)
        wv = sc_all[s:s + chunk, None] * vol
        if et.fbar:
            G = defgrad_jacobian(et, X, dNdX)
            K = np.einsum("egijak,ijlm,eglmbn,eg->eakbn", G, A, G, wv, optimize=True)
        else:
            K = np.einsum("egaj,ijkl,egbl,eg->eaibk", dNdX, A, dNdX, wv, optimize=True)
        out[s:s + chunk] = K.reshape(len(X), 3 * nn, 3 * nn)
    return out


def element_force_state(et, Xe, ue, se, mat, scale=None, Te=None, T0e=None, eig=None):
    """Block.force with a state-carrying material (gfem/model.py:70-81; stress_gp :54-64, the
    field strain subtracted from F as in element_force): returns fe (ne, nnode, 3) and the trial
    states (ne, ngp, 19)."""
    dNdX, vol = geom(et, Xe)
    ne = len(Xe)
    sc = np.ones(ne) if scale is None else np.asarray(scale, float)
    wv = sc[:, None] * vol
    F = defgrad(et, Xe, ue, dNdX) - field_shift(et, ne, Te, T0e, getattr(mat, "alpha", 0.0), eig)
    P = np.zeros_like(F)
    snew = np.zeros_like(np.asarray(se, float))
    for e in range(ne):
        for g in range(et.ngp):
            P[e, g], snew[e, g] = mat.pk1(F[e, g], se[e, g])
    if et.fbar:
        G = defgrad_jacobian(et, Xe, dNdX)
        return np.einsum("egij,egijak->eak", P * wv[:, :, None, None], G), snew
    return np.einsum("egij,egaj,eg->eai", P, dNdX, wv), snew


def element_stiffness_state(et, Xe, ue, se, mat, scale=None, Te=None, T0e=None, eig=None):
    """Block.stiff with a state-carrying material (gfem/model.py:83-111): the material tangent
    A = dP/dF per Gauss point at (F(ue) less the field strain, s_old), then K = sum_g wv G^T A G
    (B-bar) or the plain einsum. Returns (ne, 3nn, 3nn)."""
    dNdX, vol = geom(et, Xe)
    ne, nn = len(Xe), et.nnode
    sc = np.ones(ne) if scale is None else np.asarray(scale, float)
    wv = sc[:, None] * vol
    F = defgrad(et, Xe, ue, dNdX) - field_shift(et, ne, Te, T0e, getattr(mat, "alpha", 0.0), eig)
    A = np.zeros((ne, et.ngp, 3, 3, 3, 3))
    for e in range(ne):
        for g in range(et.ngp):
            A[e, g] = mat.tangent(F[e, g], se[e, g])
    if et.fbar:
        G = defgrad_jacobian(et, Xe, dNdX)
        K = np.einsum("egijak,egijlm,eglmbn,eg->eakbn", G, A, G, wv, optimize=True)
    else:
        K = np.einsum("egaj,egijkl,egbl,eg->eaibk", dNdX, A, dNdX, wv, optimize=True)
    return K.reshape(ne, 3 * nn, 3 * nn)


# ----------------------------------------------------------------------------- hyperelastic
# Strain-energy materials (gfem/materials.py:124-227) and StVenantKirchhoff (:71-80), finite
# strain. Every routine accepts complex arrays (no conjugation, det by cofactors), so the
# complex-step derivative Im f(x + i h e) / h checks it to rounding.
HYPER_KINDS = ("neohookean", "mooney_rivlin", "yeoh", "arruda_boyce", "gent", "svk")
_AB_C = (0.5, 1 / 20, 11 / 1050, 19 / 7000, 519 / 673750)         # gfem/materials.py:196


def _det3(F):
    return (F[..., 0, 0] * (F[..., 1, 1] * F[..., 2, 2] - F[..., 1, 2] * F[..., 2, 1])
            - F[..., 0, 1] * (F[..., 1, 0] * F[..., 2, 2] - F[..., 1, 2] * F[..., 2, 0])
            + F[..., 0, 2] * (F[..., 1, 0] * F[..., 2, 1] - F[..., 1, 1] * F[..., 2, 0]))


def hyper_invariants(F):
    """gfem/materials.py:134-141: J, I1bar = J^-2/3 tr C, I2bar = J^-4/3 (tr C^2 - C:C)/2."""
    J = _det3(F)
    C = np.swapaxes(F, -1, -2) @ F
    trC = np.trace(C, axis1=-2, axis2=-1)
    I1b = J ** (-2 / 3) * trC
    I2b = J ** (-4 / 3) * 0.5 * (trC ** 2 - np.sum(C * C, axis=(-2, -1)))
    return J, I1b, I2b


def hyper_energy(kind, p, F):
    """Strain energy W(F) (gfem/materials.py:158-226; svk: lam/2 tr(E)^2 + mu E:E with the
    Green-Lagrange E, whose gradient is the P = F S of :75-79). p: the material's constants -
    neohookean (C10, D1), mooney_rivlin (C10, C01, D1), yeoh (C10, C20, C30, D1, D2, D3),
    arruda_boyce (mu, lam_m, D), gent (mu, Jm, D), svk (mu, lam)."""
    if kind == "svk":
        mu, lam = p[0], p[1]
        E = 0.5 * (np.swapaxes(F, -1, -2) @ F - I3)
        trE = np.trace(E, axis1=-2, axis2=-1)
        return 0.5 * lam * trE ** 2 + mu * np.sum(E * E, axis=(-2, -1))
    J, I1b, I2b = hyper_invariants(F)
    if kind == "neohookean":
        return p[0] * (I1b - 3) + (J - 1) ** 2 / p[1]
    if kind == "mooney_rivlin":
        return p[0] * (I1b - 3) + p[1] * (I2b - 3) + (J - 1) ** 2 / p[2]
    if kind == "yeoh":
        W = sum(c * (I1b - 3) ** (i + 1) for i, c in enumerate(p[:3]))
        return W + sum((J - 1) ** (2 * (i + 1)) / d for i, d in enumerate(p[3:6]) if d > 0)
    if kind == "arruda_boyce":
        mu, lm, D = p[:3]
        W = sum(c / lm ** (2 * i) * (I1b ** (i + 1) - 3 ** (i + 1)) for i, c in enumerate(_AB_C))
        return mu * W + ((J * J - 1) / 2 - np.log(J)) / D
    if kind == "gent":
        mu, Jm, D = p[:3]
        return -0.5 * mu * Jm * np.log(1 - (I1b - 3) / Jm) + ((J * J - 1) / 2 - np.log(J)) / D
    raise ValueError(kind)


def hyper_pk1(kind, p, F):
    """P = dW/dF (the reference's jax.grad, gfem/materials.py:130-132) in closed form over the
    invariants: P = W_J J F^-T + W_1 J^-2/3 (2F - 2/3 I1 F^-T)
    + W_2 J^-4/3 (2 (I1 F - F C) - 4/3 I2 F^-T); svk P = F (lam tr E I + 2 mu E)."""
    Ft = np.swapaxes(F, -1, -2)
    C = Ft @ F
    if kind == "svk":
        mu, lam = p[0], p[1]
        E = 0.5 * (C - I3)
        return F @ (lam * np.trace(E, axis1=-2, axis2=-1)[..., None, None] * I3 + 2 * mu * E)
    J, I1b, I2b = hyper_invariants(F)
    FiT = np.swapaxes(np.linalg.inv(F), -1, -2)
    I1 = np.trace(C, axis1=-2, axis2=-1)
    I2 = 0.5 * (I1 ** 2 - np.sum(C * C, axis=(-2, -1)))
    W2 = 0.0 * J
    if kind == "neohookean":
        W1, WJ = p[0] + 0.0 * J, 2 * (J - 1) / p[1]
    elif kind == "mooney_rivlin":
        W1, W2, WJ = p[0] + 0.0 * J, p[1] + 0.0 * J, 2 * (J - 1) / p[2]
    elif kind == "yeoh":
        W1 = sum((i + 1) * c * (I1b - 3) ** i for i, c in enumerate(p[:3]))
        WJ = sum(2 * (i + 1) * (J - 1) ** (2 * i + 1) / d for i, d in enumerate(p[3:6]) if d > 0)
    elif kind == "arruda_boyce":
        mu, lm, D = p[:3]
        W1 = mu * sum(c * (i + 1) * I1b ** i / lm ** (2 * i) for i, c in enumerate(_AB_C))
        WJ = (J - 1 / J) / D
    elif kind == "gent":
        mu, Jm, D = p[:3]
        W1 = 0.5 * mu / (1 - (I1b - 3) / Jm)
        WJ = (J - 1 / J) / D
    else:
        raise ValueError(kind)
    e = lambda a: a[..., None, None]
    dI1b = e(J ** (-2 / 3)) * (2 * F - (2 / 3) * e(I1) * FiT)
    dI2b = e(J ** (-4 / 3)) * (2 * (e(I1) * F - F @ C) - (4 / 3) * e(I2) * FiT)
    return e(WJ * J) * FiT + e(W1) * dI1b + e(W2) * dI2b


def hyper_defgrad(et, Xe, ue, dNdX):
    """gfem/model.py:40-47, finite strain: F = I + u (x) dNdX, and for an F-bar element
    F (det F0 / det F)^(1/3) with F0 at the element centre."""
    F = I3 + np.einsum("eai,egaj->egij", ue, dNdX)
    if et.fbar:
        J0 = np.einsum("ean,ai->eni", Xe, et.dN0)
        b0 = np.einsum("ai,eij->eaj", et.dN0, np.linalg.inv(J0))
        F0 = I3 + np.einsum("eai,eaj->eij", ue, b0)
        F = F * ((_det3(F0)[:, None] / _det3(F)) ** (1 / 3))[..., None, None]
    return F


def hyper_defgrad_jacobian(et, Xe, ue, dNdX):
    """G[e,g,i,j,a,k] = dFbar_ij/du_ak (the jax.vjp pullback of gfem/model.py:77): with
    s = (det F0 / det F)^(1/3), h_ak = (F^-T dNdX_a)_k and h0 its centre value,
    G = s d_ik dNdX_aj + F_ij (s/3)(h0_ak - h_ak); G = d_ik dNdX_aj without F-bar."""
    ne, nn = Xe.shape[0], et.nnode
    eye = np.eye(3)
    base = np.einsum("ik,egaj->egijak", eye, dNdX)
    if not et.fbar:
        return base
    F = I3 + np.einsum("eai,egaj->egij", ue, dNdX)
    J0 = np.einsum("ean,ai->eni", Xe, et.dN0)
    b0 = np.einsum("ai,eij->eaj", et.dN0, np.linalg.inv(J0))
    F0 = I3 + np.einsum("eai,eaj->eij", ue, b0)
    s = (_det3(F0)[:, None] / _det3(F)) ** (1 / 3)
    h = np.einsum("egjk,egaj->egak", np.linalg.inv(F), dNdX)
    h0 = np.einsum("ejk,eaj->eak", np.linalg.inv(F0), b0)
    d = (h0[:, None] - h) * (s / 3)[..., None, None]
    return s[..., None, None, None, None] * base + np.einsum("egij,egak->egijak", F, d)


def hyper_element_force(et, Xe, ue, kind, p, scale=None):
    """Block.force for a hyperelastic material (gfem/model.py:70-81, no thermal / eigenstrain):
    fe = G^T (P(Fbar) sc vol) (ne, nnode, 3); complex-safe."""
    dNdX, vol = geom(et, np.asarray(Xe, float))
    sc = np.ones(len(Xe)) if scale is None else np.asarray(scale, float)
    wv = sc[:, None] * vol
    F = hyper_defgrad(et, Xe, ue, dNdX)
    G = hyper_defgrad_jacobian(et, Xe, ue, dNdX)
    return np.einsum("egij,egijak->eak", hyper_pk1(kind, p, F) * wv[:, :, None, None], G)


def hyper_element_energy(et, Xe, ue, kind, p, scale=None):
    """Element strain energy sum_g sc vol_g W(Fbar_g) (ne,), whose gradient is the force."""
    dNdX, vol = geom(et, np.asarray(Xe, float))
    sc = np.ones(len(Xe)) if scale is None else np.asarray(scale, float)
    return np.einsum("eg,eg->e", sc[:, None] * vol, hyper_energy(kind, p, hyper_defgrad(et, Xe, ue, dNdX)))


def hyper_element_stiffness(et, Xe, ue, kind, p, scale=None, h=1e-30):
    """Block.stiff (gfem/model.py:83-106: G^T A G plus the F-bar geometric part) as the
    complex-step Jacobian d fe / du of hyper_element_force, exact to rounding. (ne, 3nn, 3nn),
    row 3a+i, column 3b+k."""
    ue = np.asarray(ue, float)
    ne, nn = ue.shape[:2]
    K = np.zeros((ne, nn, 3, nn, 3))
    for b in range(nn):
        for k in range(3):
            up = ue.astype(complex)
            up[:, b, k] += 1j * h
            K[:, :, :, b, k] = hyper_element_force(et, Xe, up, kind, p, scale).imag / h
    return K.reshape(ne, 3 * nn, 3 * nn)


# ----------------------------------------------------------------------------- buckling
def element_cauchy(et, Xe, ue, material):
    """Model.stresses S (gfem/model.py:129-136,529-537): the Gauss-point Cauchy stresses
    averaged with weights vol_g / sum vol - P itself for a small-strain material, P F^T / det F
    (F the element's deformation gradient, F-bar included) otherwise. material: ("linear",
    (E, nu)) or (hyperelastic kind, constants). (ne, 3, 3)."""
    Xe = np.asarray(Xe, float)
    dNdX, vol = geom(et, Xe)
    kind, p = material
    if kind == "linear":
        sig = pk1_linear(defgrad(et, Xe, ue, dNdX), *p)
    else:
        F = hyper_defgrad(et, Xe, ue, dNdX)
        sig = hyper_pk1(kind, p, F) @ np.swapaxes(F, -1, -2) / _det3(F)[..., None, None]
    return np.einsum("eg,egij->eij", vol / vol.sum(1, keepdims=True), sig)


def geo_stiff(et, Xe, sig):
    """Block.geo_stiff (gfem/model.py:113-118): initial-stress stiffness of a constant Cauchy
    stress per element, K[a i, b j] = d_ij sum_g vol_g (dNdX_a . sig . dNdX_b). (ne, 3nn, 3nn)."""
    Xe = np.asarray(Xe, float)
    dNdX, vol = geom(et, Xe)
    k = np.einsum("egai,eij,egbj,eg->eab", dNdX, sig, dNdX, vol)
    ne, nn = k.shape[:2]
    return np.einsum("eab,ij->eaibj", k, I3).reshape(ne, 3 * nn, 3 * nn)


def buckling_factors(K0, KG, free, nmodes):
    """The eigenproblem of buckling (gfem/solvers.py:1095-1108): -KG phi = mu K0 phi on the free
    dofs, largest mu first, load factor 1/mu (inf for mu <= 0); dense, for small systems.
    Returns (load factors (nmodes,), modes (nmodes, ndof))."""
    import scipy.linalg as sla
    keep = np.nonzero(np.asarray(free) > 0)[0]
    A = sp.csr_matrix(K0)[keep][:, keep].toarray()
    B = sp.csr_matrix(KG)[keep][:, keep].toarray()
    mu, V = sla.eigh(-B, A)
    order = np.argsort(-mu)[:nmodes]
    mu, V = mu[order], V[:, order]
    modes = np.zeros((len(mu), K0.shape[0]))
    modes[:, keep] = V.T
    return np.where(mu > 0, 1.0 / np.where(mu > 0, mu, 1.0), np.inf), modes


def element_mass(et, Xe, density):
    """HRZ diagonal lumping (gfem/model.py:120-124): d_a = rho sum_g vol_g N_ga^2, scaled so that
    the element's nodal masses add up to rho V_e. Xe (ne,nnode,3) -> (ne,nnode)."""
    _, vol = geom(et, Xe)
    d = density * np.einsum("eg,ga->ea", vol, et.N ** 2)
    return d * (density * vol.sum(1))[:, None] / d.sum(1)[:, None]


def lumped_mass(X, blocks):
    """Model.lumped_mass (gfem/model.py:501-510, no point masses): nodal masses summed over the
    blocks [(etype, conn, density)], repeated for the 3 dofs of each node."""
    m = np.zeros(len(X))
    for et, conn, rho in blocks:
        np.add.at(m, conn, element_mass(et, X[conn], rho))
    return np.repeat(m, 3)


def modal_frequencies(K, M, free, nmodes):
    """modal, Lanczos branch (gfem/solvers.py:1011-1046): K phi = w^2 M phi on the free dofs with
    the lumped mass M (ndof,), dense for tiny systems, shift-invert eigsh otherwise. Returns
    (freq (nmodes,), modes (nmodes, ndof))."""
    keep = np.nonzero(np.asarray(free) > 0)[0]
    K = sp.csr_matrix(K)[keep][:, keep].tocsc()
    Mk = sp.diags(np.asarray(M)[keep]).tocsc()
    if K.shape[0] <= max(nmodes + 2, 60):
        import scipy.linalg as sla
        nm = min(nmodes, K.shape[0])
        w2, V = sla.eigh(K.toarray(), Mk.toarray())
        w2, V = w2[:nm], V[:, :nm]
    else:
        nm = min(nmodes, K.shape[0] - 2)
        sigma = -1e-6 * abs(K.diagonal()).max() / max(np.asarray(M).max(), 1e-300)
        w2, V = spla.eigsh(K, k=nm, M=Mk, sigma=sigma, which="LM")
    order = np.argsort(w2)
    w2, V = np.maximum(w2[order], 0.0), V[:, order]
    modes = np.zeros((nm, len(free)))
    modes[:, keep] = V.T
    return np.sqrt(w2) / (2 * np.pi), modes


def dofs(conn):
    """Block.dofs, gfem/model.py:157-158 (node-interleaved 3*node + comp)."""
    conn = np.asarray(conn, np.int64)
    return (conn[:, :, None] * 3 + np.arange(3)).reshape(len(conn), -1)


# ----------------------------------------------------------------------------- global operators
def ops_matvec(ops, v, ndof):
    """gfem/model.py:162-166 (scatter-add)."""
    out = np.zeros(ndof, dtype=np.asarray(v).dtype)
    for d, K in ops:
        np.add.at(out, d, np.einsum("eij,ej->ei", K, v[d]))
    return out


def ops_diagonal(ops, ndof):
    """gfem/model.py:169-173."""
    out = np.zeros(ndof)
    for d, K in ops:
        np.add.at(out, d, np.diagonal(K, axis1=1, axis2=2))
    return out


def ops_assemble(ops, ndof):
    """gfem/model.py:176-192: host COO -> CSR with summed duplicates."""
    rows, cols, vals = [], [], []
    for d, K in ops:
        d = np.asarray(d)
        nd = d.shape[1]
        rows.append(np.repeat(d, nd, axis=1).ravel())
        cols.append(np.tile(d, (1, nd)).ravel())
        vals.append(np.asarray(K).reshape(len(d), -1).ravel())
    if not rows:
        return sp.csr_matrix((ndof, ndof))
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ndof, ndof)).tocsr()
    A.sum_duplicates()
    return A


def gather_plan(conns, nn):
    """gfem/model.py:241-261: per node the stable-sorted (element, local node) slots, padded
    with a sentinel that points at a zero row."""
    slots, nodes, off = [], [], 0
    for c in conns:
        c = np.asarray(c)
        nodes.append(c.ravel())
        slots.append(off + np.arange(c.size))
        off += c.size
    nodes, slots = np.concatenate(nodes), np.concatenate(slots)
    order = np.argsort(nodes, kind="stable")
    nodes, slots = nodes[order], slots[order]
    cnt = np.bincount(nodes, minlength=nn)
    vmax = max(int(cnt.max()), 1)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    pos = np.arange(len(nodes)) - start[nodes]
    g = np.full((nn, vmax), off, np.int64)
    g[nodes, pos] = slots
    return g


def gather_matvec(plan, ops, v, ndof):
    """gfem/model.py:264-267: element products, then the atomic-free gather-reduce."""
    parts = [np.einsum("eij,ej->ei", K, v[d]).reshape(-1, 3) for d, K in ops]
    ys = np.concatenate(parts + [np.zeros((1, 3), v.dtype)])
    return np.sum(ys[plan], axis=1).reshape(-1)


def internal_force(X, blocks, u, nn):
    """Model.internal_force, gfem/model.py:466-477. blocks: [(etype, conn, E, nu, scale)]."""
    U = np.asarray(u).reshape(-1, 3)
    parts = []
    for et, conn, E, nu, sc in blocks:
        parts.append(element_force(et, X[conn], U[conn], E, nu, sc).reshape(-1, 3))
    ys = np.concatenate(parts + [np.zeros((1, 3))])
    plan = gather_plan([c for _, c, _, _, _ in blocks], nn)
    return np.sum(ys[plan], axis=1).reshape(-1)


# ----------------------------------------------------------------------------- discrete / contact
def spring_energy(kind, Xe, ue, k):
    """gfem/discrete.py:60-66 (axial: k/2 (L - L0)^2; bushing: sum k_i d_i^2 / 2, d = u1 - u0)
    and :105-112 (ground: sum k_i u_i^2 / 2)."""
    Xe, ue, k = np.asarray(Xe, float), np.asarray(ue, float), np.broadcast_to(np.asarray(k, float), (3,))
    if kind == "ground":
        return 0.5 * float(np.sum(k * ue[0] ** 2))
    if kind == "axial":
        L0 = np.linalg.norm(Xe[1] - Xe[0])
        L = np.linalg.norm(Xe[1] + ue[1] - Xe[0] - ue[0])
        return 0.5 * k[0] * (L - L0) ** 2
    d = ue[1] - ue[0]
    return 0.5 * float(np.sum(k * d * d))


def spring_element(kind, Xe, ue, k, sc=1.0):
    """Block.force / Block.stiff of SpringBlock / GroundSpring (gfem/discrete.py:68-72,
    :108-112): force = sc x grad of the energy, stiffness = its Jacobian, in closed form.
    Returns fe (nnode*3,), Ke (nd, nd)."""
    Xe, ue = np.asarray(Xe, float), np.asarray(ue, float)
    k = np.broadcast_to(np.asarray(k, float), (3,))
    if kind == "ground":
        return sc * k * ue[0], sc * np.diag(k)
    if kind == "axial":
        e0 = Xe[1] - Xe[0]
        e = e0 + ue[1] - ue[0]
        L0, L = np.linalg.norm(e0), np.linalg.norm(e)
        t = e / L
        f1 = sc * k[0] * (L - L0) * t
        B = sc * k[0] * (np.outer(t, t) + (1 - L0 / L) * (I3 - np.outer(t, t)))
    else:
        f1 = sc * k * (ue[1] - ue[0])
        B = sc * np.diag(k)
    return np.concatenate([-f1, f1]), np.block([[B, -B], [-B, B]])


def contact_gap(x):
    """_gap (gfem/contact.py:19-24) of x (4, 3) = slave, facet nodes."""
    n = np.cross(x[2] - x[1], x[3] - x[1])
    return float((x[0] - x[1]) @ (n / np.sqrt(n @ n + 1e-300)))


def contact_pair(x, k, lam, active):
    """_pair (gfem/contact.py:27-36): force (4, 3) and Gauss-Newton tangent (12, 12); dg in
    closed form (n for the slave, (x3-x1) x q and q x (x2-x1) for facet nodes 2, 3 with
    q = (d - g n)/|m|, the rest on node 1)."""
    x = np.asarray(x, float)
    e1, e2, d = x[2] - x[1], x[3] - x[1], x[0] - x[1]
    m = np.cross(e1, e2)
    mm = np.sqrt(m @ m + 1e-300)
    nh = m / mm
    g = d @ nh
    q = (d - g * nh) / mm
    d2, d3 = np.cross(e2, q), np.cross(q, e1)
    dg = np.concatenate([nh, -(nh + d2 + d3), d2, d3])
    pn = max(lam - k * g, 0.0) * active
    on = 1.0 if pn > 0 else 0.0
    return (-pn * dg).reshape(4, 3), k * on * np.outer(dg, dg)


def contact_select(X, u, slave, cand, h, depth):
    """PenaltyContact.select (gfem/contact.py:152-182): pair nodes (ns, 4), active (ns,), gap."""
    x = np.asarray(X, float) + np.asarray(u, float).reshape(-1, 3)
    xs = x[slave]
    T = x[cand]
    A, B, C = T[:, :, 0], T[:, :, 1], T[:, :, 2]
    ab, ac, ap = B - A, C - A, xs[:, None] - A
    d00, d01, d11 = (ab * ab).sum(-1), (ab * ac).sum(-1), (ac * ac).sum(-1)
    d20, d21 = (ap * ab).sum(-1), (ap * ac).sum(-1)
    den = np.maximum(d00 * d11 - d01 * d01, 1e-300)
    v = np.clip((d11 * d20 - d01 * d21) / den, 0, 1)
    w = np.clip((d00 * d21 - d01 * d20) / den, 0, 1)
    sw = np.maximum(v + w, 1.0)                 # v, w >= 0: rescaled only when v + w > 1
    v, w = v / sw, w / sw
    Q = A + v[..., None] * ab + w[..., None] * ac
    j = np.argmin(np.linalg.norm(xs[:, None] - Q, axis=-1), axis=1)
    ii = np.arange(len(slave))
    tri = cand[ii, j]
    Ab, Bb, Cb, Qb = A[ii, j], B[ii, j], C[ii, j], Q[ii, j]
    n = np.cross(Bb - Ab, Cb - Ab)
    n = n / np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-300)
    g = ((xs - Ab) * n).sum(-1)
    lateral = np.linalg.norm((xs - Qb) - g[:, None] * n, axis=1)
    active = ((lateral < 0.05 * h) & (g > -depth)).astype(float)
    return np.concatenate([np.asarray(slave)[:, None], tri], 1), active, g


def wall_evaluate(X, u, nodes, point, normal, k):
    """RigidWall.evaluate (gfem/contact.py:254-265): fe (ns, 3), K (ns, 3, 3)."""
    n = np.asarray(normal, float) / np.linalg.norm(normal)
    x = (np.asarray(X, float) + np.asarray(u, float).reshape(-1, 3))[nodes]
    g = (x - point) @ n
    pn = np.maximum(-k * g, 0.0)
    on = (pn > 0).astype(float)
    return -pn[:, None] * n[None], k * on[:, None, None] * np.outer(n, n)[None]


def op_bc(K_apply, free):
    """_LinearSystem.opA, gfem/solvers.py:257-267 without MPC (t_apply is the identity,
    gfem/constraints.py:291-302): A v = free*K(free*v) + (1-free)*v."""
    free = np.asarray(free, float)
    return lambda v: free * K_apply(free * v) + (1 - free) * v


def mpc_rows(slave, masters, weights):
    """MPC, gfem/constraints.py:14-25: u[slave] = sum_k w_k u[master_k]."""
    sl = np.asarray(slave, dtype=np.int64).ravel()
    k = np.asarray(masters).size // max(len(sl), 1)
    return (sl, np.asarray(masters, dtype=np.int64).reshape(len(sl), k),
            np.asarray(weights, dtype=np.float64).reshape(len(sl), k))


def normal_constraint(nodes, normal):
    """gfem/constraints.py:198-209: n . u = 0, slave = the dominant normal component."""
    n = np.asarray(normal, float)
    n = n / np.linalg.norm(n)
    k = int(np.argmax(np.abs(n)))
    others = [j for j in range(3) if j != k]
    nodes = np.asarray(nodes, dtype=np.int64)
    return mpc_rows(nodes * 3 + k, np.stack([nodes * 3 + j for j in others], 1),
                    np.tile([-n[j] / n[k] for j in others], (len(nodes), 1)))


def coupling(ref, nodes):
    """gfem/constraints.py:212-218: every node's translations follow node `ref`."""
    nodes = np.asarray([n for n in nodes if n != ref], dtype=np.int64)
    sd = (nodes[:, None] * 3 + np.arange(3)).ravel()
    return mpc_rows(sd, np.tile(ref * 3 + np.arange(3), len(nodes))[:, None], np.ones((len(sd), 1)))


def transformation(ndof, mpcs, fixed=None):
    """Transformation.__init__, gfem/constraints.py:224-276: T (scipy CSR, ndof x ndof) with
    u = T u_bar, and the slave mask. One constraint per slave dof (the first wins), prescribed
    slaves dropped, chains resolved by up to 8 squarings, residual cycles broken by zeroing the
    slave columns. None if there are no rows."""
    rows = [m for m in mpcs if len(m[0])]
    if not rows:
        return None, np.zeros(ndof, bool)
    sl = np.concatenate([m[0] for m in rows])
    kmax = max(m[1].shape[1] for m in rows)
    md = np.concatenate([np.pad(m[1], ((0, 0), (0, kmax - m[1].shape[1])), constant_values=-1) for m in rows])
    wt = np.concatenate([np.pad(m[2], ((0, 0), (0, kmax - m[2].shape[1]))) for m in rows])
    _, first = np.unique(sl, return_index=True)
    sl, md, wt = sl[first], md[first], wt[first]
    if fixed is not None:
        ok = ~np.asarray(fixed)[sl]
        sl, md, wt = sl[ok], md[ok], wt[ok]
    smask = np.zeros(ndof, bool)
    smask[sl] = True
    valid = md >= 0
    r = np.repeat(sl, valid.sum(1))
    keep = np.nonzero(~smask)[0]
    T = sp.csr_matrix((np.concatenate([wt[valid], np.ones(len(keep))]),
                       (np.concatenate([r, keep]), np.concatenate([md[valid], keep]))), shape=(ndof, ndof))
    for _ in range(8):
        if not T[:, smask].nnz:
            break
        T = T @ T
    T = (T @ sp.diags((~smask).astype(float))).tocsr()
    T.eliminate_zeros()
    return T, smask


def op_bc_mpc(K_apply, free, T):
    """opA with an MPC transformation (gfem/solvers.py:257-267): free*T^T K T(free*v) + (1-free)*v."""
    def A(v):
        w = T @ (free * v)
        return free * (T.T @ K_apply(w)) + (1 - free) * v
    return A


def bc_matrix(A, free):
    """gfem/linalg.py:213-217."""
    D = sp.diags(np.asarray(free, float))
    return (D @ A @ D + sp.diags(1.0 - np.asarray(free, float))).tocsr()


def jacobi_inv(diag, free):
    """gfem/solvers.py:432-434."""
    d = np.asarray(diag)
    return np.where(np.asarray(free) > 0, 1.0 / np.where(np.abs(d) > 0, d, 1.0), 1.0)


# ----------------------------------------------------------------------------- Krylov
def pcg(A, b, M=None, x0=None, tol=1e-8, maxiter=2000):
    """gfem/linalg.py:21-47, same recurrence and stopping rule (recursive residual,
    rr <= tol^2 max(b.b, 1e-300)). Returns (x, k, relres)."""
    M = M or (lambda r: r)
    x = np.zeros_like(b) if x0 is None else np.array(x0, copy=True)
    r = b - A(x)
    z = M(r)
    p = z
    rz = np.vdot(r, z)
    bn2 = max(np.vdot(b, b), 1e-300)
    tol2 = tol * tol * bn2
    k = 0
    rr = np.vdot(r, r)
    while rr > tol2 and k < maxiter:
        Ap = A(p)
        alpha = rz / np.vdot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        rr = np.vdot(r, r)
        z = M(r)
        rz_new = np.vdot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
        k += 1
    return x, k, np.sqrt(rr / bn2)


def lmax_power(A, dinv, free, its=25):
    """_LinearSystem._lmax, gfem/solvers.py:328-342."""
    n = len(free)
    v = np.sin(np.arange(n) * 0.37 + 0.1) * free + 1e-3
    lam = 1.0
    for _ in range(its):
        w = dinv * A(v)
        lam = np.linalg.norm(w) / np.linalg.norm(v)
        v = w / np.linalg.norm(w)
    return 1.1 * lam


def cheby_precond(A, dinv, lmax, steps=6):
    """Chebyshev-Jacobi preconditioner, gfem/solvers.py:289-307."""
    def M(r):
        lo, hi = lmax / 30.0, lmax
        th, de = 0.5 * (hi + lo), 0.5 * (hi - lo)
        sg = th / de
        rh = 1.0 / sg
        res = dinv * r
        d = res / th
        x = np.zeros_like(r)
        for _ in range(steps):
            x = x + d
            res = res - dinv * A(d)
            rn = 1.0 / (2 * sg - rh)
            d = rn * rh * d + 2 * rn / de * res
            rh = rn
        return x + d
    return M


# ----------------------------------------------------------------------------- AMG cycle
def cheb(Amv, dinv, rho, x, b, deg=2):
    """Chebyshev smoother on D^-1 A over [rho/30, rho], gfem/linalg.py:161-175."""
    lmax, lmin = rho, rho / 30.0
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sigma = theta / delta
    rh = 1.0 / sigma
    r = dinv * (b - Amv(x))
    d = r / theta
    for _ in range(deg):
        x = x + d
        r = r - dinv * Amv(d)
        rh_new = 1.0 / (2 * sigma - rh)
        d = rh_new * rh * d + 2 * rh_new / delta * r
        rh = rh_new
    return x + d


def vcycle(levels, Lc, b, lvl=0, A0=None, dinv0=None, deg=2):
    """gfem/linalg.py:177-192. levels: [dict(A=csr, dinv, rho, P=csr, R=csr)], Lc lower
    Cholesky factor of the coarsest matrix (two triangular solves, :178-181)."""
    import scipy.linalg as sla
    if lvl == len(levels):
        y = sla.solve_triangular(Lc, b, lower=True)
        return sla.solve_triangular(Lc.T, y, lower=False)
    L = levels[lvl]
    if lvl == 0 and A0 is not None:
        Amv, dinv = A0, dinv0
    else:
        Amv, dinv = (lambda v, A=L["A"]: A @ v), L["dinv"]
    x = cheb(Amv, dinv, L["rho"], np.zeros_like(b), b, deg)
    r = b - Amv(x)
    xc = vcycle(levels, Lc, L["R"] @ r, lvl + 1, deg=deg)
    x = x + L["P"] @ xc
    return cheb(Amv, dinv, L["rho"], x, b, deg)


def coarse_factor(Ac):
    """gfem/linalg.py:126-133."""
    Ac = np.asarray(Ac, float)
    Ac = 0.5 * (Ac + Ac.T)
    try:
        return np.linalg.cholesky(Ac + 1e-13 * np.abs(Ac).max() * np.eye(len(Ac)))
    except np.linalg.LinAlgError:
        w, V = np.linalg.eigh(Ac)
        return np.linalg.cholesky(V @ np.diag(np.maximum(w, 1e-12 * w.max())) @ V.T)


def rho_estimate(A, dinv, it=15):
    """AMG._rho, gfem/linalg.py:138-146."""
    x = np.random.default_rng(0).standard_normal(A.shape[0])
    lam = 1.0
    for _ in range(it):
        y = dinv * (A @ x)
        lam = np.linalg.norm(y) / max(np.linalg.norm(x), 1e-300)
        x = y / max(np.linalg.norm(y), 1e-300)
    return 1.1 * lam


# ----------------------------------------------------------------------------- SA set-up
# gfem/linalg.py:103-136 hands the set-up to PyAMG (an optional dependency, unpinned in the
# reference's pyproject.toml:16, not installed here): smoothed_aggregation_solver(Ab, B,
# strength=("symmetric", theta 0), smooth=("jacobi", omega 4/3, weighting "local"),
# max_coarse, keep=False, improve_candidates=None). Restated below from PyAMG's published
# algorithm: symmetric_strength_of_connection (pyamg/strength.py), standard_aggregation and
# fit_candidates (pyamg/amg_core), jacobi_prolongation_smoother (pyamg/aggregation/smooth.py)
# and the level loop of smoothed_aggregation_solver (pyamg/aggregation/aggregation.py).

def sa_strength(A, bs):
    """Node graph of the symmetric strength of connection with theta = 0 on bs x bs blocks:
    with theta = 0 every stored block is strong, so the graph is the block pattern of the BSR
    matrix (a block is stored when it holds a nonzero entry); the diagonal is dropped, it plays
    no part in the aggregation. Returns neighbour lists (ascending)."""
    Ab = sp.bsr_matrix(sp.csr_matrix(A), blocksize=(bs, bs))
    Ab.sort_indices()
    nb = Ab.shape[0] // bs
    nbr = []
    for i in range(nb):
        cols = Ab.indices[Ab.indptr[i]:Ab.indptr[i + 1]]
        blocks = Ab.data[Ab.indptr[i]:Ab.indptr[i + 1]]
        nbr.append([int(j) for j, blk in zip(cols, blocks) if j != i and np.any(blk != 0)])
    return nbr


def sa_standard_aggregation(nbr):
    """PyAMG's standard (greedy, three-pass) aggregation on neighbour lists, in the published
    encoding of its work array x (0 unmarked, k > 0 seeded in pass 1 as aggregate k, -k joined
    aggregate k in pass 2, -n isolated). Pass 1: in node order, a node whose neighbours are all
    unmarked seeds an aggregate with them; a node without neighbours is isolated. Pass 2: an
    unmarked node joins the pass-1 aggregate of its first neighbour seeded there. Pass 3: in
    node order, marked nodes are renumbered from 0 (isolated: -1) and an unmarked node seeds a
    new aggregate with its unmarked neighbours (which pass 3 then renumbers again when it
    reaches them, as the published loop does). Returns (agg, number of aggregates)."""
    n = len(nbr)
    x = [0] * n
    k = 0
    for i in range(n):
        if x[i]:
            continue
        if not nbr[i]:
            x[i] = -n
        elif not any(x[j] for j in nbr[i]):
            k += 1
            x[i] = k
            for j in nbr[i]:
                x[j] = k
    for i in range(n):
        if x[i] == 0:
            for j in nbr[i]:
                if x[j] > 0:
                    x[i] = -x[j]
                    break
    nxt = k
    for i in range(n):
        if x[i] > 0:
            x[i] -= 1
        elif x[i] == -n:
            x[i] = -1
        elif x[i] < 0:
            x[i] = -x[i] - 1
        else:
            x[i] = nxt
            for j in nbr[i]:
                if x[j] == 0:
                    x[j] = nxt
            nxt += 1
    return np.array(x, dtype=np.int64), nxt


def sa_fit_candidates(agg, nagg, B, K1, tol=1e-10):
    """Tentative prolongator: for each aggregate, the candidate rows of its member nodes
    (ascending node order, K1 rows per node) are orthonormalised column by column (modified
    Gram-Schmidt); a column whose norm after orthogonalisation is not above tol times its norm
    before is set to zero. Returns T (n K1, nagg K2) and the coarse candidates R (nagg K2, K2),
    B restricted to an aggregate == its T block times its R block."""
    B = np.asarray(B, float)
    n = len(agg)
    K2 = B.shape[1]
    T = sp.lil_matrix((n * K1, nagg * K2))
    Rc = np.zeros((nagg * K2, K2))
    for a in range(nagg):
        nodes = np.flatnonzero(agg == a)
        rows = (nodes[:, None] * K1 + np.arange(K1)).ravel()
        Q = B[rows].copy()
        Ra = np.zeros((K2, K2))
        for j in range(K2):
            thresh = tol * np.sqrt(np.sum(Q[:, j] ** 2))
            for i in range(j):
                d = np.dot(Q[:, j], Q[:, i])
                Q[:, j] -= d * Q[:, i]
                Ra[i, j] = d
            nj = np.sqrt(np.sum(Q[:, j] ** 2))
            if nj > thresh:
                Q[:, j] /= nj
                Ra[j, j] = nj
            else:
                Q[:, j] = 0.0
        for k, r in enumerate(rows):
            T[r, a * K2:(a + 1) * K2] = Q[k]
        Rc[a * K2:(a + 1) * K2] = Ra
    return T.tocsr(), Rc


def sa_smooth_prolongator(A, T, omega=4.0 / 3.0):
    """One Jacobi step on the tentative prolongator with the local (Gershgorin) weighting:
    P = T - omega D^-1 A T, D_rr = sum_c |A_rc| (rows with D = 0 are left as T)."""
    A = sp.csr_matrix(A)
    D = np.asarray(abs(A).sum(axis=1)).ravel()
    w = np.divide(omega, D, out=np.zeros_like(D), where=D != 0)
    return sp.csr_matrix(T - sp.diags(w) @ (A @ T))


def sa_hierarchy(A, B, bs, max_coarse=300, max_levels=10, omega=4.0 / 3.0):
    """The level loop: while the level has more than max_coarse block rows (and fewer than
    max_levels levels), aggregate, fit, smooth and form A_c = P^T A P; the coarse block size is
    the number of candidates. Returns ([dict(A, agg, T, P, Bc)], A_coarsest)."""
    A = sp.csr_matrix(A)
    levels = []
    while len(levels) + 1 < max_levels and A.shape[0] // bs > max_coarse:
        agg, nagg = sa_standard_aggregation(sa_strength(A, bs))
        if nagg == 0 or nagg >= A.shape[0] // bs:
            break
        T, Bc = sa_fit_candidates(agg, nagg, B, bs)
        P = sa_smooth_prolongator(A, T, omega)
        Ac = sp.csr_matrix(P.T @ (A @ P))
        levels.append(dict(A=A, agg=agg, T=T, P=P, Bc=Bc))
        A, B, bs = Ac, Bc, B.shape[1]
    return levels, A


def rigid_body_modes(X):
    """gfem/linalg.py:80-98 (no rotation rows)."""
    X = np.asarray(X)
    n = len(X)
    c = X - X.mean(0)
    B = np.zeros((3 * n, 6))
    for i in range(3):
        B[i::3, i] = 1.0
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    B[0::3, 3], B[1::3, 3] = -y, x
    B[1::3, 4], B[2::3, 4] = -z, y
    B[0::3, 5], B[2::3, 5] = z, -x
    return B


def mixed_solve(A64, inner, b, cg_tol, max_ref=12):
    """_LinearSystem._mixed, gfem/solvers.py:371-384: x=0, r=b; repeat <=12: sc=|r| or 1,
    d = inner(r/sc) (FP32), x += sc d, r = b - A64 x, stop on |r| <= cg_tol |b|.
    Returns (x, total inner iterations, refinement steps)."""
    x = np.zeros_like(b)
    r = b.copy()
    bn = float(np.linalg.norm(b)) or 1.0
    its = 0
    for step in range(max_ref):
        sc = float(np.linalg.norm(r)) or 1.0
        d, k = inner((r / sc).astype(np.float32))
        its += int(k)
        x = x + d.astype(np.float64) * sc
        r = b - A64(x)
        if float(np.linalg.norm(r)) <= cg_tol * bn:
            return x, its, step + 1
    return x, its, max_ref


def multi_load_cases(ops, ndof, free, F, cg_tol=1e-8, maxiter=20000):
    """gfem/optimize.py:116-143: every case is pcg(A, free*f, Jacobi, tol) with the scatter
    ops_matvec operator; the vmapped loop is restated as a loop over cases."""
    free = np.asarray(free, float)
    A = op_bc(lambda w: ops_matvec(ops, w, ndof), free)
    inv = jacobi_inv(ops_diagonal(ops, ndof), free)
    out = []
    for f in np.atleast_2d(F):
        x, _, _ = pcg(A, free * f, lambda r: inv * r, tol=cg_tol, maxiter=maxiter)
        out.append(x)
    return np.array(out)


# ----------------------------------------------------------------------------- meshes / loads
def box_hex(L, H, W, nx, ny, nz, origin=(0, 0, 0)):
    """gfem/mesh.py:5-16."""
    xs = np.linspace(0, L, nx + 1) + origin[0]
    ys = np.linspace(0, H, ny + 1) + origin[1]
    zs = np.linspace(0, W, nz + 1) + origin[2]
    X = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"), -1).reshape(-1, 3)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()

    def nid(a, b, c):
        return (a * (ny + 1) + b) * (nz + 1) + c
    conn = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                     nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                     nid(i, j + 1, k + 1)], -1)
    return X, conn


_KUHN = np.array([[0, 1, 2, 6], [0, 2, 3, 6], [0, 3, 7, 6], [0, 7, 4, 6], [0, 4, 5, 6], [0, 5, 1, 6]])


def hex_to_tet4(X, conn):
    """gfem/mesh.py:23-29 (Kuhn split, positive orientation)."""
    tets = conn[:, _KUHN].reshape(-1, 4)
    v = np.einsum("ij,ij->i", np.cross(X[tets[:, 1]] - X[tets[:, 0]], X[tets[:, 2]] - X[tets[:, 0]]),
                  X[tets[:, 3]] - X[tets[:, 0]])
    neg = v < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    return X, tets


def tet4_to_tet10(X, conn):
    """gfem/mesh.py:32-38."""
    e = np.concatenate([np.sort(conn[:, list(p)], 1) for p in _EDGES10])
    uniq, inv = np.unique(e, axis=0, return_inverse=True)
    mids = 0.5 * (X[uniq[:, 0]] + X[uniq[:, 1]])
    newid = inv.reshape(len(_EDGES10), -1).T + len(X)
    return np.vstack([X, mids]), np.hstack([conn, newid])


def quad4_weights(X, facets):
    """Surface.integrate for QUAD4 facets (gfem/surfaces.py:103-120, 2x2 Gauss rule of
    gfem/elements.py:78-79,95-96): nodal weights w_a = int N_a dA."""
    g = 1 / np.sqrt(3)
    pts = [[-g, -g], [g, -g], [g, g], [-g, g]]
    QC = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], float)
    w = np.zeros(len(X))
    xe = X[facets]
    for s, t in pts:
        N = 0.25 * (1 + QC[:, 0] * s) * (1 + QC[:, 1] * t)
        dNs = 0.25 * QC[:, 0] * (1 + QC[:, 1] * t)
        dNt = 0.25 * QC[:, 1] * (1 + QC[:, 0] * s)
        a1 = np.einsum("fai,a->fi", xe, dNs)
        a2 = np.einsum("fai,a->fi", xe, dNt)
        da = np.linalg.norm(np.cross(a1, a2), axis=1)
        np.add.at(w, facets, da[:, None] * N[None, :])
    return w


def traction_vector(weights, t):
    """gfem/loads.py:54-57."""
    return (weights[:, None] * np.asarray(t, float)[None]).ravel()
