"""Host-side logic of the product and the C ABI surface (CPU only, no compute calls on a GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import gfem_oracle as O
from exfem_amd import _lib as L
from exfem_amd import amg_setup, mesh, workloads
from exfem_amd.model import gather_plan, ops_assemble

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_declares_exactly_the_bound_abi():
    hdr = open(os.path.join(ROOT, "include", "exfem.h")).read()
    declared = set(re.findall(r"\b(exfem_[a-z0-9_]+)\s*\(", hdr))
    bound = set(L.SIGNATURES)
    assert declared == bound, (declared ^ bound)


def test_library_loads_and_exports_every_symbol():
    lib = L.load()
    for name in L.SIGNATURES:
        assert hasattr(lib, name), name
    assert lib.exfem_abi_version() == 1
    # a host-only query: packed Ke size of the 1M-dof config (chunks of 4096 elements from 65536
    # elements on, else 64)
    assert lib.exfem_ke_bytes(L.HEX8, 320000, L.KE_FULL, L.F64) == 79 * 4096 * 576 * 8
    assert lib.exfem_ke_bytes(L.HEX8, 65535, L.KE_FULL, L.F64) == 65536 * 576 * 8
    assert lib.exfem_ke_bytes(L.HEX8, 65, L.KE_SYM, L.F32) == 128 * 300 * 4
    assert lib.exfem_ke_bytes(L.TET10, 1, L.KE_SYM, L.F64) == 64 * 465 * 8
    # bad arguments are errors, not crashes
    assert lib.exfem_ke_bytes(7, 1, 0, 0) < 0
    assert b"element type" in lib.exfem_last_error()


def test_mesh_generators_match_the_oracle():
    X1, c1 = mesh.box_hex(100, 20, 20, 20, 4, 4)
    X0, c0 = O.box_hex(100, 20, 20, 20, 4, 4)
    assert np.array_equal(X1, X0) and np.array_equal(c1, c0)
    X1, t1 = mesh.tet4_to_tet10(*mesh.hex_to_tet4(X1, c1))
    X0, t0 = O.tet4_to_tet10(*O.hex_to_tet4(X0, c0))
    assert np.array_equal(X1, X0) and np.array_equal(t1, t0)


def test_gather_plan_matches_the_oracle():
    X, c = mesh.box_hex(3, 2, 2, 3, 2, 2)
    assert np.array_equal(gather_plan([c], len(X)), O.gather_plan([c], len(X)))


def test_workload_sizes_and_loads():
    p = workloads.small_cantilever()
    assert (p.nn, p.ndof, p.ne) == (525, 1575, 320)
    fz = p.f.reshape(-1, 3)[:, 2]
    assert np.isclose(fz.sum(), -400.0)
    assert np.allclose(np.sort(np.unique(np.round(-fz[fz < 0], 9))), [6.25, 12.5, 25.0])
    assert p.free.reshape(-1, 3)[p.X[:, 0] < 1e-9].sum() == 0
    big = workloads.cantilever(200, 40, 40, jitter=0.1, seed=0)
    assert (big.nn, big.ndof, big.ne) == (337881, 1013643, 320000)
    # jitter stays within +-10% h and only moves interior nodes; faces stay flat
    X0, _ = mesh.box_hex(100, 20, 20, 200, 40, 40)
    d = np.abs(big.X - X0)
    assert d.max() <= 0.05 + 1e-12 and d.max() > 0.04
    assert np.isclose(big.f.reshape(-1, 3)[:, 2].sum(), -400.0)


def test_tet10_workload_has_no_flat_tets():
    p = workloads.tet10_box(6, 6, 24, seed=0)
    X, c = p.X, p.conn
    v = np.einsum("ij,ij->i", np.cross(X[c[:, 1]] - X[c[:, 0]], X[c[:, 2]] - X[c[:, 0]]),
                  X[c[:, 3]] - X[c[:, 0]]) / 6
    assert v.min() > 1e-6 and np.isclose(v.sum(), 4.0)
    assert p.scale.min() >= 1e3 and p.scale.max() <= 1e6
    assert np.isclose(p.f.reshape(-1, 3)[:, 0].sum(), -1.0)


def _small_system(n=(6, 2, 2)):
    X, c = mesh.box_hex(3.0, 1.0, 1.0, *n)
    ops = [(O.dofs(c), O.element_stiffness(O.HEX8, X[c], 200e3, 0.3))]
    nd = 3 * len(X)
    free = np.ones(nd)
    free[np.nonzero(X[:, 0] < 1e-9)[0][:, None] * 3 + np.arange(3)] = 0
    return X, c, O.bc_matrix(O.ops_assemble(ops, nd), free), free


def test_sa_aggregation_covers_connected_nodes():
    X, c, A, free = _small_system((8, 3, 3))
    ptr, idx = amg_setup.block_graph(A, 3)
    agg, nagg = amg_setup.aggregate(ptr, idx, len(X))
    clamped = np.nonzero(X[:, 0] < 1e-9)[0]
    # fully clamped nodes are isolated (identity rows) and stay out of every aggregate
    assert (agg[clamped] == -1).all()
    others = np.setdiff1d(np.arange(len(X)), clamped)
    assert (agg[others] >= 0).all() and agg.max() == nagg - 1
    # every aggregate is connected to its seed: members are graph neighbours of one node
    G = sp.csr_matrix((np.ones(len(idx)), idx, ptr), shape=(len(X), len(X)))
    for a in range(nagg):
        mem = np.nonzero(agg == a)[0]
        assert len(mem) >= 1
        reach = np.asarray(G[mem].sum(0)).ravel() > 0
        assert reach[mem].all()


def test_sa_tentative_prolongator_reproduces_candidates():
    X, c, A, free = _small_system((8, 3, 3))
    B = amg_setup.rigid_body_modes(X)
    ptr, idx = amg_setup.block_graph(A, 3)
    agg, nagg = amg_setup.aggregate(ptr, idx, len(X))
    T, Bc = amg_setup.fit_candidates(agg, nagg, B, 3)
    rows = np.repeat(agg >= 0, 3)
    # T has orthonormal columns and T Bc = B on every aggregated row
    TtT = (T.T @ T).toarray()
    nzc = np.abs(TtT).sum(0) > 0
    assert np.allclose(TtT[np.ix_(nzc, nzc)], np.eye(nzc.sum()), atol=1e-12)
    assert np.allclose((T @ Bc)[rows], B[rows], atol=1e-10)


def test_sa_hierarchy_is_galerkin_and_coarsens():
    X, c, A, free = _small_system((12, 3, 3))
    B = amg_setup.rigid_body_modes(X)
    levels, Ac = amg_setup.build_hierarchy(A, B, 3, max_coarse=20)
    assert len(levels) >= 1
    for lv in levels:
        P, R, Al = lv["P"], lv["R"], lv["A"]
        assert (R - P.T).nnz == 0 or abs(R - P.T).max() == 0
        assert P.shape[0] == Al.shape[0]
    nxt = [lv["A"] for lv in levels[1:]] + [Ac]
    for lv, An in zip(levels, nxt):
        G = (lv["R"] @ lv["A"] @ lv["P"]).toarray()
        assert np.allclose(G, An.toarray(), rtol=1e-10, atol=1e-10 * np.abs(G).max())
        assert np.allclose(An.toarray(), An.toarray().T, atol=1e-8 * np.abs(G).max())
    assert Ac.shape[0] // 6 <= 20


def test_ops_assemble_matches_the_oracle():
    X, c = mesh.box_hex(2, 1, 1, 2, 1, 1)
    Ke = O.element_stiffness(O.HEX8, X[c], 200e3, 0.3)
    ops = [(O.dofs(c), Ke)]
    A1 = ops_assemble(ops, 3 * len(X))
    A0 = O.ops_assemble(ops, 3 * len(X))
    assert abs(A1 - A0).max() == 0


def _patches(conn, nn, nmax, max_elems=64):
    ne, nnode = conn.shape
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    cptr = np.zeros(ne + 1, np.int32)
    elem = np.zeros(max(ne, 1), np.int32)
    lmap = np.zeros(max(ne * nnode, 1), np.uint8)
    rbase = np.zeros(ne + 1, np.int32)
    rown = np.zeros(max(ne * nnode, 1), np.int32)
    nc = L.load().exfem_patch_build(ne, nnode, nn, conn.ctypes.data, nmax, max_elems, 7,
                                    cptr.ctypes.data, elem.ctypes.data, lmap.ctypes.data,
                                    rbase.ctypes.data, rown.ctypes.data)
    assert nc >= 0
    return nc, cptr[:nc + 1], elem[:ne], lmap[:ne * nnode].reshape(ne, nnode), rbase[:nc + 1], rown


@pytest.mark.parametrize("kind", ["hex", "tet10"])
@pytest.mark.parametrize("nmax", [24, 48, 120])
def test_patch_plan_covers_every_element_once(kind, nmax):
    """Patches of the batched element pass: each element in exactly one patch, at most 64
    elements and nmax rows per patch, every local node mapped to the partial row of its node,
    and (for the structured box) far fewer rows than element slots."""
    if kind == "hex":
        p = workloads.cantilever(12, 6, 5, jitter=0.1, seed=1)
    else:
        p = workloads.tet10_box(4, 4, 8, seed=2)
    conn = p.conn.astype(np.int32)
    ne, nnode = conn.shape
    nc, cptr, elem, lmap, rbase, rown = _patches(conn, p.nn, nmax)
    assert cptr[0] == 0 and cptr[-1] == ne and rbase[0] == 7
    assert np.array_equal(np.sort(elem), np.arange(ne))
    sizes = np.diff(cptr)
    rows = np.diff(rbase)
    assert (sizes >= 1).all() and (sizes <= 64).all()
    assert (rows >= nnode).all() and (rows <= nmax).all()
    for c in range(nc):
        ks = np.arange(cptr[c], cptr[c + 1])
        assert (lmap[ks] < rows[c]).all()
        prow = rbase[c] - 7 + lmap[ks].astype(np.int64)
        assert np.array_equal(rown[prow], conn[elem[ks]])
        assert len(np.unique(rown[rbase[c] - 7:rbase[c + 1] - 7])) == rows[c]
    if kind == "hex" and nmax >= 48:
        assert rows.sum() <= 0.45 * ne * nnode
    assert L.load().exfem_patch_build(ne, nnode, p.nn, conn.ctypes.data, nnode - 1, 64, 0,
                                      None, None, None, None, None) < 0


def test_native_sa_hierarchy_matches_the_scipy_route():
    """exfem_sa_level_build (native host SA level: graph, aggregation, fit_candidates, Jacobi
    smoothing, Galerkin product on blocks) == the scipy-product route of the same algorithm:
    identical aggregates, P and coarse matrices to rounding, same rho; the result does not
    depend on the thread count; bc_bsr == bc_matrix."""
    X, c, A, free = _small_system((16, 4, 4))
    Ab = sp.bsr_matrix(O.ops_assemble([(O.dofs(c), O.element_stiffness(O.HEX8, X[c], 200e3, 0.3))],
                                      3 * len(X)), blocksize=(3, 3))
    Abc = amg_setup.bc_bsr(Ab, free)
    assert abs(Abc.tocsr() - sp.csr_matrix(A)).max() == 0
    B = amg_setup.rigid_body_modes(X)
    L1, A1 = amg_setup.build_hierarchy(sp.csr_matrix(A), B, 3, max_coarse=20)
    L2, A2 = amg_setup.build_hierarchy_native(Abc, B, 3, max_coarse=20, nthreads=4)
    L3, A3 = amg_setup.build_hierarchy_native(Abc, B, 3, max_coarse=20, nthreads=1)
    assert [l["n"] for l in L1] == [l["n"] for l in L2] and len(L1) >= 1
    for a, b, c1 in zip(L1, L2, L3):
        assert np.array_equal(a["agg"], b["agg"])
        Pa = a["P"].toarray()
        assert np.abs(Pa - b["P"].toarray()).max() <= 1e-12 * np.abs(Pa).max()
        Aa = sp.csr_matrix(a["A"]).toarray()
        assert np.abs(Aa - b["A"].toarray()).max() <= 1e-12 * np.abs(Aa).max()
        assert abs(a["rho"] - b["rho"]) <= 1e-12 * a["rho"]
        assert np.array_equal(b["P"].toarray(), c1["P"].toarray()) and b["rho"] == c1["rho"]
    Ac = A1.toarray()
    assert np.abs(Ac - A2.toarray()).max() <= 1e-12 * np.abs(Ac).max()
    assert np.array_equal(A2.toarray(), A3.toarray())


def test_transformation_host_build_matches_the_oracle():
    """The library's native Transformation build (exfem_mpc_build: components of the slave graph,
    first constraint wins, prescribed slaves dropped) equals the oracle's restatement of the
    reference on random constraint sets whose slave cycles contract (|weights| summing below 1
    per row), where the reference's repeated squaring converges: to 1e-9 of max |T| (the
    squaring stops after 2^8 terms of the series the native build sums exactly)."""
    from exfem_amd.constraints import MPC, Transformation, coupling, normal_constraint
    rng = np.random.default_rng(11)
    nd = 60
    compared = 0
    for trial in range(40):
        mpcs = []
        for _ in range(rng.integers(1, 4)):
            k = int(rng.integers(1, 3))
            nc = int(rng.integers(1, 8))
            mpcs.append(MPC(rng.integers(0, nd, nc), rng.integers(0, nd, (nc, k)),
                            rng.uniform(-0.45, 0.45, (nc, k))))
        if trial % 3 == 0:
            mpcs.append(normal_constraint(rng.integers(0, nd // 3, 3), rng.standard_normal(3)))
        if trial % 4 == 0:
            mpcs.append(coupling(int(rng.integers(0, nd // 3)), rng.integers(0, nd // 3, 4)))
        fixed = rng.random(nd) < 0.1
        tr = Transformation(nd, mpcs, fixed)
        T, smask = O.transformation(nd, [(m.slave, m.masters, m.weights) for m in mpcs], fixed)
        assert np.array_equal(tr.slave_mask, smask)
        assert not tr.T[:, smask].nnz
        if tr.broken:
            continue
        compared += 1
        D = abs(tr.T - T)
        assert D.max() <= 1e-9 * max(abs(T).max(), 1.0), trial
        r, c, v = tr.triplets()
        assert np.array_equal(sp.coo_matrix((v, (r, c)), shape=(nd, nd)).toarray(), tr.T.toarray())
    assert compared >= 30
    assert not Transformation(nd, [], None).active


def test_transformation_chains_and_cycles_closed_forms():
    """Closed forms of the native build: a 40-deep chain u_i = 0.9 u_(i+1), u_39 = 0.9 u_m gives
    u_i = 0.9^(40-i) u_m (also the oracle's); the contracting cycle u_0 = 0.5 u_1 + 0.5 u_m,
    u_1 = u_0 has the fixed point u_0 = u_1 = u_m (the limit the oracle's squaring approaches);
    a pure cycle and a divergent one (u_0 = u_1 + u_m, u_1 = u_0) are broken: empty rows."""
    from exfem_amd.constraints import MPC, Transformation
    nd = 50
    m = 45
    chain = MPC(np.arange(40), np.append(np.arange(1, 40), m), np.full(40, 0.9))
    tr = Transformation(nd, [chain])
    Td = tr.T.toarray()
    assert tr.broken == 0 and tr.slave_mask[:40].all()
    assert np.allclose(Td[np.arange(40), m], 0.9 ** (40 - np.arange(40)), rtol=1e-14, atol=0)
    T, _ = O.transformation(nd, [(chain.slave, chain.masters, chain.weights)])
    assert abs(tr.T - T).max() <= 1e-15
    cyc = MPC([0, 1], [[1, m], [0, -1]], [[0.5, 0.5], [1.0, 0.0]])
    tr = Transformation(nd, [cyc])
    Td = tr.T.toarray()
    assert tr.broken == 0 and abs(Td[0, m] - 1.0) <= 1e-15 and abs(Td[1, m] - 1.0) <= 1e-15
    T, _ = O.transformation(nd, [(cyc.slave, cyc.masters, cyc.weights)])
    assert abs(tr.T - T).max() <= 1e-12
    for rows in (MPC([0, 1], [[1], [0]], [[1.0], [1.0]]),
                 MPC([0, 1], [[1, m], [0, -1]], [[1.0, 1.0], [1.0, 0.0]])):
        tr = Transformation(nd, [rows])
        assert tr.broken == 1
        assert not tr.T.toarray()[:2].any()
    # prescribed slaves drop out and keep identity rows; errors are reported, not crashed on
    tr = Transformation(nd, [chain], fixed=np.arange(nd) == 3)
    assert tr.dropped == 1 and not tr.slave_mask[3] and tr.T[3, 3] == 1.0
    assert tr.T[2, 3] == 0.9
    from exfem_amd import _lib as L
    with pytest.raises(L.ExfemError):
        Transformation(nd, [MPC([nd + 1], [[0]], [[1.0]])])


def test_j2_material_validates_its_input():
    import exfem_amd as X
    with pytest.raises(ValueError):
        X.J2Plasticity(210e3, 0.3, [[250, 0.1], [300, 0.05]])          # strains must increase
    with pytest.raises(ValueError):
        X.J2Plasticity(210e3, 0.3)                                      # no hardening given
    with pytest.raises(NotImplementedError):
        X.J2Plasticity(210e3, 0.3, johnson_cook=dict(A=1, B=1, n=0.5, Tmelt=1500))
    m = X.J2Plasticity(210e3, 0.3, [[250, 0]], Hk=10.0)
    assert m.nstate == 19 and not m.linear and m.sy0 == 250.0


def test_sa_aggregation_equals_the_oracle_on_directed_graphs():
    """exfem_sa_aggregate == the oracle's restatement of PyAMG's standard aggregation on random
    graphs, directed ones included (they reach pass 3 and its renumbering)."""
    rng = np.random.default_rng(7)
    for trial in range(40):
        n = int(rng.integers(2, 40))
        nbr = []
        for i in range(n):
            k = int(rng.integers(0, 4)) if rng.random() > 0.15 else 0
            nbr.append(sorted(set(int(j) for j in rng.integers(0, n, k)) - {i}))
        ptr = np.zeros(n + 1, np.int64)
        ptr[1:] = np.cumsum([len(r) for r in nbr])
        idx = np.array([j for r in nbr for j in r] or [0], np.int32)
        agg, nagg = amg_setup.aggregate(ptr, idx, n)
        ragg, rn = O.sa_standard_aggregation(nbr)
        assert nagg == rn and np.array_equal(agg, ragg), trial


@pytest.mark.parametrize("kind", ["hex", "tet10"])
def test_native_sa_hierarchy_equals_the_oracle(kind):
    """The library's SA hierarchy (exfem_sa_level_build, every level) == the oracle's
    restatement of PyAMG's set-up on a clamped C3D8 box and a C3D10 box: identical aggregates,
    P and the coarse matrices to 1e-12, the same level count."""
    if kind == "hex":
        X, c, A, free = _small_system((10, 4, 4))
        B = O.rigid_body_modes(X)
    else:
        p = workloads.tet10_box(3, 3, 6, seed=2)
        X, c = p.X, p.conn
        ops = [(O.dofs(c), O.element_stiffness(O.TET10, X[c], 1.0, 0.3, scale=p.scale))]
        free = p.free.astype(float)
        A = O.bc_matrix(O.ops_assemble(ops, 3 * len(X)), free)
        B = O.rigid_body_modes(X)
    ref, Ac_ref = O.sa_hierarchy(A, B, 3, max_coarse=8)
    got, Ac_got = amg_setup.build_hierarchy_native(sp.csr_matrix(A), B, 3, max_coarse=8, nthreads=2)
    assert len(ref) == len(got) >= 2
    for r, g in zip(ref, got):
        assert np.array_equal(r["agg"], g["agg"])
        Pr = r["P"].toarray()
        assert np.abs(Pr - g["P"].toarray()).max() <= 1e-12 * np.abs(Pr).max()
        Ar = r["A"].toarray()
        assert np.abs(Ar - sp.csr_matrix(g["A"]).toarray()).max() <= 1e-12 * np.abs(Ar).max()
    Ar = Ac_ref.toarray()
    assert np.abs(Ar - Ac_got.toarray()).max() <= 1e-12 * np.abs(Ar).max()


def test_rho_start_is_the_reference_draw():
    """The cached power-estimate start vector equals default_rng(0).standard_normal(n) for every
    n, before and after the cache grows."""
    for n in (5, 1000, 3, 250000, 7, 99991):
        assert np.array_equal(amg_setup.rho_start(n), np.random.default_rng(0).standard_normal(n))
