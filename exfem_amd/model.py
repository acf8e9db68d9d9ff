"""Model, element blocks and EbE operators on the MI355X path (gfem/model.py:21-499).

Device arrays are torch tensors on ``cuda`` (PyTorch is used for memory and streams); all
arithmetic runs in the HIP library through the C ABI. Element matrices live on the device in
the packed element-interleaved layout of the library (``KeBlock``); ``KeBlock.numpy()`` returns
the reference's (ne, nd, nd) layout.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .elements import ElementType

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("exfem_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return t.device("cuda")


def stream_ptr():
    return C.c_void_p(torch().cuda.current_stream().cuda_stream)


def vptr(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _tdt(dtype):
    return torch().float64 if dtype == L.F64 else torch().float32


def as_dev(a, dtype=L.F64):
    """Contiguous device tensor of `dtype` (F64/F32) from numpy / torch / scalars."""
    t = torch()
    td = _tdt(dtype)
    if isinstance(a, t.Tensor):
        return a.to(device=device(), dtype=td).contiguous()
    return t.as_tensor(np.ascontiguousarray(a), dtype=td, device=device()).contiguous()


def _check_out(y, x, what):
    """An output buffer the library writes as many values of x's dtype as x holds: refuse one of
    another dtype, size, device or layout instead of letting the kernel write past its end."""
    t = torch()
    if not isinstance(y, t.Tensor) or y.dtype != x.dtype or y.numel() != x.numel() \
            or y.device != x.device or not y.is_contiguous():
        got = (f"{tuple(y.shape)} {y.dtype} on {y.device}" if isinstance(y, t.Tensor)
               else type(y).__name__)
        raise ValueError(f"{what}: output must be a contiguous {x.dtype} device tensor of "
                         f"{x.numel()} values, got {got}")


def to_numpy(a):
    t = torch()
    if isinstance(a, t.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


class KeBlock:
    """Packed element matrices of one element block, resident on the device."""

    def __init__(self, etype, ne, storage, dtype, buf):
        self.etype, self.ne, self.storage, self.dtype, self.buf = etype, ne, storage, dtype, buf

    @property
    def nd(self):
        return self.etype.nd

    @property
    def shape(self):
        return (self.ne, self.nd, self.nd)

    def ptr(self):
        return vptr(self.buf)

    def nbytes(self):
        return self.buf.numel()

    @classmethod
    def empty(cls, etype, ne, storage=L.KE_FULL, dtype=L.F64):
        lib = L.load()
        nb = lib.exfem_ke_bytes(etype.code, ne, storage, dtype)
        L.check(nb)
        buf = torch().empty(max(nb, 1), dtype=torch().uint8, device=device())
        return cls(etype, ne, storage, dtype, buf)

    @classmethod
    def from_array(cls, etype, Ke, storage=L.KE_FULL, dtype=L.F64):
        """Pack reference-layout matrices (ne, nd, nd)."""
        Kd = as_dev(Ke, L.F64)
        ne = Kd.shape[0]
        out = cls.empty(etype, ne, storage, dtype)
        L.check(L.load().exfem_ke_pack(etype.nd, ne, vptr(Kd), storage, dtype, out.ptr(),
                                       stream_ptr()))
        return out

    def to_array(self):
        """(ne, nd, nd) float64 device tensor in the reference layout."""
        out = torch().empty((self.ne, self.nd, self.nd), dtype=torch().float64, device=device())
        L.check(L.load().exfem_ke_unpack(self.nd, self.ne, self.ptr(), self.storage, self.dtype,
                                         vptr(out), stream_ptr()))
        return out

    def numpy(self):
        return to_numpy(self.to_array())

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a.astype(dtype) if dtype is not None else a

    def astype(self, dtype, storage=None):
        storage = self.storage if storage is None else storage
        if dtype == self.dtype and storage == self.storage:
            return self
        return KeBlock.from_array(self.etype, self.to_array(), storage, dtype)


class Block:
    """One element type + one material (gfem/model.py:21-26,154-158)."""

    def __init__(self, conn, etype, material, name="block"):
        self.conn = np.ascontiguousarray(np.asarray(conn, dtype=np.int32))
        self.etype, self.material, self.name = etype, material, name
        self.ne = int(self.conn.shape[0])
        self._conn_dev = None

    def conn_dev(self):
        if self._conn_dev is None:
            self._conn_dev = torch().as_tensor(self.conn, device=device())
        return self._conn_dev

    def dofs(self):
        """(ne, 3 nnode) node-interleaved dofs (gfem/model.py:157-158). The same array on every
        call, so operators cached by it are found again."""
        if getattr(self, "_dofs", None) is None:
            self._dofs = (self.conn.astype(np.int64)[:, :, None] * 3 + np.arange(3)).reshape(self.ne, -1)
            self._dofs.flags.writeable = False
        return self._dofs

    @property
    def plastic(self):
        """A state-carrying material (J2)."""
        return not getattr(self.material, "linear", True) and getattr(self.material, "nstate", 1) > 1

    @property
    def hyper(self):
        """A strain-energy (hyperelastic) material."""
        return getattr(self.material, "kind", -1) >= 0 and hasattr(self.material, "params")

    @property
    def nonlinear(self):
        """Its tangent depends on u (J2 or hyperelastic)."""
        return not getattr(self.material, "linear", True)

    def init_state(self):
        """Zero state (ne, ngp, nstate) (gfem/model.py:150-151) for a state-carrying material;
        None for a linear one. The tensor is a view of the library's structure-of-arrays
        storage (nstate, ngp, ne)."""
        if not self.plastic:
            return None
        soa = torch().zeros((self.material.nstate, self.etype.ngp, self.ne), dtype=torch().float64,
                            device=device())
        return soa.permute(2, 1, 0)


def state_soa(s):
    """Structure-of-arrays (nstate, ngp, ne) storage of a state array (ne, ngp, nstate): the
    tensor itself when it is such a view, else a contiguous copy."""
    T = torch()
    if isinstance(s, T.Tensor) and s.is_cuda and s.dtype == T.float64:
        soa = s.permute(2, 1, 0)
        if soa.is_contiguous():
            return soa                     # the library's own layout: no copy
    return as_dev(s).permute(2, 1, 0).contiguous()


class EbeOperator:
    """Handle of the device EbE operator over element blocks sharing nn nodes."""

    def __init__(self, nn, etypes, conns):
        lib = L.load()
        self.nn, self.ndof = int(nn), 3 * int(nn)
        self.etypes = list(etypes)
        self._conns = [np.ascontiguousarray(np.asarray(c, np.int32)) for c in conns]
        nb = len(self._conns)
        et = (C.c_int * nb)(*[e.code for e in self.etypes])
        ne = (C.c_int64 * nb)(*[len(c) for c in self._conns])
        cp = (C.c_void_p * nb)(*[c.ctypes.data for c in self._conns])
        h = C.c_void_p()
        device()
        L.check(lib.exfem_op_create(C.byref(h), self.nn, nb, et, ne, cp))
        self.h = h
        self.ke = [None] * nb
        # device buffers the library reads through raw pointers: one attached element-matrix set
        # per (block, dtype) and the current geometry; replacing one releases the previous
        self._ke_ref = {}
        self._geom_ref = ()
        self._rest = {}                 # rest slot -> its connectivity (host)
        self.has_bc = False
        self.has_geometry = False
        self.assembled = False

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and L is not None and L._lib is not None:
            try:
                L._lib.exfem_op_destroy(h)
            except Exception:
                pass
            self.h = None

    def set_ke(self, block, ke):
        """Attach block's element matrices of ke.dtype, replacing (and releasing) the ones
        attached before for that block and dtype."""
        L.check(L.load().exfem_op_set_ke(self.h, block, ke.ptr(), ke.storage, ke.dtype))
        self._ke_ref[(block, ke.dtype)] = ke
        return self

    def clear_ke(self, block, dtype):
        """Detach block's element matrices of `dtype` (the matrix-free operator needs none)."""
        L.check(L.load().exfem_op_set_ke(self.h, block, None, L.KE_FULL, dtype))
        self._ke_ref.pop((block, dtype), None)
        return self

    def attached_ke(self, block, dtype):
        """The KeBlock attached for (block, dtype), or None."""
        return self._ke_ref.get((block, dtype))

    def set_rest(self, slot, conn, Ke=None):
        """Rest operator `slot` (exfem_op_set_rest): elements conn (ne, nnode) with matrices Ke
        (ne, 3 nnode, 3 nnode) added to every product (None: zero matrices, kept only for
        rest_assemble). An empty conn clears the slot."""
        conn = np.ascontiguousarray(np.asarray(conn, dtype=np.int32))
        if conn.size == 0:
            L.check(L.load().exfem_op_set_rest(self.h, slot, 0, 0, None, None, stream_ptr()))
            self._rest.pop(slot, None)
            return self
        Kd = None if Ke is None else as_dev(Ke)
        L.check(L.load().exfem_op_set_rest(self.h, slot, conn.shape[0], conn.shape[1],
                                           conn.ctypes.data, vptr(Kd), stream_ptr()))
        self._rest[slot] = conn
        return self

    def ensure_rest(self, slot, conn):
        """Slot `slot` holds elements conn (set with zero matrices if it holds others)."""
        cur = self._rest.get(slot)
        conn = np.asarray(conn, dtype=np.int32)
        if cur is None or cur.shape != conn.shape or not np.array_equal(cur, conn):
            self.set_rest(slot, conn, None)
        return self

    def clear_rest(self, first=0):
        """Clear every rest slot from `first` on."""
        for slot in [k for k in self._rest if k >= first]:
            self.set_rest(slot, np.zeros((0, 1), np.int32))
        return self

    def rest_assemble(self, slot, fe, y):
        """y += assembled element vectors fe (ne, 3 nnode) of rest slot `slot`."""
        L.check(L.load().exfem_op_rest_assemble(self.h, slot, vptr(as_dev(fe)), vptr(y), stream_ptr()))
        return y

    def set_assembled(self, on=True):
        """Products on the device-assembled 3x3 block CSR of the element blocks
        (operator="csr", gfem/solvers.py:406-414; csr_matvec gfem/model.py:237-238) instead of
        element by element; the rest operators stay element by element, the matrix-free
        products and the diagonal are unchanged."""
        on = bool(on)
        if on != self.assembled:
            L.check(L.load().exfem_op_set_assembled(self.h, int(on)))
            self.assembled = on
        return self

    def set_bc(self, free):
        free = None if free is None else np.ascontiguousarray(np.asarray(free) > 0, dtype=np.uint8)
        L.check(L.load().exfem_op_set_bc(self.h, None if free is None else free.ctypes.data))
        self.has_bc = free is not None and not free.all()
        return self

    def set_geometry(self, X, scales, E, nu):
        nb = len(self._conns)
        Xd = as_dev(X)
        sc = [None if s is None else as_dev(s) for s in scales]
        arr = (C.c_void_p * nb)(*[0 if s is None else s.data_ptr() for s in sc])
        Ea = (C.c_double * nb)(*E)
        na = (C.c_double * nb)(*nu)
        L.check(L.load().exfem_op_set_geometry(self.h, vptr(Xd), arr, Ea, na))
        self._geom_ref = tuple([Xd] + [s for s in sc if s is not None])
        self.has_geometry = True
        return self

    def matvec(self, x, y=None, bc=False, matrix_free=False, dot=False, dtype=L.F64):
        xd = as_dev(x, dtype)
        if y is None:
            y = torch().empty_like(xd)
        else:
            _check_out(y, xd, "matvec")
        d = torch().empty(1, dtype=torch().float64, device=device()) if dot else None
        L.check(L.load().exfem_op_matvec(self.h, vptr(xd), vptr(y), dtype, int(bc),
                                         int(matrix_free), vptr(d), stream_ptr()))
        return (y, d) if dot else y

    def matvec_multi(self, X, Y=None, bc=False, dtype=L.F64):
        Xd = as_dev(X, dtype)
        if Y is None:
            Y = torch().empty_like(Xd)
        else:
            _check_out(Y, Xd, "matvec_multi")
        L.check(L.load().exfem_op_matvec_multi(self.h, Xd.shape[0], vptr(Xd), vptr(Y), dtype,
                                               int(bc), stream_ptr()))
        return Y

    def multi_shape(self, nrhs, dtype=L.F64):
        """Variant the autotuner chose for the batched product: dict(index, ntw, blocks, nmax,
        ms) - a slot variant has ntw/blocks, a patch variant its node cap nmax, the product on
        the assembled operator nmax = -1 - or None before the first batched product with this
        dtype and case-tile count."""
        v = [C.c_int() for _ in range(4)]
        ms = C.c_float()
        L.check(L.load().exfem_op_multi_shape(self.h, nrhs, dtype, *[C.byref(x) for x in v],
                                              C.byref(ms)))
        if v[0].value < 0:
            return None
        return dict(index=v[0].value, ntw=v[1].value, blocks=v[2].value, nmax=v[3].value,
                    ms=ms.value)

    def set_multi_shape(self, nrhs, index, dtype=L.F64):
        """Pin the batched product to variant `index` (None: let the autotuner choose again)."""
        L.check(L.load().exfem_op_set_multi_shape(self.h, nrhs, dtype,
                                                  -1 if index is None else int(index)))

    def diagonal(self):
        d = torch().empty(self.ndof, dtype=torch().float64, device=device())
        L.check(L.load().exfem_op_diagonal(self.h, vptr(d), stream_ptr()))
        return d

    def assemble(self, bc=False):
        """Assembled matrix sum_e P^T Ke P (3x3 block CSR, scipy) summed on the device from the
        attached FP64 element matrices (gfem/model.py:176-234). bc=True returns bc_matrix of it
        (gfem/linalg.py:213-217) for the free mask given to set_bc, formed on the device."""
        import scipy.sparse as sp
        lib = L.load()
        nblk = C.c_int64()
        L.check(lib.exfem_op_assemble_bsr(self.h, C.byref(nblk), None, None, None, int(bc), stream_ptr()))
        indptr = np.empty(self.nn + 1, np.int32)
        indices = np.empty(max(nblk.value, 1), np.int32)
        vals = np.empty((max(nblk.value, 1), 3, 3))
        L.check(lib.exfem_op_assemble_bsr(self.h, C.byref(nblk), indptr.ctypes.data,
                                          indices.ctypes.data, vals.ctypes.data, int(bc), stream_ptr()))
        n = nblk.value
        return sp.bsr_matrix((vals[:n], indices[:n], indptr), shape=(self.ndof, self.ndof))

    def bsr_nblocks(self):
        """Number of 3x3 blocks of the assembled matrix (builds the pattern on first use)."""
        nblk = C.c_int64()
        L.check(L.load().exfem_op_assemble_bsr(self.h, C.byref(nblk), None, None, None, 0, stream_ptr()))
        return nblk.value

    def internal_force(self, u):
        ud = as_dev(u)
        f = torch().empty_like(ud)
        L.check(L.load().exfem_op_internal_force(self.h, vptr(ud), vptr(f), stream_ptr()))
        return f

    def stress(self, block, u):
        """Volume-averaged Cauchy stress of each element of a block at u (exfem_op_stress):
        (ne, 3, 3) device f64."""
        ud = as_dev(u)
        sig = torch().empty((len(self._conns[block]), 9), dtype=torch().float64, device=device())
        L.check(L.load().exfem_op_stress(self.h, block, vptr(ud), vptr(sig), stream_ptr()))
        return sig.view(-1, 3, 3)

    def geo_ke(self, block, etype, sig, storage=L.KE_FULL, dtype=L.F64):
        """Initial-stress element matrices of a block for a constant Cauchy stress per element
        (exfem_op_geo_ke)."""
        kb = KeBlock.empty(etype, len(self._conns[block]), storage, dtype)
        sd = as_dev(sig).reshape(-1, 9).contiguous()
        L.check(L.load().exfem_op_geo_ke(self.h, block, vptr(sd), storage, dtype, kb.ptr(), stream_ptr()))
        kb._keep = sd
        return kb

    def set_fields(self, T, T0, alphas, eigs):
        """Thermal strain and eigenstrain fields (exfem_op_set_fields): nodal T and T0 (None = 0),
        the blocks' expansion coefficients and per-block eigenstrains ((ne, 3, 3) / (ne, 9) or
        None); T = None and eigs = None clear them."""
        nb = len(self._conns)
        Td = None if T is None else as_dev(T).contiguous()
        T0d = None if (T is None or T0 is None) else as_dev(T0).contiguous()
        eigs = list(eigs) if eigs is not None else [None] * nb
        ed = [None if e is None else as_dev(e).reshape(-1, 9).contiguous() for e in eigs]
        al = (C.c_double * nb)(*([0.0] * nb if alphas is None else [float(a) for a in alphas]))
        ep = (C.c_void_p * nb)(*[0 if e is None else e.data_ptr() for e in ed])
        L.check(L.load().exfem_op_set_fields(self.h, vptr(Td), vptr(T0d), al,
                                             ep if any(e is not None for e in ed) else None))
        self._fields_ref = (Td, T0d, ed)
        return self

    def set_hyper(self, block, mat):
        """Strain-energy material of a block (exfem_op_set_hyper)."""
        p = (C.c_double * 6)(*(list(mat.params()) + [0.0] * 6)[:6])
        L.check(L.load().exfem_op_set_hyper(self.h, block, mat.kind, p))

    def set_j2(self, block, mat):
        """J2 material of a block (exfem_op_set_j2)."""
        if mat.jc is not None:
            jc = (C.c_double * 3)(mat.jc["A"], mat.jc["B"], mat.jc["n"])
            L.check(L.load().exfem_op_set_j2(self.h, block, mat.E, mat.nu, mat.Hk, 0, None, jc))
        else:
            t = np.ascontiguousarray(mat.table, dtype=np.float64)
            L.check(L.load().exfem_op_set_j2(self.h, block, mat.E, mat.nu, mat.Hk, t.shape[0],
                                             t.ctypes.data, None))

    def internal_force_state(self, u, states):
        """(f, new states) with J2 blocks returning from `states` (exfem_op_internal_force_state);
        states: per block (ne, ngp, 19) or None for linear blocks."""
        ud = as_dev(u)
        f = torch().empty_like(ud)
        nb = len(self._conns)
        old = [None if st is None else state_soa(st) for st in states]
        new = [None if st is None else torch().empty_like(st) for st in old]
        so = (C.c_void_p * nb)(*[0 if t is None else t.data_ptr() for t in old])
        sn = (C.c_void_p * nb)(*[0 if t is None else t.data_ptr() for t in new])
        L.check(L.load().exfem_op_internal_force_state(self.h, vptr(ud), so, sn, vptr(f),
                                                       stream_ptr()))
        return f, [None if t is None else t.permute(2, 1, 0) for t in new]

    def set_state(self, u, states):
        """The point (u, states) the matrix-free form linearises J2 blocks at
        (exfem_op_set_state): the return runs here, on the current stream, and its tangent data
        stays on the device; u None clears it."""
        lib = L.load()
        if u is None:
            L.check(lib.exfem_op_set_state(self.h, None, None, None))
            return self
        ud = as_dev(u)
        soa = [None if st is None else state_soa(st) for st in states]
        arr = (C.c_void_p * len(self._conns))(*[0 if t is None else t.data_ptr() for t in soa])
        L.check(lib.exfem_op_set_state(self.h, vptr(ud), arr, stream_ptr()))
        return self

    def ke_state(self, block, etype, u, state, storage=L.KE_FULL, dtype=L.F64):
        """Element matrices of a J2 block at (u, state) or a hyperelastic block at u (state
        None) (exfem_op_ke_state)."""
        kb = KeBlock.empty(etype, len(self._conns[block]), storage, dtype)
        ud = as_dev(u)
        so = None if state is None else state_soa(state)
        L.check(L.load().exfem_op_ke_state(self.h, block, vptr(ud), vptr(so), storage, dtype,
                                           kb.ptr(), stream_ptr()))
        kb._keep = (ud, so)
        return kb


def _conn_from_dofs(d, nnode):
    d = np.asarray(to_numpy(d))
    if d.ndim != 2 or d.shape[1] != 3 * nnode:
        raise ValueError("dofs must be (ne, 3*nnode)")
    c = d[:, 0::3] // 3
    if not (np.all(d[:, 1::3] == 3 * c + 1) and np.all(d[:, 2::3] == 3 * c + 2)):
        raise ValueError("the device operator needs node-interleaved dofs (3*node + comp)")
    return c.astype(np.int32)


_OP_CACHE = {}


def _operator_for(ops, ndof):
    """Device operator for a list of (dofs, Ke) pairs (cached by dofs identity): pairs of a
    solid element type become element blocks, the others (springs, contact pairs: 1-4 nodes)
    rest operators."""
    from .elements import BY_NNODE
    blocks, rest = [], []
    for d, K in ops:
        nd = (K.nd if isinstance(K, KeBlock) else np.shape(K)[1])
        (blocks if (isinstance(K, KeBlock) or nd // 3 in BY_NNODE) else rest).append((d, K))
    key = (ndof,) + tuple((id(d), K.shape if hasattr(K, "shape") else None) for d, K in blocks)
    ent = _OP_CACHE.get(key)
    kes = []
    for d, K in blocks:
        nd = (K.nd if isinstance(K, KeBlock) else np.shape(K)[1])
        et = K.etype if isinstance(K, KeBlock) else BY_NNODE[nd // 3]
        kes.append(K if isinstance(K, KeBlock) else KeBlock.from_array(et, K))
    if ent is None:
        conns = [_conn_from_dofs(d, k.nd // 3) for (d, _), k in zip(blocks, kes)]
        op = EbeOperator(ndof // 3, [k.etype for k in kes], conns)
        _OP_CACHE.clear()
        _OP_CACHE[key] = (op, [d for d, _ in blocks])
    else:
        op = ent[0]
    for b, k in enumerate(kes):
        op.set_ke(b, k)
    set_rest_ops(op, rest)
    return op


def set_rest_ops(op, rest):
    """Attach (dofs, Ke) pairs as the operator's rest operators, slot by slot; clear the slots
    after them (gfem/solvers.py:257-267: ops_matvec of ops["rest"])."""
    if len(rest) > 8:
        raise NotImplementedError("more than 8 rest operator groups")
    for j, (d, K) in enumerate(rest):
        nd = np.shape(K)[-1]
        op.set_rest(j, _conn_from_dofs(d, nd // 3), K)
    op.clear_rest(len(rest))
    return op


def ops_matvec(ops, v, ndof):
    """y = sum_blocks P^T Ke P v (gfem/model.py:162-166), deterministic device assembly."""
    if not ops:
        return torch().zeros(ndof, dtype=torch().float64, device=device())
    return _operator_for(ops, ndof).matvec(v)


def ops_diagonal(ops, ndof):
    """diag(K) (gfem/model.py:169-173)."""
    if not ops:
        return torch().zeros(ndof, dtype=torch().float64, device=device())
    return _operator_for(ops, ndof).diagonal()


def ops_assemble(ops, ndof):
    """Host scipy CSR with summed duplicates (gfem/model.py:176-192)."""
    import scipy.sparse as sp
    rows, cols, vals = [], [], []
    for d, K in ops:
        d = np.asarray(to_numpy(d), np.int64)
        Kh = K.numpy() if isinstance(K, KeBlock) else np.asarray(to_numpy(K))
        nd = d.shape[1]
        rows.append(np.repeat(d, nd, axis=1).ravel())
        cols.append(np.tile(d, (1, nd)).ravel())
        vals.append(Kh.reshape(len(d), -1).ravel())
    if not rows:
        return sp.csr_matrix((ndof, ndof))
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ndof, ndof)).tocsr()
    A.sum_duplicates()
    return A


def gather_plan(conns, nn):
    """Host gather plan in the reference format (gfem/model.py:241-261): per node the stable
    (element, local node) slots padded with the sentinel `total slots`."""
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
    return g.astype(np.int32 if off < 2 ** 31 else np.int64)


def gather_matvec(plan, ops, v, ndof):
    """gfem/model.py:264-267. The device operator builds the same per-node slot order as
    `plan` (stable (block, element, local node) order), so `plan` is accepted and unused."""
    return ops_matvec(ops, v, ndof)


class Model:
    """Nodes + element blocks (gfem/model.py:270-499, solid subset: linear-elastic and J2
    blocks, MPC constraints)."""

    def __init__(self, nodes):
        self.X = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64))
        self.nn = int(self.X.shape[0])
        self.ndof = 3 * self.nn
        self.blocks = []
        self.nsets, self.elsets, self.surfaces = {}, {}, {}
        self.constraints, self.interactions = [], []
        self.discrete = []              # springs (gfem/discrete.py), rest operators of the device op
        self.rot_of = {}
        self.T0 = None
        self._X_dev = None
        self._op = None
        self._op_nb = 0
        self._op_key = None

    # ------------------------------------------------------------------ building
    def add_constraint(self, c):
        """An MPC (constraints.MPC, normal_constraint, coupling) (gfem/model.py:372-374)."""
        self.constraints.append(c)
        return c

    def add_spring(self, conn, k, kind="axial", name=None):
        """Two-node springs (gfem/model.py:345-349): kind "axial" or "bushing"."""
        from .discrete import SpringBlock
        b = SpringBlock(conn, k, kind, name or f"spring{len(self.blocks) + len(self.discrete)}")
        self.discrete.append(b)
        return b

    def add_ground_spring(self, nodes, k, name=None):
        """Node-to-ground springs (gfem/model.py:351-355)."""
        from .discrete import GroundSpring
        b = GroundSpring(nodes, k, name or f"gspring{len(self.blocks) + len(self.discrete)}")
        self.discrete.append(b)
        return b

    def add_interaction(self, c):
        """A contact or rigid wall (gfem/model.py:376-377)."""
        self.interactions.append(c)
        return c

    def add_block(self, conn, etype, material, name=None):
        if isinstance(etype, str):
            from .elements import ELEMENTS
            etype = ELEMENTS[etype]
        if not isinstance(etype, ElementType):
            raise TypeError("etype must be an exfem_amd ElementType")
        b = Block(conn, etype, material, name or f"block{len(self.blocks)}")
        self.blocks.append(b)
        return b

    def init_state(self):
        return [b.init_state() for b in self.blocks]

    def rot_mask(self):
        return np.zeros(self.nn, bool)

    @property
    def nelem(self):
        return sum(b.ne for b in self.blocks)

    def X_dev(self):
        if self._X_dev is None:
            self._X_dev = as_dev(self.X)
        return self._X_dev

    def gather(self):
        return gather_plan([b.conn for b in self.blocks], self.nn)

    def summary(self):
        lines = [f"{self.nn} nodes, {self.nelem} elements, {self.ndof} dofs"]
        for b in self.blocks:
            lines.append(f"  {b.name}: {b.ne} x {b.etype.name}, {b.material.describe()}")
        return "\n".join(lines)

    # ------------------------------------------------------------------ fields
    def _scales(self, fields):
        f = fields or {}
        sc = f.get("scale") or [None] * len(self.blocks)
        out = []
        for b, s in zip(self.blocks, sc):
            out.append(None if s is None else as_dev(s))
        return out

    def _field_args(self, fields):
        """Nodal T, T0 and per-block eigenstrains of Model._fields (gfem/model.py:446-463), or
        None when neither a temperature nor an eigenstrain is set."""
        f = fields or {}
        T = f.get("T")
        eig = f.get("eig") or getattr(self, "eigenstrain", None)
        if T is None and eig is None:
            return None
        T0 = None if self.T0 is None else self.T0
        return T, T0, eig

    def operator(self, fields=None):
        """The cached device operator of this model (plan, packed connectivity, geometry,
        fields)."""
        op = self._operator_geometry(fields)
        fa = self._field_args(fields)
        fkey = None if fa is None else (id(fa[0]), id(fa[2]), id(self.T0))
        if getattr(self, "_fkey", None) != fkey:
            if fa is None:
                op.set_fields(None, None, None, None)
            else:
                T, T0, eig = fa
                op.set_fields(T, T0, [getattr(b.material, "alpha", 0.0) for b in self.blocks], eig)
            self._fkey = fkey
        return op

    def _operator_geometry(self, fields=None):
        key = (len(self.blocks), id(fields.get("scale")) if fields else None)
        if self._op is None or self._op_nb != len(self.blocks):
            self._op = EbeOperator(self.nn, [b.etype for b in self.blocks],
                                   [b.conn for b in self.blocks])
            self._op_nb = len(self.blocks)
            self._op_key = None
            self._fkey = None
        if self._op_key != key:
            self._op.set_geometry(self.X_dev(), self._scales(fields),
                                  [b.material.E for b in self.blocks],
                                  [b.material.nu for b in self.blocks])
            for i, b in enumerate(self.blocks):
                if b.plastic:
                    self._op.set_j2(i, b.material)
                elif b.hyper:
                    self._op.set_hyper(i, b.material)
            self._op_key = key
        return self._op

    def live_operator(self):
        """The cached device operator as it stands (geometry and fields unchanged; made with
        the unscaled geometry if there is none yet). Products with stored element matrices do
        not read the geometry, so they need not reset it."""
        if self._op is None or self._op_nb != len(self.blocks) or self._op_key is None:
            return self.operator()
        return self._op

    @property
    def plastic(self):
        """Any block with a state-carrying (J2) material."""
        return any(b.plastic for b in self.blocks)

    @property
    def nonlinear(self):
        """Any block whose tangent depends on u (J2 or hyperelastic)."""
        return any(b.nonlinear for b in self.blocks)

    # ------------------------------------------------------------------ element kernels
    def element_stiffness(self, u=None, states=None, fields=None, storage=L.KE_FULL,
                          dtype=L.F64):
        """Block.stiff per block (gfem/model.py:479-483) as packed device matrices. A linear
        material's tangent does not depend on u."""
        out = []
        scales = self._scales(fields)
        lib = L.load()
        for i, (b, sc) in enumerate(zip(self.blocks, scales)):
            if b.nonlinear:         # consistent tangent at (u, s_old), gfem/model.py:83-111
                op = self.operator(fields)
                uu = torch().zeros(self.ndof, dtype=torch().float64, device=device()) if u is None else u
                st = None
                if b.plastic:
                    st = b.init_state() if states is None or states[i] is None else states[i]
                out.append(op.ke_state(i, b.etype, uu, st, storage, dtype))
                continue
            kb = KeBlock.empty(b.etype, b.ne, storage, dtype)
            L.check(lib.exfem_ke_compute(b.etype.code, b.ne, self.nn, vptr(self.X_dev()),
                                         vptr(b.conn_dev()), vptr(sc), b.material.E,
                                         b.material.nu, storage, dtype, kb.ptr(), stream_ptr()))
            out.append(kb)
        return out

    def stiffness_ops(self, u=None, states=None, fields=None, storage=L.KE_FULL, dtype=L.F64):
        """[(dofs, Ke)] (gfem/model.py:485-486): the element blocks (packed KeBlock), then the
        discrete springs ((ne, nd, nd) device tensors at u)."""
        ops = list(zip([b.dofs() for b in self.blocks],
                       self.element_stiffness(u, states, fields, storage, dtype)))
        if self.discrete:
            uu = torch().zeros(self.ndof, dtype=torch().float64, device=device()) if u is None else u
            for b in self.discrete:
                ops.append((b.dofs(), b.evaluate(self.X_dev(), uu, force=False, stiff=True)[1]))
        return ops

    def internal_force(self, u, states=None, fields=None):
        """Model.internal_force (gfem/model.py:466-477): returns (f, new states). J2 blocks
        return from `states` (the converged states of the last increment)."""
        op = self.operator(fields)
        if not self.plastic:
            f, st = op.internal_force(u), states
        else:
            f, st = op.internal_force_state(u, self.init_state() if states is None else states)
        for i, b in enumerate(self.discrete):      # spring forces, assembled per node
            fe, _ = b.evaluate(self.X_dev(), u, force=True, stiff=False)
            op.ensure_rest(i, b.conn)
            op.rest_assemble(i, fe, f)
        return f, st

    def stresses(self, u, states=None, fields=None):
        """Model.stresses (gfem/model.py:529-537): per block dict(S (ne, 3, 3) volume-averaged
        Cauchy stress, MISES (ne,), PEEQ (ne,) zeros), on the device. J2 blocks are not
        supported."""
        t = torch()
        op = self.operator(fields)
        out = []
        for i, b in enumerate(self.blocks):
            S = op.stress(i, u)
            dev = S - (S.diagonal(dim1=1, dim2=2).sum(-1) / 3)[:, None, None] * t.eye(3, dtype=S.dtype, device=S.device)
            out.append(dict(S=S, MISES=t.sqrt(1.5 * (dev * dev).sum((1, 2))), PEEQ=t.zeros(b.ne, dtype=S.dtype, device=S.device)))
        return out

    def geometric_ops(self, stresses, storage=L.KE_FULL, dtype=L.F64):
        """Model.geometric_ops for solids (gfem/model.py:541-553): [(dofs, KG)] per block, the
        initial-stress stiffness of each element's stress."""
        op = self.operator()
        return [(b.dofs(), op.geo_ke(i, b.etype, st["S"], storage, dtype))
                for i, (b, st) in enumerate(zip(self.blocks, stresses))]

    def tangent_jvp(self, u, states, v, fields=None):
        """K(u) v (gfem/model.py:488-490), matrix-free. Linear material: internal_force(v); J2
        blocks apply their consistent tangent at (u, states) on the fly."""
        op = self.operator(fields)
        uu = torch().zeros(self.ndof, dtype=torch().float64, device=device()) if u is None else u
        if self.nonlinear:
            op.set_state(uu, self.init_state() if states is None else states)
        set_rest_ops(op, [(b.dofs(), b.evaluate(self.X_dev(), uu, force=False)[1]) for b in self.discrete])
        return op.matvec(v, matrix_free=True)

    def _attach(self, Kes):
        op = self.live_operator().set_assembled(False)
        nb = len(self.blocks)
        for i, k in enumerate(Kes[:nb]):
            op.set_ke(i, k if isinstance(k, KeBlock) else KeBlock.from_array(self.blocks[i].etype, k))
        set_rest_ops(op, [(b.dofs(), k) for b, k in zip(self.discrete, Kes[nb:])])
        return op

    def matvec(self, Kes, v):
        """K v from element matrices per block (gfem/model.py:492-494), springs included."""
        return self._attach(Kes).matvec(v)

    def diagonal(self, Kes):
        return self._attach(Kes).diagonal()

    def lumped_mass(self):
        """Model.lumped_mass (gfem/model.py:501-510; HRZ lumping of Block.mass,
        gfem/model.py:120-124): (ndof,) device f64, each node's mass on its three dofs."""
        op = self.operator()
        dens = (C.c_double * len(self.blocks))(*[b.material.density for b in self.blocks])
        m = torch().empty(self.ndof, dtype=torch().float64, device=device())
        L.check(L.load().exfem_op_lumped_mass(op.h, dens, vptr(m), stream_ptr()))
        return m

    def total_mass(self):
        """gfem/model.py:515-516: sum of rho V over the blocks (= the sum of the lumped masses)."""
        return float(self.lumped_mass().sum()) / 3.0


__all__ = ["Model", "Block", "KeBlock", "EbeOperator", "ops_matvec", "ops_diagonal",
           "ops_assemble", "gather_plan", "gather_matvec"]
