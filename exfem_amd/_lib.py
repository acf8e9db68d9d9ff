"""ctypes binding of libexfem.so (the C ABI declared in include/exfem.h).

The library is built in-tree by ``python -m exfem_amd.build`` (or ``__graft_entry__.build()``).
Importing this module never builds: a missing library raises ``ExfemLibraryMissing`` so the
device path fails loudly instead of falling back to anything else.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("EXFEM_LIB") or os.path.join(HERE, "_lib", "libexfem.so")

HEX8, HEX8_PLAIN, TET10 = 0, 1, 2
KE_FULL, KE_SYM = 0, 1
F64, F32 = 0, 1
PC_NONE, PC_JACOBI, PC_AMG, PC_CHEBY, PC_AMG_MIXED, PC_CHEBY_MIXED = 0, 1, 2, 3, 4, 5
OK, NOT_CONVERGED = 0, 1


class ExfemLibraryMissing(RuntimeError):
    pass


class ExfemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"exfem error {code}: {msg}")
        self.code = code


_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_d = C.c_double

# name -> (restype, argtypes); this list is the exported ABI (tests check it against the header)
SIGNATURES = {
    "exfem_last_error": (C.c_char_p, []),
    "exfem_abi_version": (_i32, []),
    "exfem_ke_bytes": (_i64, [_i32, _i64, _i32, _i32]),
    "exfem_ke_compute": (_i32, [_i32, _i64, _i64, _p, _p, _p, _d, _d, _i32, _i32, _p, _p]),
    "exfem_ke_pack": (_i32, [_i32, _i64, _p, _i32, _i32, _p, _p]),
    "exfem_ke_unpack": (_i32, [_i32, _i64, _p, _i32, _i32, _p, _p]),
    "exfem_op_create": (_i32, [C.POINTER(_p), _i64, _i32, C.POINTER(_i32), C.POINTER(_i64),
                               C.POINTER(_p)]),
    "exfem_op_destroy": (_i32, [_p]),
    "exfem_op_set_ke": (_i32, [_p, _i32, _p, _i32, _i32]),
    "exfem_op_set_bc": (_i32, [_p, _p]),
    "exfem_op_set_geometry": (_i32, [_p, _p, C.POINTER(_p), C.POINTER(_d), C.POINTER(_d)]),
    "exfem_op_matvec": (_i32, [_p, _p, _p, _i32, _i32, _i32, _p, _p]),
    "exfem_op_matvec_multi": (_i32, [_p, _i32, _p, _p, _i32, _i32, _p]),
    "exfem_op_multi_shape": (_i32, [_p, _i32, _i32, _p, _p, _p, _p, _p]),
    "exfem_op_set_multi_shape": (_i32, [_p, _i32, _i32, _i32]),
    "exfem_patch_build": (_i64, [_i64, _i32, _i64, _p, _i32, _i32, _i64, _p, _p, _p, _p, _p]),
    "exfem_op_set_assembled": (_i32, [_p, _i32]),
    "exfem_op_set_hyper": (_i32, [_p, _i32, _i32, _p]),
    "exfem_op_stress": (_i32, [_p, _i32, _p, _p, _p]),
    "exfem_op_geo_ke": (_i32, [_p, _i32, _p, _i32, _i32, _p, _p]),
    "exfem_op_set_fields": (_i32, [_p, _p, _p, _p, _p]),
    "exfem_op_diagonal": (_i32, [_p, _p, _p]),
    "exfem_op_set_rest": (_i32, [_p, _i32, _i64, _i32, _p, _p, _p]),
    "exfem_op_rest_assemble": (_i32, [_p, _i32, _p, _p, _p]),
    "exfem_spring_eval": (_i32, [_i32, _i64, _p, _p, _p, _p, _p, _p, _p, _p]),
    "exfem_contact_pairs": (_i32, [_i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "exfem_contact_select": (_i32, [_i64, _i32, _p, _p, _p, _p, _d, _d, _p, _p, _p, _p]),
    "exfem_wall_eval": (_i32, [_i64, _p, _p, _p, _p, _p, _d, _p, _p, _p, _p]),
    "exfem_op_profile": (_i32, [_p, _p, _p, _i32, _i32, _i32, _i32, C.POINTER(C.c_float),
                                C.POINTER(C.c_float), _p]),
    "exfem_op_lumped_mass": (_i32, [_p, _p, _p, _p]),
    "exfem_op_internal_force": (_i32, [_p, _p, _p, _p]),
    "exfem_op_assemble_bsr": (_i32, [_p, C.POINTER(_i64), _p, _p, _p, _i32, _p]),
    "exfem_amg_create": (_i32, [C.POINTER(_p), _i32, _i32]),
    "exfem_amg_destroy": (_i32, [_p]),
    "exfem_amg_set_level": (_i32, [_p, _i32, _i64, _i64, _d, _p, _p, _p, _p, _p, _p, _p, _p,
                                   _p, _p]),
    "exfem_amg_set_level_matrix": (_i32, [_p, _i32, _p, _p, _p]),
    "exfem_amg_set_coarse": (_i32, [_p, _i64, _p]),
    "exfem_amg_vcycle": (_i32, [_p, _p, _p, _p, _p, _i32, _p]),
    "exfem_amg_set_precision": (_i32, [_p, _i32]),
    "exfem_op_set_j2": (_i32, [_p, _i32, _d, _d, _d, _i32, _p, _p]),
    "exfem_op_set_mpc": (_i32, [_p, _i64, _p, _p, _p]),
    "exfem_op_mpc_apply": (_i32, [_p, _i32, _p, _p, _p]),
    "exfem_mpc_build": (_i32, [_i64, _i64, _i32, _p, _p, _p, _p, C.POINTER(_p)]),
    "exfem_mpc_plan_info": (_i32, [_p, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64),
                                   C.POINTER(_i64)]),
    "exfem_mpc_plan_get": (_i32, [_p, _p, _p, _p, _p]),
    "exfem_mpc_plan_free": (None, [_p]),
    "exfem_op_internal_force_state": (_i32, [_p, _p, C.POINTER(_p), C.POINTER(_p), _p, _p]),
    "exfem_op_ke_state": (_i32, [_p, _i32, _p, _p, _i32, _i32, _p, _p]),
    "exfem_op_set_state": (_i32, [_p, _p, C.POINTER(_p), _p]),
    "exfem_amg_vcycle_multi": (_i32, [_p, _p, _p, _i32, _p, _p, _i32, _i32, _p]),
    "exfem_amg_set_level_blocks": (_i32, [_p, _i32, _i32, _i64, _i32, _i32, _p, _p, _p]),
    "exfem_amg_smooth": (_i32, [_p, _p, _p, _i32, _p, _p, _i32, _i32, _i32, _p]),
    "exfem_csr_mv": (_i32, [_i64, _i64, _i32, _p, _p, _p, _p, _p, _i32, _p]),
    "exfem_pcg_create": (_i32, [C.POINTER(_p), _i64]),
    "exfem_pcg_destroy": (_i32, [_p]),
    "exfem_pcg_set_graph": (_i32, [_p, _i32]),
    "exfem_pcg_begin": (_i32, [_p, _p, _p, _i32, _i32, _p, _d, _p, _i32, _d, _i32, _p]),
    "exfem_pcg_advance": (_i32, [_p, _i32]),
    "exfem_pcg_query": (_i32, [_p, C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i32),
                               C.POINTER(_d)]),
    "exfem_pcg_finish": (_i32, [_p, _p, _p]),
    "exfem_pcg_solve": (_i32, [_p, _p, _p, _p, _i32, _i32, _p, _d, _p, _i32, _d, _i32,
                               C.POINTER(_i32), C.POINTER(_d), _p]),
    "exfem_pcg_solve_noalloc": (_i32, [_p, _p, _p, _i32, _i32, _p, _d, _p, _i32, _d, _i32,
                                       C.POINTER(_i32), C.POINTER(_d), _p]),
    "exfem_pcg_precond": (_i32, [_p, _p, _i32, _p, _d, _p, _i32, _p, _p, _i32, _p]),
    "exfem_pcg_lmax": (_i32, [_p, _p, _p, _i32, _i32, C.POINTER(_d), _p]),
    "exfem_block_pcg_solve": (_i32, [_p, _i32, _p, _p, _i32, _p, _d, _i32, C.POINTER(_i32),
                                     C.POINTER(_d), _p]),
    "exfem_block_pcg_solve_pc": (_i32, [_p, _i32, _p, _p, _i32, _i32, _p, _p, _d, _i32,
                                        C.POINTER(_i32), C.POINTER(_d), _p]),
    "exfem_sa_aggregate": (_i64, [_i64, _p, _p, _p]),
    "exfem_sa_fit_candidates": (_i32, [_i64, _i32, _i32, _p, _i64, _p, _d, _p, _p]),
    "exfem_sa_level_build": (_i32, [_i64, _i32, _p, _p, _p, _i32, _p, _d, _d, _i32, _p]),
    "exfem_sa_level_info": (_i32, [_p, _p, _p, _p]),
    "exfem_sa_level_get": (_i32, [_p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "exfem_sa_level_free": (None, [_p]),
    "exfem_sa_rho": (_i32, [_i64, _i32, _p, _p, _p, _p, _i32, _i32, _p]),
    "exfem_sa_dev_begin": (_i32, [_p, _p, _i32, _p, _p]),
    "exfem_sa_dev_begin_rbm": (_i32, [_p, _p, _p, _p, _p]),
    "exfem_sa_dev_level": (_i32, [_p, _d, _d, _p, _i32, _i32, _p, _p, _p]),
    "exfem_sa_dev_info": (_i32, [_p, _i32, _p, _p, _p, _p, _p, _p]),
    "exfem_sa_dev_export": (_i32, [_p, _i32, _i32, _p, _p, _p, _p]),
    "exfem_sa_dev_coarse": (_i32, [_p, _p, _p]),
    "exfem_sa_dev_to_amg": (_i32, [_p, _p, _i32, _p]),
    "exfem_sa_dev_destroy": (_i32, [_p]),
    "exfem_sa_dev_coarse_solve": (_i32, [_p, _p, _p, _p, _p]),
    "exfem_amg_level0_from_op": (_i32, [_p, _p, _p]),
    "exfem_amg_set_coarse_inverse": (_i32, [_p, _i64, _p]),
}

_lib = None


def load():
    """Load libexfem.so (raises ExfemLibraryMissing if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ExfemLibraryMissing(
            f"{LIB_PATH} not found: build it with `python -m exfem_amd.build`")
    # the HIP runtime must be the one torch already loaded (same SONAME); import torch first
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code):
    """Raise on a negative status; return the status otherwise (1 = not converged)."""
    if code < 0:
        msg = load().exfem_last_error()
        raise ExfemError(code, msg.decode() if msg else "")
    return code
