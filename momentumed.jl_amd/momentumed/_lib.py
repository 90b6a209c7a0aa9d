"""ctypes binding of libedm.so (include/edm.h). Fails loudly when the library is missing:
there is no CPU fallback for the device path."""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libedm.so")

EDM_OK = 0
EDM_EINVAL = -1
EDM_ENODEV = -2
EDM_ENOMEM = -3
EDM_EHIP = -4
EDM_ENOCONV = -5
EDM_EDIM = -6

c_i32, c_i64, c_u64, c_dbl, c_vp = (ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64,
                                    ctypes.c_double, ctypes.c_void_p)
P = ctypes.POINTER


class EdmSystem(ctypes.Structure):
    _fields_ = [("dim", c_i32), ("nk", c_i32), ("nc_mix", c_i32), ("nc_conserve", c_i32),
                ("k_list", P(c_i64)), ("gk", P(c_i64)), ("n_each", P(c_i64))]


class EdmSolveInfo(ctypes.Structure):
    _fields_ = [("iterations", c_i32), ("matvecs", c_i32), ("converged", c_i32),
                ("krylovdim", c_i32), ("max_residual", c_dbl), ("seconds", c_dbl)]


# (name, restype, argtypes) for every entry point of include/edm.h
_SIGS = [
    ("edm_device_available", c_i32, []),
    ("edm_ctx_create", c_i32, [c_i32, c_u64, P(c_vp)]),
    ("edm_ctx_destroy", c_i32, [c_vp]),
    ("edm_last_error", ctypes.c_char_p, [c_vp]),
    ("edm_synchronize", c_i32, [c_vp]),
    ("edm_ctx_stream", c_i32, [c_vp, P(c_vp)]),
    ("edm_ctx_memory", c_i32, [c_vp, P(c_u64), P(c_u64)]),
    ("edm_num_sectors", c_i32, [P(EdmSystem), P(c_i64)]),
    ("edm_sector_sizes", c_i32, [c_vp, P(EdmSystem), c_vp]),
    ("edm_basis_enumerate", c_i32, [c_vp, P(EdmSystem), c_i64, P(c_vp)]),
    ("edm_basis_enumerate_all", c_i32, [c_vp, P(EdmSystem), c_vp]),
    ("edm_basis_upload", c_i32, [c_vp, c_vp, c_i64, c_i32, P(EdmSystem), c_i64, P(c_vp)]),
    ("edm_basis_size", c_i32, [c_vp, P(c_i64)]),
    ("edm_basis_download", c_i32, [c_vp, c_vp]),
    ("edm_basis_device_ptr", c_i32, [c_vp, P(c_vp)]),
    ("edm_basis_rank_mode", c_i32, [c_vp, P(c_i32)]),
    ("edm_basis_drop_rank", c_i32, [c_vp]),
    ("edm_basis_destroy", c_i32, [c_vp]),
    ("edm_op_upload", c_i32, [c_vp, c_i32, c_vp, c_vp, c_vp, c_i64, c_i32, P(c_vp)]),
    ("edm_op_size", c_i32, [c_vp, P(c_i64)]),
    ("edm_op_full_list", c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("edm_op_destroy", c_i32, [c_vp]),
    ("edm_matvec", c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("edm_matvec_adjoint", c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("edm_matvec_block", c_i32, [c_vp, c_vp, c_vp, c_i32, c_vp, c_vp, c_vp]),
    ("edm_matvec_host", c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("edm_connectivity_counts", c_i32, [c_vp, c_vp, c_vp, c_i64, c_i64, c_vp]),
    ("edm_connectivity", c_i32, [c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp]),
    ("edm_connectivity_totals", c_i32, [c_vp, c_vp, c_vp, P(c_i64), P(c_i64)]),
    ("edm_eigsolve", c_i32, [c_vp, c_vp, c_vp, c_i32, c_i32, c_dbl, c_i32, c_u64, c_vp, c_vp,
                             c_vp, P(EdmSolveInfo)]),
    ("edm_eigsolve_general", c_i32, [c_vp, c_vp, c_vp, c_i32, c_i32, c_dbl, c_i32, c_u64, c_vp,
                                     c_vp, c_vp, P(EdmSolveInfo)]),
    ("edm_eigsolve_block", c_i32, [c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_dbl, c_i32, c_u64,
                                   c_vp, c_vp, c_vp, P(EdmSolveInfo)]),
    ("edm_batch_create", c_i32, [c_vp, P(c_vp)]),
    ("edm_batch_destroy", c_i32, [c_vp]),
    ("edm_batch_add", c_i32, [c_vp, c_vp, c_vp, c_i32, c_i32, c_dbl, c_i32, c_u64, c_vp, c_i32,
                              P(c_i64)]),
    ("edm_batch_step", c_i32, [c_vp, c_i32, P(c_i32), P(c_i32)]),
    ("edm_batch_finished", c_i32, [c_vp, P(c_i64), c_i32, P(c_i32)]),
    ("edm_batch_status", c_i32, [c_vp, c_i64, P(c_i32), P(EdmSolveInfo)]),
    ("edm_batch_result", c_i32, [c_vp, c_i64, c_vp, c_vp, P(EdmSolveInfo)]),
    ("edm_batch_remove", c_i32, [c_vp, c_i64]),
    ("edm_batch_count", c_i32, [c_vp, P(c_i32), P(c_i32)]),
    ("edm_batch_set_quantum", c_i32, [c_vp, c_i64]),
    ("edm_batch_set_weight", c_i32, [c_vp, c_i64, ctypes.c_double]),
    ("edm_batch_problem_bytes", c_i32, [c_i64, c_i32, c_i32, P(c_u64)]),
    ("edm_matvec_tune", c_i32, [c_vp, c_vp, c_vp, P(c_i32), c_vp]),
    ("edm_matvec_kernel", c_i32, [c_vp, c_vp, c_vp, P(c_i32)]),
    ("edm_matvec_between", c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("edm_bracket2", c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("edm_bracket2_noalloc", c_i32, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("edm_pes_block_matrix", c_i32, [c_vp, c_vp, c_vp, c_i64, c_vp, c_u64, c_vp, c_vp, c_vp]),
    ("edm_rdm1", c_i32, [c_vp, c_vp, c_vp, c_i32, c_vp, c_vp]),
    ("edm_step_inner_prod", c_i32, [c_vp, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp]),
    ("edm_csr_build", c_i32, [c_vp, c_vp, c_vp, P(c_vp)]),
    ("edm_csr_size", c_i32, [c_vp, P(c_i64), P(c_i64)]),
    ("edm_csr_download", c_i32, [c_vp, c_vp, c_vp, c_vp]),
    ("edm_csr_matvec", c_i32, [c_vp, c_vp, c_vp, c_vp]),
    ("edm_csr_eigsolve", c_i32, [c_vp, c_vp, c_i32, c_i32, c_dbl, c_i32, c_u64, c_vp, c_vp, c_vp,
                                 P(EdmSolveInfo)]),
    ("edm_csr_eigsolve_general", c_i32, [c_vp, c_vp, c_i32, c_i32, c_dbl, c_i32, c_u64, c_vp,
                                         c_vp, c_vp, P(EdmSolveInfo)]),
    ("edm_csr_dense", c_i32, [c_vp, c_vp, c_vp]),
    ("edm_csr_destroy", c_i32, [c_vp]),
    ("edm_projected_eigh", c_i32, [c_i32, c_vp, c_vp, c_vp]),
    ("edm_projected_heigh", c_i32, [c_i32, c_vp, c_vp, c_vp]),
    ("edm_projected_schur", c_i32, [c_i32, c_vp, c_vp, c_vp]),
    ("edm_fill_random", c_i32, [c_vp, c_vp, c_i64, c_u64, c_i32]),
]

EXPORTED = [s[0] for s in _SIGS]

_lib = None


class EDMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[edm {code}] {msg}")
        self.code = code


def load(path: str | None = None) -> ctypes.CDLL:
    """Load libedm.so and bind every entry point. Raises if the library is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    # PyTorch-ROCm ships its own libamdhip64.so.7 and loads it by path; libedm.so asks for the
    # soname. Loading torch first makes libedm.so bind to the runtime torch already holds, so
    # a process that uses both sees one HIP runtime (two runtimes cannot share the device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(p):
        raise RuntimeError(f"libedm.so is not built ({p}); run `python -c \"import "
                           f"__graft_entry__ as g; g.build()\"` or `make -C "
                           f"momentumed.jl_amd/csrc`")
    L = ctypes.CDLL(p)
    for name, res, args in _SIGS:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = L
    return L


def check(rc: int, ctx=None):
    """Raise EDMError for a non-zero status, with the library's message."""
    if rc == EDM_OK:
        return
    L = load()
    msg = L.edm_last_error(ctx)
    msg = msg.decode() if msg else ""
    if rc == EDM_ENODEV:
        raise EDMError(rc, msg or "no usable GPU")
    if rc == EDM_EDIM:
        raise EDMError(rc, msg)
    raise EDMError(rc, msg)
