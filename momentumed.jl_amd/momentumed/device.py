"""Device backend: the MI355X counterpart of the reference's GPU plugin contract.

Mirrors MomentumED/src/method/gpu_interface.jl (flags, activation, the abstract linear-map
hooks) and the per-vendor extension MomentumED/ext/MomentumEDCUDAExt.jl (device arrays,
create_gpu_linearmap, the callable map, krylov_map_solve, release_gpu_memory), all backed by
libedm.so through its C-ABI. The device symbol of this backend is "rocm"; "gpu" resolves to
DEFAULT_GPU as in the reference (MomentumED.jl:446-453).
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _lib
from ._lib import EdmSolveInfo, EdmSystem, check

__all__ = [
    "GPU_AVAILABLE", "DEFAULT_GPU", "activate_gpu_device", "release_gpu_memory", "Context",
    "get_context", "DeviceBasis", "DeviceOperator", "ROCmLinearMap", "create_gpu_linearmap",
    "krylov_map_solve", "system_struct", "DeviceCSR", "create_gpu_matrix", "krylov_matrix_solve",
]

# gpu_interface.jl:25-30,39
GPU_AVAILABLE = {"rocm": False}
DEFAULT_GPU = {"value": "nogpu"}
# gpu_interface.jl:36: when set, every device eigensolve reports the context's HBM use
GPU_MEMORY_MONITOR = {"value": False}


def _memory_report(ctx):
    if GPU_MEMORY_MONITOR["value"]:
        used, budget = ctx.memory()
        print(f"rocm context {ctx.device}: {used / 2**30:.3f} GiB in use"
              + (f" of a {budget / 2**30:.3f} GiB budget" if budget else ""))


def activate_gpu_device(device: str = "rocm", set_default_gpu: bool = True) -> bool:
    """gpu_interface.jl:41-60 + MomentumEDCUDAExt.jl:26-39 for the ROCm backend."""
    if device not in GPU_AVAILABLE:
        raise ValueError(f"Unknown GPU {device}. Valid options are: {list(GPU_AVAILABLE)}.")
    L = _lib.load()
    ok = bool(L.edm_device_available())
    GPU_AVAILABLE[device] = ok
    if ok and set_default_gpu:
        DEFAULT_GPU["value"] = device
    return ok


def throw_gpu_unavailable(device: str):
    """gpu_interface.jl:63-80."""
    if GPU_AVAILABLE.get(device):
        return
    try:
        activate_gpu_device(device, set_default_gpu=False)
    except RuntimeError:
        raise
    if not GPU_AVAILABLE.get(device):
        raise ValueError(f"device=:{device} requires functional GPU(s) of the type. "
                         "Load the required package(s) and verify GPU availability.")


# ------------------------------------------------------------------------- context
class Context:
    """One libedm context (one HIP stream) on one device."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self.L = _lib.load()
        h = ctypes.c_void_p()
        check(self.L.edm_ctx_create(device, hbm_budget_bytes, ctypes.byref(h)))
        self.h = h
        self.device = device
        self._fin = weakref.finalize(self, self.L.edm_ctx_destroy, h)

    def close(self):
        self._fin()

    def synchronize(self):
        check(self.L.edm_synchronize(self.h), self.h)

    @property
    def stream(self) -> int:
        s = ctypes.c_void_p()
        check(self.L.edm_ctx_stream(self.h, ctypes.byref(s)), self.h)
        return s.value or 0

    def memory(self):
        u, b = ctypes.c_uint64(), ctypes.c_uint64()
        check(self.L.edm_ctx_memory(self.h, ctypes.byref(u), ctypes.byref(b)), self.h)
        return u.value, b.value


_CONTEXTS: dict[int, Context] = {}


def get_context(device: int = 0) -> Context:
    ctx = _CONTEXTS.get(device)
    if ctx is None:
        ctx = Context(device)
        _CONTEXTS[device] = ctx
    return ctx


def release_gpu_memory(device: str = "rocm", level: int = 2):
    """gpu_interface.jl:87 / MomentumEDCUDAExt.jl:43-47: drop cached device state."""
    import gc
    if level >= 1:
        gc.collect()
    if level >= 2:
        for ctx in list(_CONTEXTS.values()):
            ctx.synchronize()


def system_struct(k_list, Gk, Nc_mix, Nc_conserve, N_each):
    """Build an edm_system (arrays are kept alive on the returned object)."""
    k_list = np.asarray(k_list, dtype=np.int64)
    if k_list.ndim == 1:
        k_list = k_list.reshape(1, -1)
    dim, nk = k_list.shape
    kl = np.ascontiguousarray(k_list.T.reshape(-1))          # column-major
    gk = np.ascontiguousarray(np.asarray(Gk, dtype=np.int64).reshape(-1))
    ne = np.ascontiguousarray(np.asarray(N_each, dtype=np.int64).reshape(-1))
    s = EdmSystem(dim, nk, Nc_mix, Nc_conserve,
                  kl.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                  gk.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                  ne.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    s._keep = (kl, gk, ne)
    return s


# ------------------------------------------------------------------------- basis
TUNE_AMP_LDS = 0x10000   # edm.h EDM_TUNE_AMP_LDS
TUNE_NARROW = 0x20000    # edm.h EDM_TUNE_NARROW
TUNE_SPLIT_SHIFT = 20    # edm.h EDM_TUNE_SPLIT_SHIFT
TUNE_CANDIDATES = 12     # edm.h EDM_TUNE_CANDIDATES


class DeviceBasis:
    """A sorted basis resident on the device (with rank tables when it is a complete sector)."""

    def __init__(self, ctx: Context, handle: ctypes.c_void_p):
        self.ctx = ctx
        self.h = handle
        D = ctypes.c_int64()
        check(ctx.L.edm_basis_size(handle, ctypes.byref(D)), ctx.h)
        self.D = D.value
        self._fin = weakref.finalize(self, ctx.L.edm_basis_destroy, handle)

    @classmethod
    def enumerate(cls, ctx: Context, system, sector: int) -> "DeviceBasis":
        h = ctypes.c_void_p()
        check(ctx.L.edm_basis_enumerate(ctx.h, ctypes.byref(system), sector, ctypes.byref(h)),
              ctx.h)
        return cls(ctx, h)

    @classmethod
    def enumerate_all(cls, ctx: Context, system) -> list:
        """Every sector's basis (momentum_list order, empty sectors included) from one
        device enumeration of the whole occupation space (edm_basis_enumerate_all)."""
        n = ctypes.c_int64()
        check(ctx.L.edm_num_sectors(ctypes.byref(system), ctypes.byref(n)), ctx.h)
        nsec = n.value
        hs = (ctypes.c_void_p * nsec)()
        check(ctx.L.edm_basis_enumerate_all(ctx.h, ctypes.byref(system), hs), ctx.h)
        return [cls(ctx, ctypes.c_void_p(h)) for h in hs]

    @classmethod
    def upload(cls, ctx: Context, states: np.ndarray, nbits: int, system=None,
               sector: int = -1) -> "DeviceBasis":
        states = np.ascontiguousarray(states, dtype=np.uint64)
        h = ctypes.c_void_p()
        sysp = ctypes.byref(system) if system is not None else None
        check(ctx.L.edm_basis_upload(ctx.h, states.ctypes.data if states.size else None,
                                     states.size, nbits, sysp, sector, ctypes.byref(h)), ctx.h)
        return cls(ctx, h)

    def download(self) -> np.ndarray:
        out = np.empty(max(self.D, 1), dtype=np.uint64)
        check(self.ctx.L.edm_basis_download(self.h, out.ctypes.data), self.ctx.h)
        return out[: self.D]

    @property
    def rank_mode(self) -> int:
        """0 binary search, 1 rank tables of a complete sector of the given system, 2 rank
        tables derived from the state list (used for an operator that keeps every target in
        the basis)."""
        m = ctypes.c_int32()
        check(self.ctx.L.edm_basis_rank_mode(self.h, ctypes.byref(m)), self.ctx.h)
        return m.value

    def drop_rank(self):
        """Release the rank tables: partner indices by binary search from then on."""
        check(self.ctx.L.edm_basis_drop_rank(self.h), self.ctx.h)

    @property
    def device_ptr(self) -> int:
        p = ctypes.c_void_p()
        check(self.ctx.L.edm_basis_device_ptr(self.h, ctypes.byref(p)), self.ctx.h)
        return p.value or 0

    def close(self):
        self._fin()


# ------------------------------------------------------------------------- operator
class DeviceOperator:
    """A term list on the device, expanded to the full list when upper-Hermitian
    (linear_map.jl:22-41) and bucketed by creation mask."""

    def __init__(self, ctx: Context, nbits: int, amp, out, inn, upper_hermitian: bool):
        amp = np.ascontiguousarray(amp, dtype=np.complex128)
        out = np.ascontiguousarray(out, dtype=np.uint64)
        inn = np.ascontiguousarray(inn, dtype=np.uint64)
        h = ctypes.c_void_p()
        S = amp.size
        check(ctx.L.edm_op_upload(ctx.h, nbits, amp.ctypes.data if S else None,
                                  out.ctypes.data if S else None, inn.ctypes.data if S else None,
                                  S, int(bool(upper_hermitian)), ctypes.byref(h)), ctx.h)
        self.ctx = ctx
        self.h = h
        self.nbits = nbits
        self.upper_hermitian = bool(upper_hermitian)
        n = ctypes.c_int64()
        check(ctx.L.edm_op_size(h, ctypes.byref(n)), ctx.h)
        self.S_full = n.value
        self._fin = weakref.finalize(self, ctx.L.edm_op_destroy, h)

    @classmethod
    def from_mboperator(cls, ctx: Context, op, nbits: int) -> "DeviceOperator":
        amp, out, inn = op.arrays()
        return cls(ctx, nbits, amp, out, inn, op.upper_hermitian)

    def full_list(self):
        n = self.S_full
        amp = np.empty(max(n, 1), dtype=np.complex128)
        out = np.empty(max(n, 1), dtype=np.uint64)
        inn = np.empty(max(n, 1), dtype=np.uint64)
        par = np.empty(max(n, 1), dtype=np.uint64)
        check(self.ctx.L.edm_op_full_list(self.h, amp.ctypes.data, out.ctypes.data,
                                          inn.ctypes.data, par.ctypes.data), self.ctx.h)
        return amp[:n], out[:n], inn[:n], par[:n]

    def close(self):
        self._fin()


# ------------------------------------------------------------------------- linear map
class ROCmLinearMap:
    """The device linear map (the CuLinearMap counterpart, MomentumEDCUDAExt.jl:55-63,119-139).

    ``A(y, x)`` computes y = H x for device tensors (torch, complex128, on the GPU) or host
    numpy arrays; ``A.adjoint()`` gives the H^dagger map."""

    def __init__(self, op: DeviceOperator, basis: DeviceBasis, space=None, adjoint: bool = False):
        self.op = op
        self.basis = basis
        self.space = space
        self.is_adjoint = adjoint

    @property
    def ctx(self) -> Context:
        return self.op.ctx

    def size(self):
        return (self.basis.D, self.basis.D)

    @property
    def dtype(self):
        return np.complex128

    def adjoint(self) -> "ROCmLinearMap":
        return ROCmLinearMap(self.op, self.basis, self.space, not self.is_adjoint)

    def _check(self, y, x):
        D = self.basis.D
        if len(x) != D or len(y) != D:
            raise ValueError("Dimension of Hamiltonian linear map mismatches vector length.")

    def tune(self):
        """Pick the matvec kernel for this operator and basis by timing (edm.h
        edm_matvec_tune); returns (chosen kernel code or 0, {candidate: ms per matvec}).
        The code is the rows per workgroup, | TUNE_AMP_LDS when the amplitudes are staged in
        LDS, | TUNE_NARROW when the low-half rank table is, | lanes << TUNE_SPLIT_SHIFT for the
        split-row kernel; candidates are named 256, 512, 1024, "lds256", "lds512", "lds1024",
        "split2", "split4", "split8", "split16", "rank512", "rank1024"."""
        code = ctypes.c_int32()
        ms = (ctypes.c_double * TUNE_CANDIDATES)()
        check(self.ctx.L.edm_matvec_tune(self.ctx.h, self.op.h, self.basis.h,
                                         ctypes.byref(code), ms), self.ctx.h)
        names = [256, 512, 1024, "lds256", "lds512", "lds1024", "split2", "split4", "split8",
                 "split16", "rank512", "rank1024"]
        return code.value, {n: ms[k] for k, n in enumerate(names) if ms[k] >= 0}

    def kernel(self) -> int:
        """The kernel code edm_matvec launches for this operator and basis now (edm.h
        edm_matvec_kernel; same encoding as tune(), 0 for the global-table kernels)."""
        code = ctypes.c_int32()
        check(self.ctx.L.edm_matvec_kernel(self.ctx.h, self.op.h, self.basis.h,
                                           ctypes.byref(code)), self.ctx.h)
        return code.value

    def matvec_device(self, y_ptr: int, x_ptr: int):
        """Stream-ordered y = H x on raw device pointers (no synchronisation)."""
        L = self.ctx.L
        f = L.edm_matvec_adjoint if self.is_adjoint else L.edm_matvec
        check(f(self.ctx.h, self.op.h, self.basis.h, x_ptr, y_ptr), self.ctx.h)

    def __call__(self, y, x=None):
        if x is None:
            x, y = y, None
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, dtype=np.complex128)
            if y is None:
                y = np.empty_like(x)
            self._check(y, x)
            if self.is_adjoint:
                import torch
                xd = torch.from_numpy(x).to(f"cuda:{self.ctx.device}")
                yd = torch.empty_like(xd)
                self(yd, xd)
                y[:] = yd.cpu().numpy()
                return y
            check(self.ctx.L.edm_matvec_host(self.ctx.h, self.op.h, self.basis.h,
                                             x.ctypes.data, y.ctypes.data), self.ctx.h)
            return y
        import torch
        if y is None:
            y = torch.empty_like(x)
        self._check(y, x)
        if x.dtype != torch.complex128 or y.dtype != torch.complex128:
            raise TypeError("device vectors must be complex128")
        if not (x.is_contiguous() and y.is_contiguous()):
            raise ValueError("device vectors must be contiguous")
        torch.cuda.current_stream(x.device).synchronize()
        self.matvec_device(y.data_ptr(), x.data_ptr())
        self.ctx.synchronize()   # as gpu_matvec! synchronizes after each launch (:218)
        return y


class LinearMap:
    """LinearMap(H::MBOperator, subspace) (linear_map.jl:18-42): the operator and subspace a
    map is made of. Here it only names them for create_gpu_linearmap(A, "rocm"); the full-list
    expansion and every product run on the device."""

    def __init__(self, op, space):
        self.op = op
        self.space = space

    def size(self):
        return (len(self.space), len(self.space))


class SparseHmltMatrix:
    """SparseHmltMatrix(subspace, scat_list; element_type, index_type)
    (sparse_matrix.jl:65-191): the Hermitian matrix of upper-Hermitian terms on a subspace.
    The reference assembles it on the host; here it names the terms and subspace for
    create_gpu_matrix(H, "rocm"), which assembles it on the device."""

    def __init__(self, subspace, scat_list, element_type=float, index_type=np.int32):
        from .core import MBOperator
        if len(subspace) > np.iinfo(index_type).max:
            raise AssertionError(f"Hilbert space dimension {len(subspace)} exceeds limit of "
                                 f"index_type={index_type}.")
        self.space = subspace
        self.op = MBOperator(list(scat_list), upper_hermitian=True)

    def size(self):
        return (len(self.space), len(self.space))


def create_gpu_linearmap(op, space=None, device: str = "rocm", ctx: Context | None = None):
    """gpu_interface.jl:85 / MomentumEDCUDAExt.jl:94-113: upload the operator (and the basis,
    unless the subspace is already on the device) and return the device map. Called as
    create_gpu_linearmap(op, space) or, as in the reference, create_gpu_linearmap(A, "rocm")
    with A a LinearMap."""
    if isinstance(op, LinearMap):
        if isinstance(space, str):
            device = space
        op, space = op.op, op.space
    if device != "rocm":
        raise ValueError(f"Unknown GPU {device}.")
    ctx = ctx or get_context(0)
    basis = device_basis_of(space, ctx)
    dop = DeviceOperator.from_mboperator(ctx, op, space.bits)
    return ROCmLinearMap(dop, basis, space)


def device_basis_of(space, ctx: Context) -> DeviceBasis:
    """The device copy of a HilbertSubspace (uploaded once and cached on the subspace)."""
    dev = getattr(space, "_device", None)
    if dev is not None and dev.ctx is ctx:
        return dev
    sysinfo = getattr(space, "sector_info", None)
    if sysinfo is not None:
        system, sector = sysinfo
        dev = DeviceBasis.upload(ctx, space.list, space.bits, system, sector)
    else:
        dev = DeviceBasis.upload(ctx, space.list, space.bits)
    space._device = dev
    return dev


def krylov_map_solve(H: ROCmLinearMap, N_eigen: int, *, ishermitian: bool = True, vec0=None,
                     krylovdim: int = 30, tol: float = 1e-12, maxiter: int = 100,
                     seed: int = 0x5EED, want_vectors: bool = True, blocksize: int = 1,
                     **_ignored):
    """gpu_interface.jl:92-94 / MomentumEDCUDAExt.jl:159-179 with KrylovKit's defaults
    (krylovdim 30, tol 1e-12, maxiter 100). Returns (vals, vecs, info); vecs is a list of
    host complex128 arrays (or None). blocksize > 1 runs block Lanczos (edm_eigsolve_block:
    `blocksize` vectors per step through the block matvec; krylovdim is then the basis size,
    rounded up to a multiple of the block). ishermitian=False runs the general eigensolver
    (edm_eigsolve_general, Krylov-Schur) as KrylovKit does for a non-Hermitian map: the
    N_eigen eigenvalues of smallest real part, complex, ascending real part."""
    if not ishermitian:
        if blocksize != 1:
            raise NotImplementedError("block Lanczos is for Hermitian maps")
        if H.is_adjoint:
            raise NotImplementedError("the general eigensolver applies the map itself, not "
                                      "its adjoint")
        return _general_solve(H.ctx, H.basis.D, N_eigen, vec0, krylovdim, tol, maxiter, seed,
                              want_vectors,
                              lambda *a: H.ctx.L.edm_eigsolve_general(H.ctx.h, H.op.h,
                                                                      H.basis.h, *a))
    if H.is_adjoint:
        H = H.adjoint()   # H is Hermitian: the adjoint map has the same spectrum
    D = H.basis.D
    N_eigen = min(N_eigen, D)
    vals = np.empty(N_eigen, dtype=np.float64)
    vecs = np.empty((N_eigen, D), dtype=np.complex128) if want_vectors else None
    v0 = None
    if vec0 is not None:
        v0 = np.ascontiguousarray(np.asarray(vec0, dtype=np.complex128).reshape(-1))
        if v0.size != D:
            raise ValueError("Dimension of Hamiltonian linear map mismatches vector length.")
    info = EdmSolveInfo()
    ctx = H.ctx
    v0p = v0.ctypes.data if v0 is not None else None
    vp = vecs.ctypes.data if vecs is not None else None
    if blocksize == 1:
        rc = ctx.L.edm_eigsolve(ctx.h, H.op.h, H.basis.h, N_eigen, krylovdim, float(tol),
                                maxiter, seed, v0p, vals.ctypes.data, vp, ctypes.byref(info))
    else:
        rc = ctx.L.edm_eigsolve_block(ctx.h, H.op.h, H.basis.h, N_eigen, int(blocksize),
                                      krylovdim, float(tol), maxiter, seed, v0p,
                                      vals.ctypes.data, vp, ctypes.byref(info))
    if rc not in (_lib.EDM_OK, _lib.EDM_ENOCONV):
        check(rc, ctx.h)
    out_info = {"iterations": info.iterations, "matvecs": info.matvecs,
                "converged": info.converged, "krylovdim": info.krylovdim,
                "max_residual": info.max_residual, "seconds": info.seconds}
    _memory_report(ctx)
    return vals, (list(vecs) if vecs is not None else None), out_info


def _general_solve(ctx, D, N_eigen, vec0, krylovdim, tol, maxiter, seed, want_vectors, call):
    """Shared driver of the general eigensolvers (edm_eigsolve_general and
    edm_csr_eigsolve_general): complex eigenvalues, ascending real part."""
    N_eigen = min(N_eigen, D)
    vals = np.empty(N_eigen, dtype=np.complex128)
    vecs = np.empty((N_eigen, D), dtype=np.complex128) if want_vectors else None
    v0 = None
    if vec0 is not None:
        v0 = np.ascontiguousarray(np.asarray(vec0, dtype=np.complex128).reshape(-1))
        if v0.size != D:
            raise ValueError("Dimension of Hamiltonian linear map mismatches vector length.")
    info = EdmSolveInfo()
    rc = call(N_eigen, krylovdim, float(tol), maxiter, seed,
              v0.ctypes.data if v0 is not None else None, vals.ctypes.data,
              vecs.ctypes.data if vecs is not None else None, ctypes.byref(info))
    if rc not in (_lib.EDM_OK, _lib.EDM_ENOCONV):
        check(rc, ctx.h)
    out_info = {"iterations": info.iterations, "matvecs": info.matvecs,
                "converged": info.converged, "krylovdim": info.krylovdim,
                "max_residual": info.max_residual, "seconds": info.seconds}
    _memory_report(ctx)
    return vals, (list(vecs) if vecs is not None else None), out_info


# ------------------------------------------------------------------------- stored matrix
class DeviceCSR:
    """The Hamiltonian of a subspace stored on the device as CSR (the CuSparseMatrixCSR
    counterpart, MomentumEDCUDAExt.jl:147-152), built on the device from the term tables."""

    def __init__(self, op: DeviceOperator, basis: DeviceBasis, space=None):
        ctx = op.ctx
        h = ctypes.c_void_p()
        check(ctx.L.edm_csr_build(ctx.h, op.h, basis.h, ctypes.byref(h)), ctx.h)
        self.ctx = ctx
        self.h = h
        self.space = space
        D, nnz = ctypes.c_int64(), ctypes.c_int64()
        check(ctx.L.edm_csr_size(h, ctypes.byref(D), ctypes.byref(nnz)), ctx.h)
        self.D, self.nnz = D.value, nnz.value
        self._fin = weakref.finalize(self, ctx.L.edm_csr_destroy, h)

    def size(self):
        return (self.D, self.D)

    def download(self):
        """(rowptr int64[D+1], cols int32[nnz], vals complex128[nnz]) on the host."""
        rp = np.empty(self.D + 1, dtype=np.int64)
        cols = np.empty(max(self.nnz, 1), dtype=np.int32)
        vals = np.empty(max(self.nnz, 1), dtype=np.complex128)
        check(self.ctx.L.edm_csr_download(self.h, rp.ctypes.data, cols.ctypes.data,
                                          vals.ctypes.data), self.ctx.h)
        return rp, cols[: self.nnz], vals[: self.nnz]

    def matvec_device(self, y_ptr: int, x_ptr: int):
        check(self.ctx.L.edm_csr_matvec(self.ctx.h, self.h, x_ptr, y_ptr), self.ctx.h)

    def __call__(self, y, x=None):
        """y = A x for torch complex128 device tensors or host numpy arrays."""
        import torch
        if x is None:
            x, y = y, None
        host = isinstance(x, np.ndarray)
        xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).to(
            f"cuda:{self.ctx.device}") if host else x
        if len(xd) != self.D:
            raise ValueError("Dimension of Hamiltonian matrix mismatches vector length.")
        yd = torch.empty_like(xd)
        torch.cuda.current_stream(xd.device).synchronize()
        self.matvec_device(yd.data_ptr(), xd.data_ptr())
        self.ctx.synchronize()
        if host:
            out = yd.cpu().numpy()
            if y is not None:
                y[:] = out
                return y
            return out
        if y is not None:
            y.copy_(yd)
            return y
        return yd

    def close(self):
        self._fin()


def create_gpu_matrix(op, space=None, device: str = "rocm", ctx: Context | None = None) -> DeviceCSR:
    """create_gpu_matrix(SparseHmltMatrix(subspace, scats), Val(device)) in one step
    (sparse_matrix.jl:65-191 + MomentumEDCUDAExt.jl:147-152): the matrix is built on the
    device from the operator's term list and the subspace's device basis. Called as
    create_gpu_matrix(op, space) or create_gpu_matrix(H, "rocm") with H a SparseHmltMatrix."""
    if isinstance(op, SparseHmltMatrix):
        if isinstance(space, str):
            device = space
        op, space = op.op, op.space
    if device != "rocm":
        raise ValueError(f"Unknown GPU {device}.")
    ctx = ctx or get_context(0)
    basis = device_basis_of(space, ctx)
    dop = DeviceOperator.from_mboperator(ctx, op, space.bits)
    return DeviceCSR(dop, basis, space)


def dense_matrix_solve(H: DeviceCSR, N_eigen: int, *, want_vectors: bool = True):
    """The :dense method (MomentumED.jl:340-369: eigen(Hermitian(Matrix(H)), 1:N)) on the
    device: the stored matrix is written as a dense D x D array by edm_csr_dense and
    diagonalised by the GPU's dense Hermitian eigensolver (torch.linalg.eigh, rocSOLVER).
    Returns (vals, vecs) with the N lowest eigenpairs."""
    import torch
    D = H.D
    N_eigen = min(N_eigen, D)
    ctx = H.ctx
    A = torch.empty((D, D), dtype=torch.complex128, device=f"cuda:{ctx.device}")
    torch.cuda.current_stream(A.device).synchronize()
    check(ctx.L.edm_csr_dense(ctx.h, H.h, A.data_ptr()), ctx.h)
    ctx.synchronize()
    w, V = torch.linalg.eigh(A)
    vals = w[:N_eigen].cpu().numpy()
    vecs = [V[:, i].cpu().numpy() for i in range(N_eigen)] if want_vectors else None
    del A, w, V
    return vals, vecs


def krylov_matrix_solve(H: DeviceCSR, N_eigen: int, *, ishermitian: bool = True, vec0=None,
                        krylovdim: int = 30, tol: float = 1e-12, maxiter: int = 100,
                        seed: int = 0x5EED, want_vectors: bool = True, blocksize: int = 1,
                        **_ignored):
    """krylov_matrix_solve (MomentumED.jl:316-333) on the device matrix; the same Lanczos
    iteration and defaults as krylov_map_solve. Returns (vals, vecs, info)."""
    if blocksize != 1:
        raise NotImplementedError("block Lanczos is provided for the matrix-free map "
                                  "(method=\"map\")")
    if not ishermitian:
        return _general_solve(H.ctx, H.D, N_eigen, vec0, krylovdim, tol, maxiter, seed,
                              want_vectors,
                              lambda *a: H.ctx.L.edm_csr_eigsolve_general(H.ctx.h, H.h, *a))
    D = H.D
    N_eigen = min(N_eigen, D)
    vals = np.empty(N_eigen, dtype=np.float64)
    vecs = np.empty((N_eigen, D), dtype=np.complex128) if want_vectors else None
    v0 = None
    if vec0 is not None:
        v0 = np.ascontiguousarray(np.asarray(vec0, dtype=np.complex128).reshape(-1))
        if v0.size != D:
            raise ValueError("Dimension of Hamiltonian matrix mismatches vector length.")
    info = EdmSolveInfo()
    ctx = H.ctx
    rc = ctx.L.edm_csr_eigsolve(ctx.h, H.h, N_eigen, krylovdim, float(tol), maxiter, seed,
                                v0.ctypes.data if v0 is not None else None, vals.ctypes.data,
                                vecs.ctypes.data if vecs is not None else None, ctypes.byref(info))
    if rc not in (_lib.EDM_OK, _lib.EDM_ENOCONV):
        check(rc, ctx.h)
    out_info = {"iterations": info.iterations, "matvecs": info.matvecs,
                "converged": info.converged, "krylovdim": info.krylovdim,
                "max_residual": info.max_residual, "seconds": info.seconds}
    _memory_report(ctx)
    return vals, (list(vecs) if vecs is not None else None), out_info
