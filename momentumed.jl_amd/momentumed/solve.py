"""EDsolve: the user-facing solve (MomentumED/src/MomentumED.jl:217-480), device map path."""
from __future__ import annotations

import logging
import time
import warnings

import numpy as np

from . import device as dev
from .core import HilbertSubspace, MBOperator, MBS64Vector, Scatter, isupper

__all__ = ["EDsolve"]

_DEVICES = ("cpu", "gpu", "rocm", "cuda", "oneapi", "metal", "multi_cuda")
_log = logging.getLogger("momentumed")


def EDsolve(subspace: HilbertSubspace, hamiltonian, *more_scats, N: int, method: str = "sparse",
            device: str = "gpu", showtime: bool = False, ishermitian: bool = True,
            element_type=float, index_type=np.int32, min_sparse_dim: int = 100,
            max_dense_dim: int | None = None, map_warning_dim: int = 20000,
            method_info: bool = True, **krylovkit_kwargs):
    """Lowest N eigenpairs of `hamiltonian` in `subspace`.

    Same call shape and return value as the reference: ``energies, vectors = EDsolve(subspace,
    H; N, method, device, ...)`` or ``EDsolve(subspace, scat_list1, scat_list2, ...; ...)``, with
    energies ascending (float64[N]) and vectors a list of MBS64Vector (MomentumED.jl:217-480).
    All methods run on the device: method="sparse" (the default, as in the reference)
    builds the stored matrix on the device and solves with it (MomentumED.jl:259-333);
    method="map" applies the matrix-free map (MomentumED.jl:375-480); method="dense" writes
    the stored matrix as a dense array on the device and calls the GPU's dense Hermitian
    eigensolver (the reference runs LAPACK on the host, MomentumED.jl:340-369). As in the
    reference, "sparse" switches to "dense" below min_sparse_dim and "dense" to "sparse"
    above max_dense_dim (default 200 for an MBOperator, 500 for scatter lists,
    MomentumED.jl:221,379). element_type Float64 or Float32 (sparse and dense methods; the
    map ignores it, as the reference's does): the device computes in complex FP64 and
    Float32 results are rounded to Float32 / ComplexF32. KrylovKit keywords krylovdim, tol,
    maxiter, vec0 are honoured,
    plus `seed` for the start vector and, for method="map", `blocksize` (block Lanczos with
    that many vectors per step; 1, the default, is the single-vector iteration KrylovKit
    runs)."""
    if N < 1:
        raise AssertionError("N should be at least 1.")
    single = _single_precision(element_type)
    if method not in ("sparse", "dense", "map"):
        raise AssertionError(f"Unknown method: {method}. Use :sparse, :dense, :map.")
    if device not in _DEVICES:
        raise AssertionError(f"Unknown device: {device}.")
    if not isinstance(hamiltonian, MBOperator):
        # EDsolve(subspace, scat_lists...) form (MomentumED.jl:217-231): upper-Hermitian terms
        if method == "map":
            raise ValueError("Linear map methods are only supported when input Hamitonian is "
                             "MBOperator instead of Vector{Scatter}.")
        lists = [hamiltonian, *more_scats]
        if not all(isinstance(s, Scatter) for lst in lists for s in lst):
            raise TypeError("expected MBOperator or lists of Scatter")
        if not ishermitian:
            raise AssertionError("Current Hamiltonian matrix construction assumes it being "
                                 "Hermitian.")
        hamiltonian = MBOperator(*lists, upper_hermitian=True)
        if max_dense_dim is None:
            max_dense_dim = 500
    if max_dense_dim is None:
        max_dense_dim = 200
    if ishermitian and not isupper(hamiltonian):
        raise AssertionError("Use upper_hermitian form of Hamiltonian operator when "
                             "ishermitian = true.")
    if length_limit := np.iinfo(index_type).max if index_type is not None else None:
        if len(subspace) > length_limit:
            raise AssertionError(f"Hilbert space dimension {len(subspace)} exceeds limit of "
                                 f"index_type={index_type}.")
    dim = len(subspace)
    if method in ("sparse", "dense") and not ishermitian:
        # the MBOperator form forwards to the scatter-list form, which asserts it
        # (MomentumED.jl:247)
        raise AssertionError("Current Hamiltonian matrix construction assumes it being "
                             "Hermitian.")
    if method in ("sparse", "dense"):
        # MomentumED.jl:233-244
        if min_sparse_dim > max_dense_dim:
            if method_info:
                _log.info("Trying to set min_sparse_dim as %d, larger than max_dense_dim(=%d). "
                          "Reset to %d automatically.", min_sparse_dim, max_dense_dim,
                          max_dense_dim)
            min_sparse_dim = max_dense_dim
        if method == "sparse" and dim < min_sparse_dim:
            if method_info:
                _log.info("Hilbert space dimension < %d; switch to method=:dense "
                          "automatically.", min_sparse_dim)
            method = "dense"
        if method == "dense" and dim > max_dense_dim:
            if method_info:
                _log.info("Hilbert space dimension > %d; switch to method=:sparse "
                          "automatically.", max_dense_dim)
            method = "sparse"
    if method == "map" and dim < map_warning_dim and method_info:
        warnings.warn(f"Linear map may be slow for dim={dim}. Consider using :sparse method.")
    if device == "cpu":
        raise NotImplementedError("device='cpu': the CPU methods are the reference's; this "
                                  "package runs on the GPU (device='gpu' or 'rocm')")
    if device == "gpu":
        if dev.DEFAULT_GPU["value"] == "nogpu":
            dev.activate_gpu_device("rocm")
        device = dev.DEFAULT_GPU["value"]
        if device == "nogpu":
            raise ValueError("No GPU available. Please load a supported GPU extension and "
                             "verify availability.")
    if device != "rocm":
        raise ValueError(f"device=:{device} is not available in this package (use 'rocm').")
    dev.throw_gpu_unavailable("rocm")

    t0 = time.perf_counter()
    if method == "dense":
        if not ishermitian:
            raise NotImplementedError("the device dense method is the Hermitian one")
        H_gpu = dev.create_gpu_matrix(hamiltonian, subspace, "rocm")
        t1 = time.perf_counter()
        vals, vecs = dev.dense_matrix_solve(H_gpu, N)
        H_gpu.close()
        if showtime:
            print(f"  {t1 - t0:.6f} seconds (device matrix, {H_gpu.nnz} stored entries)")
            print(f"  {time.perf_counter() - t1:.6f} seconds (device dense eigensolver)")
        n = min(N, dim)
        return _as_requested(np.asarray(vals[:n]), [vecs[i] for i in range(n)], subspace,
                             single)
    if method == "sparse":
        H_gpu = dev.create_gpu_matrix(hamiltonian, subspace, "rocm")
        what = f"device matrix, {H_gpu.nnz} stored entries"
    else:
        H_gpu = dev.create_gpu_linearmap(hamiltonian, subspace, "rocm")
        what = "device linear map"
    t1 = time.perf_counter()
    solve = dev.krylov_matrix_solve if method == "sparse" else dev.krylov_map_solve
    vals, vecs, info = solve(H_gpu, N, ishermitian=ishermitian, **krylovkit_kwargs)
    t2 = time.perf_counter()
    if showtime:
        print(f"  {t1 - t0:.6f} seconds ({what})")
        print(f"  {t2 - t1:.6f} seconds (device Krylov solve, {info['matvecs']} matvecs)")
    if method == "sparse":
        H_gpu.close()   # release_gpu_memory after the solve (MomentumED.jl:329-331)
    if len(vals) < N:
        raise RuntimeError(f"Krylov method fails. Cannot find {N} eigenvectors.")
    if info["converged"] < min(N, dim):
        warnings.warn(f"Lanczos eigsolve stopped after {info['iterations']} restarts with "
                      f"{info['converged']} of {N} eigenpairs converged "
                      f"(max residual {info['max_residual']:.2e})")
    if method == "map":
        single = False   # the map ignores element_type (MomentumED.jl:389 builds LinearMap
        #                  with its Float64 default)
    return _as_requested(np.asarray(vals[:N]), vecs[:N], subspace, single)


def _single_precision(element_type) -> bool:
    """element_type as the reference accepts it (sparse_matrix.jl:72): Float64 or Float32
    (also given as numpy's float64 / float32 or Python's float)."""
    if element_type in (float, np.float64, "Float64"):
        return False
    if element_type in (np.float32, "Float32"):
        return True
    raise AssertionError(f"element_type={element_type}. Use element_type Float64, Float32.")


def _as_requested(energies, vecs, subspace, single):
    """The device computes in complex FP64; element_type=Float32 returns the results rounded
    to Float32 / ComplexF32, the types the reference's single-precision matrix gives."""
    if single:
        energies = energies.astype(np.float32)
        vecs = [np.asarray(v).astype(np.complex64) for v in vecs]
    return energies, [MBS64Vector(v, subspace) for v in vecs]
