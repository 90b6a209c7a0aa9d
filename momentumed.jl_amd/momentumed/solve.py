"""EDsolve: the user-facing solve (MomentumED/src/MomentumED.jl:217-480), device map path."""
from __future__ import annotations

import time
import warnings

import numpy as np

from . import device as dev
from .core import HilbertSubspace, MBOperator, MBS64Vector, Scatter, isupper

__all__ = ["EDsolve"]

_DEVICES = ("cpu", "gpu", "rocm", "cuda", "oneapi", "metal", "multi_cuda")


def EDsolve(subspace: HilbertSubspace, hamiltonian, *more_scats, N: int, method: str = "sparse",
            device: str = "gpu", showtime: bool = False, ishermitian: bool = True,
            element_type=float, index_type=np.int32, min_sparse_dim: int = 100,
            max_dense_dim: int = 200, map_warning_dim: int = 20000, method_info: bool = True,
            **krylovkit_kwargs):
    """Lowest N eigenpairs of `hamiltonian` in `subspace`.

    Same call shape and return value as the reference: ``energies, vectors = EDsolve(subspace,
    H; N, method, device, ...)`` or ``EDsolve(subspace, scat_list1, scat_list2, ...; ...)``, with
    energies ascending (float64[N]) and vectors a list of MBS64Vector (MomentumED.jl:217-480).
    Both Krylov methods run on the device: method="sparse" (the default, as in the reference)
    builds the stored matrix on the device and solves with it (MomentumED.jl:259-333);
    method="map" applies the matrix-free map (MomentumED.jl:375-480). KrylovKit keywords
    krylovdim, tol, maxiter, vec0 are honoured, plus `seed` for the start vector and, for
    method="map", `blocksize` (block Lanczos with that many vectors per step; 1, the default,
    is the single-vector iteration KrylovKit runs). The dense
    method (host LAPACK) is not part of this package; the reference's automatic switch from
    :sparse to :dense below min_sparse_dim is therefore not taken: small subspaces stay on
    the stored-matrix path, which gives the same eigenpairs."""
    if N < 1:
        raise AssertionError("N should be at least 1.")
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
    if method == "dense":
        raise NotImplementedError("method='dense' (host LAPACK) is not part of this package; "
                                  "use method='sparse' or 'map' on the device")
    if ishermitian and not isupper(hamiltonian):
        raise AssertionError("Use upper_hermitian form of Hamiltonian operator when "
                             "ishermitian = true.")
    if length_limit := np.iinfo(index_type).max if index_type is not None else None:
        if len(subspace) > length_limit:
            raise AssertionError(f"Hilbert space dimension {len(subspace)} exceeds limit of "
                                 f"index_type={index_type}.")
    dim = len(subspace)
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
        raise ValueError(f"device=:{device} is not available in this build (use 'rocm').")
    dev.throw_gpu_unavailable("rocm")

    t0 = time.perf_counter()
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
    energies = np.asarray(vals[:N])
    vectors = [MBS64Vector(vecs[i], subspace) for i in range(N)]
    return energies, vectors
