"""Many eigensolves advanced together on one device (libedm's batch engine).

The reference runs one krylov_map_solve at a time (MomentumED/src/MomentumED.jl:375-480,
method/linear_map.jl:179-191); sweeps over sectors and parameter points are host loops of
EDsolve. A :class:`BatchSolver` holds many independent problems, each exactly the solve
:func:`momentumed.device.krylov_map_solve` would run for it (same arguments, bitwise the same
result), and moves all running problems one Lanczos step per launch of each kernel. Problems
can be added and collected between steps, which is what the serving loop
(:mod:`momentumed.serve`) builds on.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import EdmSolveInfo, check
from .device import Context, ROCmLinearMap, get_context

__all__ = ["BatchSolver", "BatchResult", "problem_bytes", "EDsolve_batch"]

RUNNING, CONVERGED, NOT_CONVERGED = 0, 1, 2


def problem_bytes(D: int, nev: int, krylovdim: int = 30) -> int:
    """Device bytes one batched problem of dimension D allocates (its Krylov workspace)."""
    L = _lib.load()
    b = ctypes.c_uint64()
    check(L.edm_batch_problem_bytes(int(D), int(nev), int(krylovdim), ctypes.byref(b)))
    return b.value


def _info_dict(info: EdmSolveInfo) -> dict:
    return {"iterations": info.iterations, "matvecs": info.matvecs,
            "converged": info.converged, "krylovdim": info.krylovdim,
            "max_residual": info.max_residual, "seconds": info.seconds}


@dataclass
class BatchResult:
    values: np.ndarray
    vectors: list | None
    info: dict
    converged: bool
    tag: object = None


class BatchSolver:
    """A set of Hermitian lowest-eigenpair problems solved together on one context."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or get_context(0)
        self.L = self.ctx.L
        h = ctypes.c_void_p()
        check(self.L.edm_batch_create(self.ctx.h, ctypes.byref(h)), self.ctx.h)
        self.h = h
        self._maps: dict[int, tuple] = {}    # id -> (map, nev, want_vectors, tag)
        self._closed = False

    # ------------------------------------------------------------------ lifecycle
    def close(self):
        if not self._closed:
            self.L.edm_batch_destroy(self.h)
            self._maps.clear()
            self._closed = True

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------ problems
    def add(self, H: ROCmLinearMap, N_eigen: int, *, krylovdim: int = 30, tol: float = 1e-12,
            maxiter: int = 100, seed: int = 0x5EED, vec0=None, want_vectors: bool = False,
            tag=None) -> int:
        """Admit one problem (arguments as krylov_map_solve). Raises EDMError with code
        EDM_ENOMEM when the context's HBM budget cannot hold its workspace."""
        if H.is_adjoint:
            H = H.adjoint()
        D = H.basis.D
        N_eigen = min(int(N_eigen), D)
        v0 = None
        if vec0 is not None:
            v0 = np.ascontiguousarray(np.asarray(vec0, dtype=np.complex128).reshape(-1))
            if v0.size != D:
                raise ValueError("Dimension of Hamiltonian linear map mismatches vector length.")
        pid = ctypes.c_int64()
        check(self.L.edm_batch_add(self.h, H.op.h, H.basis.h, N_eigen, int(krylovdim),
                                   float(tol), int(maxiter), int(seed) & ((1 << 64) - 1),
                                   v0.ctypes.data if v0 is not None else None,
                                   int(bool(want_vectors)), ctypes.byref(pid)), self.ctx.h)
        self._maps[pid.value] = (H, N_eigen, bool(want_vectors), tag)
        return pid.value

    def step(self, max_steps: int = 1) -> tuple[int, int]:
        """Advance every running problem by up to max_steps Lanczos steps (stopping after a
        step in which any problem finished). Returns (running, finished during the call)."""
        r, f = ctypes.c_int32(), ctypes.c_int32()
        check(self.L.edm_batch_step(self.h, int(max_steps), ctypes.byref(r), ctypes.byref(f)),
              self.ctx.h)
        return r.value, f.value

    def finished(self) -> list[int]:
        n = ctypes.c_int32()
        check(self.L.edm_batch_finished(self.h, None, 0, ctypes.byref(n)), self.ctx.h)
        if n.value == 0:
            return []
        ids = (ctypes.c_int64 * n.value)()
        check(self.L.edm_batch_finished(self.h, ids, n.value, ctypes.byref(n)), self.ctx.h)
        return list(ids)

    def status(self, pid: int) -> tuple[int, dict]:
        st, info = ctypes.c_int32(), EdmSolveInfo()
        check(self.L.edm_batch_status(self.h, int(pid), ctypes.byref(st), ctypes.byref(info)),
              self.ctx.h)
        return st.value, _info_dict(info)

    def result(self, pid: int) -> BatchResult:
        """Collect a finished problem (releases its device workspace)."""
        H, nev, want, tag = self._maps[pid]
        D = H.basis.D
        vals = np.empty(nev, dtype=np.float64)
        vecs = np.empty((nev, D), dtype=np.complex128) if want else None
        info = EdmSolveInfo()
        rc = self.L.edm_batch_result(self.h, int(pid), vals.ctypes.data,
                                     vecs.ctypes.data if vecs is not None else None,
                                     ctypes.byref(info))
        if rc not in (_lib.EDM_OK, _lib.EDM_ENOCONV):
            check(rc, self.ctx.h)
        del self._maps[pid]
        return BatchResult(vals, list(vecs) if vecs is not None else None, _info_dict(info),
                           rc == _lib.EDM_OK, tag)

    def remove(self, pid: int):
        check(self.L.edm_batch_remove(self.h, int(pid)), self.ctx.h)
        self._maps.pop(pid, None)

    def set_quantum(self, rows: int):
        """Deficit-round-robin step pacing (0: every problem every step); see edm.h."""
        check(self.L.edm_batch_set_quantum(self.h, int(rows)), self.ctx.h)

    def set_weight(self, pid: int, weight: float):
        """Round-robin weight of one problem (credit weight * quantum rows per step)."""
        check(self.L.edm_batch_set_weight(self.h, int(pid), float(weight)), self.ctx.h)

    def count(self) -> tuple[int, int]:
        t, r = ctypes.c_int32(), ctypes.c_int32()
        check(self.L.edm_batch_count(self.h, ctypes.byref(t), ctypes.byref(r)), self.ctx.h)
        return t.value, r.value

    def run(self, max_steps_per_call: int = 64) -> dict[int, BatchResult]:
        """Step until no problem is running; collect and return every result by id."""
        out = {}
        while True:
            running, _ = self.step(max_steps_per_call)
            for pid in self.finished():
                out[pid] = self.result(pid)
            if running == 0:
                break
        return out


def EDsolve_batch(problems, N: int = 6, *, krylovdim: int = 30, tol: float = 1e-12,
                  maxiter: int = 100, seed: int = 0x5EED, want_vectors: bool = True,
                  ctx: Context | None = None):
    """Solve many (HilbertSubspace, MBOperator) pairs together; the batched counterpart of a
    loop of EDsolve(subspace, H; N, method=:map, device=:rocm) calls (MomentumED.jl:375-480).

    Returns a list of (energies, vectors) in input order, vectors as MBS64Vector (or None
    when want_vectors is False). Bases and operators shared between problems are uploaded
    once."""
    from .core import MBS64Vector
    from .device import DeviceOperator, device_basis_of
    ctx = ctx or get_context(0)
    op_cache: dict[int, DeviceOperator] = {}
    maps = []
    for space, H in problems:
        dop = op_cache.get(id(H))
        if dop is None:
            dop = DeviceOperator.from_mboperator(ctx, H, space.bits)
            op_cache[id(H)] = dop
        maps.append(ROCmLinearMap(dop, device_basis_of(space, ctx), space))
    with BatchSolver(ctx) as bs:
        ids = [bs.add(m, N, krylovdim=krylovdim, tol=tol, maxiter=maxiter, seed=seed,
                      want_vectors=want_vectors) for m in maps]
        res = bs.run()
    out = []
    for pid, (space, _) in zip(ids, problems):
        r = res[pid]
        if not r.converged:
            raise RuntimeError("Krylov method fails. Cannot find N eigenvectors.")
        vecs = [MBS64Vector(v, space) for v in r.vectors] if r.vectors is not None else None
        out.append((r.values, vecs))
    return out
