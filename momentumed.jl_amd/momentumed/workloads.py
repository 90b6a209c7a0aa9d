"""Synthetic benchmark systems: spinless (or two-component) fermions on a k-mesh with a
Coulomb-like interaction V(q) = 2*pi/|q| * exp(-c |q|^2) * (1 + a cos 6 theta_q), or its
screened form with 1/|q| -> 1/sqrt(|q|^2 + lambda^-2).

Geometry: a square torus of area 2*pi*Nk, so the momentum step is 2*pi/L with
L = sqrt(2*pi*Nk); the transfer q = k_f1 - k_i1 is taken at its minimum image on the
n1 x n2 mesh, and q = 0 is dropped. The interaction is density-density in the component
index. Perturbations are seeded and keep the interaction Hermitian and symmetric under vertex
exchange, so the term lists stay those of a Hermitian two-body operator.

Workloads (sizes computed by enumeration):
  small   3x5 mesh, N=5, H_one = cos(2 pi k1) + cos(2 pi k2), V with c=1/4 plus a complex
          Gaussian perturbation of amplitude 0.01 (seed 1234); 15 sectors of 200/201 states.
  medium  3x9 mesh, N=9, flat band, V with c=1/4; sector (0,0) has 173,613 states.
  large   6x6 mesh, N=12, V with c=1/2, a=0.1, then a phase e^{i phi}, phi ~ U(-pi,pi)*1e-3
          (seed 4321), on every non-diagonal upper term; sector (0,0) has 34,770,492 states.
  batched two conserved components on a 3x4 mesh each (24 orbitals), V with c=1/4 within a
          component and r*V between components; fillings (N1,N2) in (2,3), (2,4), (3,3),
          (3,4), (4,4), all 12 momentum sectors of each (1,210 ... 20,419 states), and 64
          ratios r ~ U(0.5, 2.0) (numpy default_rng seed 99): 3,840 solves, lowest 4 each.
          The 4x6-mesh, N1=N2=4 system has 4.7 million states per sector, not the
          1.5k-20k range this sweep is meant to cover, so the mesh and fillings are chosen to
          give that range.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .core import MBOperator, Scatter
from .preparation import EDPara, ED_scatterlist_onebody, ED_scatterlist_twobody

__all__ = ["mesh", "CoulombGaussian", "Workload", "small", "medium", "large", "get",
           "BATCH_FILLINGS", "batched_para", "batched_ratios"]

_M64 = (1 << 64) - 1


def _splitmix(x: int) -> int:
    """SplitMix64 finaliser (Steele, Lea & Flood 2014) on a 64-bit counter."""
    z = (x + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def _gauss2(seed: int, key: tuple) -> complex:
    """A complex number with independent N(0,1) parts, a pure function of (seed, key)."""
    h = seed & _M64
    for v in key:
        h = _splitmix(h ^ (int(v) & _M64))
    u1 = ((_splitmix(h) >> 11) + 1) * 2.0 ** -53
    u2 = (_splitmix(h ^ 0x5851F42D4C957F2D) >> 11) * 2.0 ** -53
    r = math.sqrt(-2.0 * math.log(u1))
    return complex(r * math.cos(2 * math.pi * u2), r * math.sin(2 * math.pi * u2))


def mesh(n1: int, n2: int) -> np.ndarray:
    """2 x (n1*n2) integer k_list, first coordinate fastest (the reference notebooks' order)."""
    return np.array([[i % n1 for i in range(n1 * n2)], [i // n1 for i in range(n1 * n2)]],
                    dtype=np.int64)


class CoulombGaussian:
    """Index-signature H_two(kf1, kf2, ki2, ki1, cf1, cf2, ci2, ci1) (1-based indices)."""

    def __init__(self, k_list: np.ndarray, dims: tuple, c: float = 0.25, aniso: float = 0.0,
                 perturb: float = 0.0, seed: int = 0, inter_ratio: float = 1.0,
                 screening: float | None = None):
        self.k = np.asarray(k_list, dtype=np.int64)
        self.dims = dims
        self.Nk = self.k.shape[1]
        self.step = 2 * math.pi / math.sqrt(2 * math.pi * self.Nk)
        self.c = c
        self.aniso = aniso
        self.perturb = perturb
        self.seed = seed
        self.inter_ratio = inter_ratio
        self.screening = screening     # length lambda: 1/|q| -> 1/sqrt(|q|^2 + lambda^-2)
        self._cache: dict = {}

    def vq(self, kf: int, ki: int) -> float:
        key = (kf, ki)
        v = self._cache.get(key)
        if v is None:
            q = []
            for d, n in enumerate(self.dims):
                dq = int(self.k[d, kf - 1] - self.k[d, ki - 1]) % n
                if dq > n // 2:
                    dq -= n
                q.append(dq * self.step)
            qq = math.hypot(q[0], q[1])
            if qq == 0.0:
                v = 0.0
            else:
                den = qq if self.screening is None else math.sqrt(qq * qq + self.screening ** -2)
                v = 2 * math.pi / den * math.exp(-self.c * qq * qq)
                if self.aniso:
                    v *= 1.0 + self.aniso * math.cos(6.0 * math.atan2(q[1], q[0]))
            self._cache[key] = v
        return v

    def _delta(self, t: tuple) -> complex:
        kf1, kf2, ki2, ki1, cf1, cf2, ci2, ci1 = t
        ex = (kf2, kf1, ki1, ki2, cf2, cf1, ci1, ci2)
        he = (ki1, ki2, kf2, kf1, ci1, ci2, cf2, cf1)
        hex_ = (ki2, ki1, kf1, kf2, ci2, ci1, cf1, cf2)
        g = (_gauss2(self.seed, t) + _gauss2(self.seed, ex)
             + _gauss2(self.seed, he).conjugate() + _gauss2(self.seed, hex_).conjugate())
        return 0.5 * g

    def __call__(self, kf1, kf2, ki2, ki1, cf1, cf2, ci2, ci1):
        v = 0j
        if cf1 == ci1 and cf2 == ci2:
            v = complex(self.vq(kf1, ki1))
            if cf1 != cf2:
                v *= self.inter_ratio
        # the perturbation keeps the particle number of every component
        if self.perturb and sorted((cf1, cf2)) == sorted((ci1, ci2)):
            v += self.perturb * self._delta((kf1, kf2, ki2, ki1, cf1, cf2, ci2, ci1))
        return v


@dataclass
class Workload:
    name: str
    para: EDPara
    N_each: tuple
    sector: tuple              # target total momentum (for the single-sector workloads)
    nev: int
    krylovdim: int = 30
    tol: float = 1e-12
    phase_disorder: float = 0.0
    phase_seed: int = 0
    notes: str = ""
    _op: MBOperator | None = field(default=None, repr=False)

    def operator(self) -> MBOperator:
        """The upper-Hermitian MBOperator of the workload (built once)."""
        if self._op is None:
            lists = [ED_scatterlist_twobody(self.para)]
            if self.para._H_one_given:
                lists.insert(0, ED_scatterlist_onebody(self.para))
            op = MBOperator(*lists, upper_hermitian=True)
            if self.phase_disorder:
                new = []
                for idx, s in enumerate(op.scats):
                    if s.in_ != s.out:
                        g = _splitmix(_splitmix(self.phase_seed) ^ idx)
                        u = (g >> 11) * 2.0 ** -53
                        phi = (2 * u - 1) * math.pi * self.phase_disorder
                        a = s.Amp * complex(math.cos(phi), math.sin(phi))
                        s = Scatter(None, bits=s.bits, _raw=(a, s.out, s.in_))
                    new.append(s)
                op.scats = new
            self._op = op
        return self._op


def small() -> Workload:
    k = mesh(3, 5)

    def h1(kc, cf, ci):
        return complex(math.cos(2 * math.pi * kc[0]) + math.cos(2 * math.pi * kc[1])) if cf == ci else 0j
    para = EDPara(k_list=k, Gk=(3, 5), H_one=h1,
                  H_two=CoulombGaussian(k, (3, 5), c=0.25, perturb=0.01, seed=1234),
                  two_momentum_coordinate=False)
    return Workload("small_3x5_N5", para, (5,), (0, 0), nev=8, krylovdim=30)


def medium() -> Workload:
    k = mesh(3, 9)
    para = EDPara(k_list=k, Gk=(3, 9), H_two=CoulombGaussian(k, (3, 9), c=0.25),
                  two_momentum_coordinate=False)
    return Workload("medium_3x9_N9", para, (9,), (0, 0), nev=6, krylovdim=40)


def large() -> Workload:
    k = mesh(6, 6)
    para = EDPara(k_list=k, Gk=(6, 6),
                  H_two=CoulombGaussian(k, (6, 6), c=0.5, aniso=0.1),
                  two_momentum_coordinate=False)
    return Workload("large_6x6_N12", para, (12,), (0, 0), nev=6, krylovdim=40,
                    phase_disorder=1e-3, phase_seed=4321)


BATCH_FILLINGS = ((2, 3), (2, 4), (3, 3), (3, 4), (4, 4))
BATCH_MESH = (3, 4)
BATCH_NEV = 4


def batched_para(ratio: float) -> EDPara:
    """The two-component 3x4 system of the batched sweep at inter-component ratio r."""
    k = mesh(*BATCH_MESH)
    return EDPara(k_list=k, Gk=BATCH_MESH, Nc_conserve=2,
                  H_two=CoulombGaussian(k, BATCH_MESH, c=0.25, inter_ratio=float(ratio)),
                  two_momentum_coordinate=False)


def literal_batched_para(ratio: float) -> EDPara:
    """The batched configuration as written: two conserved components on a 4x6 mesh each (48
    orbitals), intra-component V with c = 1/4 and inter-component ratio * V. With
    N_up = N_down = 4 its 24 sectors hold about 4.7 M states each."""
    k = mesh(4, 6)
    return EDPara(k_list=k, Gk=(4, 6), Nc_conserve=2,
                  H_two=CoulombGaussian(k, (4, 6), c=0.25, inter_ratio=float(ratio)),
                  two_momentum_coordinate=False)


def batched_ratios(n: int = 64, seed: int = 99) -> np.ndarray:
    return np.random.default_rng(seed).uniform(0.5, 2.0, n)


def get(name: str) -> Workload:
    return {"small": small, "medium": medium, "large": large}[name]()
