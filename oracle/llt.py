"""Lowest-Landau-level torus interaction — TEST INFRASTRUCTURE ONLY (see oracle.py header).

Restatement of the reference example's LandauInteraction
(examples/landau level torus/Landau level torus.jl): ReciprocalLattice (:35-60), BZ fold
(:121-141), V_Coulomb (:188-211), V_Haldane (:217-232), and the callable (:243-347) with the
magnetic-translation phases. Scattering within one Landau level n is restated (n_f = n_i for
both form factors): level_form_factor(n, n, ql) (:157-173) is exp(-|ql|^2/4) L_n(|ql|^2/2)
with the ordinary Laguerre polynomial L_n = laguerrel(n, 0, .) of ClassicalOrthogonalPolynomials
(absent here; computed by its three-term recurrence, pinned against the closed forms of
L_0..L_3); at n = 0 the factor is exp(-|ql|^2/4) exactly. Level-changing form factors are not
restated.

Used to reproduce numbers the reference's notebooks print: the 1734 upper two-body terms of the
4x6, Ne=8 triangular Coulomb system ("4. cuda support.ipynb"), the lowest energies of the
3x5, Ne=5 sector (0,0) ("1.FQH_Laughlin.ipynb"), the bilayer Halperin energies, term count
and many-body Chern numbers ("2.BilayerFQH_Halperin.ipynb"; orbital_overlap, :352-393), and
the n = 2 Landau-level charge-density-wave system ("1.CDW.ipynb": term count, sector sizes,
energies).
"""
from __future__ import annotations

import cmath
import math
from fractions import Fraction

W0 = 1.0
L_MAG = 1.0
GL_CUTOFF = 10.0


class ReciprocalLattice:
    """Landau level torus.jl:35-60 (G1 phase 0)."""

    def __init__(self, G2_ratio_G1, cos_angle):
        G2_ratio_G1 = Fraction(G2_ratio_G1)
        cos_angle = Fraction(cos_angle)
        sin_sq = 1.0 - float(cos_angle) ** 2
        assert sin_sq > 0
        sin_a = math.sqrt(sin_sq)
        g1 = math.sqrt(2 * math.pi / L_MAG ** 2 / abs(float(G2_ratio_G1)) / sin_a)
        self.G1 = complex(g1, 0.0)
        self.G2 = self.G1 * float(G2_ratio_G1) * complex(float(cos_angle), sin_a)
        self.r22 = G2_ratio_G1 ** 2                 # exact |G2|^2 / |G1|^2
        self.r12 = G2_ratio_G1 * cos_angle          # exact G1.G2 / |G1|^2

    @classmethod
    def named(cls, s: str):
        return {"square": cls(1, 0), "triangular": cls(1, Fraction(-1, 2))}[s]


def _rem1(x: Fraction) -> Fraction:
    return x - math.floor(x)


def bz(k, lat: ReciprocalLattice):
    """Landau level torus.jl:121-141 with exact rationals."""
    k1, k2 = _rem1(k[0]), _rem1(k[1])
    kdG1 = k1 + k2 * lat.r12
    kdG2 = k1 * lat.r12 + k2 * lat.r22
    d00 = k1 ** 2 + 2 * k1 * k2 * lat.r12 + k2 ** 2 * lat.r22
    d01 = d00 - 2 * kdG2 + lat.r22
    d10 = d00 - 2 * kdG1 + 1
    d11 = d00 - 2 * kdG1 - 2 * kdG2 + 1 + 2 * lat.r12 + lat.r22
    if d11 <= d10 and d11 <= d01 and d11 <= d00:
        return (k1 - 1, k2 - 1)
    if d10 <= d01 and d10 <= d00:
        return (k1 - 1, k2)
    if d01 <= d00:
        return (k1, k2 - 1)
    return (k1, k2)


def v_coulomb(ql: float, same_layer: bool, d_l: float, D_l: float) -> float:
    """Landau level torus.jl:188-211."""
    DD = 2 * D_l
    V = W0
    if ql == 0.0:
        V *= (DD + d_l) * (DD - d_l) / (2 * DD) if same_layer else (DD - d_l) ** 2 / (2 * DD)
    else:
        ed = math.exp(-ql * d_l)
        eD = math.exp(-ql * DD)
        if same_layer:
            V *= (1 / ed - eD) * (ed - eD) / (1 - eD ** 2) / ql
        else:
            V *= (ed - eD) ** 2 / (1 - eD ** 2) / ql / ed
    return V


def _laguerre0(m: int, x: float) -> float:
    """L_m(x) by the three-term recurrence (Abramowitz & Stegun 22.7.12)."""
    a, b = 1.0, 1.0 - x
    if m == 0:
        return a
    for n in range(1, m):
        a, b = b, ((2 * n + 1 - x) * b - n * a) / (n + 1)
    return b


def v_haldane(ql: float, same_layer: bool, V_intra, V_inter) -> float:
    """Landau level torus.jl:217-232."""
    Vm = V_intra if same_layer else V_inter
    return W0 * sum(_laguerre0(i, ql * ql) * v for i, v in enumerate(Vm))


def _ita(g1: int, g2: int) -> int:
    return 1 if (g1 % 2 == 0 and g2 % 2 == 0) else -1


def _cross(a, b1, b2):
    return a[0] * b2 - a[1] * b1


class LandauInteraction:
    """Landau level torus.jl:243-347 for components (layer, level, Chern, other)."""

    def __init__(self, lattice: ReciprocalLattice, *components):
        self.lattice = lattice
        G1, G2 = lattice.G1, lattice.G2
        glmin = math.sqrt(min(abs(G1) ** 2, abs(G2) ** 2, abs(G1 + G2) ** 2, abs(G1 - G2) ** 2)) * L_MAG
        self.Nshell = math.ceil(max(2.5, GL_CUTOFF / glmin))
        self.D_l = 10.0
        self.d_l = 0.0
        self.V_intra = [0.0, 0.7]
        self.V_inter = [1.5, 0.0]
        self.mix = 0.0
        self.layer = [c[0] for c in components]
        self.level = [c[1] for c in components]
        self.chern = [c[2] for c in components]
        self.other = [c[3] for c in components]

    def __call__(self, kf1, kf2, ki2, ki1, cf1=1, cf2=1, ci2=1, ci1=1) -> complex:
        rd = lambda v: (v[cf1 - 1], v[cf2 - 1], v[ci2 - 1], v[ci1 - 1])  # noqa: E731
        lf1, lf2, li2, li1 = rd(self.layer)
        nf1, nf2, ni2, ni1 = rd(self.level)
        Cf1, Cf2, Ci2, Ci1 = rd(self.chern)
        sf1, sf2, si2, si1 = rd(self.other)
        if li1 != lf1 or li2 != lf2 or Ci1 != Cf1 or Ci2 != Cf2 or si1 != sf1 or si2 != sf2:
            return 0j
        if nf1 != ni1 or nf2 != ni2:
            raise NotImplementedError("level-changing form factors are not restated")
        kf1 = tuple(Fraction(v) for v in kf1)
        kf2 = tuple(Fraction(v) for v in kf2)
        ki1 = tuple(Fraction(v) for v in ki1)
        ki2 = tuple(Fraction(v) for v in ki2)
        d = (ki1[0] - kf1[0], ki1[1] - kf1[1])
        q = bz(d, self.lattice)
        gs1 = (round(d[0] - q[0]), round(d[1] - q[1]))
        gs2 = (round(kf2[0] - ki2[0] - q[0]), round(kf2[1] - ki2[1] - q[1]))
        q1, q2 = float(q[0]), float(q[1])
        G1, G2 = self.lattice.G1, self.lattice.G2
        same = li1 == li2
        c_i1f1 = _cross(ki1, kf1[0], kf1[1])
        c_i2f2 = _cross(ki2, kf2[0], kf2[1])
        s1 = (ki1[0] + kf1[0], ki1[1] + kf1[1])
        s2 = (ki2[0] + kf2[0], ki2[1] + kf2[1])
        total = 0j
        N = self.Nshell
        for g1 in range(-N, N + 1):
            for g2 in range(-N, N + 1):
                if abs(g1 * G1 + g2 * G2) ** 2 > GL_CUTOFF ** 2:
                    continue
                qq1, qq2 = q1 + g1, q2 + g2
                ql = (qq1 * G1 + qq2 * G2) * L_MAG
                aql = abs(ql)
                V = 0.0
                if self.mix != 0.0:
                    V += self.mix * v_haldane(aql, same, self.V_intra, self.V_inter)
                if self.mix != 1.0:
                    V += (1 - self.mix) * v_coulomb(aql, same, self.d_l, self.D_l)
                F = math.exp(-aql * aql / 4.0) ** 2      # two form factors, exp part
                if ni1 or ni2:                           # Laguerre parts (1 at n = 0)
                    F *= _laguerre0(ni1, aql * aql / 2.0) * _laguerre0(ni2, aql * aql / 2.0)
                ang = float(-Ci1 * c_i1f1)
                ang += -Ci1 * (float(s1[0]) * qq2 - float(s1[1]) * qq1)
                ang += float(-Ci2 * c_i2f2)
                ang += -Ci2 * (float(s2[0]) * (-qq2) - float(s2[1]) * (-qq1))
                phase = cmath.exp(1j * math.pi * ang)
                sign = _ita(g1 + gs1[0], g2 + gs1[1]) * _ita(g1 + gs2[0], g2 + gs2[1])
                total += sign * V * F * phase
        return total


def orbital_overlap(k_list, Gk, V: LandauInteraction, shift_f, shift_i):
    """Landau level torus.jl:352-393: <u_{k+shift_f} | u_{k+shift_i}> for every orbital
    (k fastest, then component), for the level-0 components of V. shift_f / shift_i: one
    (s1, s2) for all components or one per component. The form factor is the reference's
    one-argument level_form_factor (:174-181), exp(-1/4) L_n(|ql|^2 / 2) as written there
    (a real constant for n = 0)."""
    nc = len(V.layer)
    nk = len(k_list[0])
    if not isinstance(shift_f, list):
        shift_f = [shift_f] * nc
    if not isinstance(shift_i, list):
        shift_i = [shift_i] * nc
    G1, G2 = V.lattice.G1, V.lattice.G2
    out = []
    for c in range(nc):
        sf = tuple(Fraction(v) for v in shift_f[c])
        si = tuple(Fraction(v) for v in shift_i[c])
        q = ((sf[0] - si[0]) / Gk[0], (sf[1] - si[1]) / Gk[1])
        ql = (float(q[0]) * G1 + float(q[1]) * G2) * L_MAG
        F = math.exp(-0.25) * _laguerre0(0, 0.5 * abs(ql) ** 2)
        for k in range(nk):
            kf = ((k_list[0][k] + sf[0]) / Gk[0], (k_list[1][k] + sf[1]) / Gk[1])
            ki = ((k_list[0][k] + si[0]) / Gk[0], (k_list[1][k] + si[1]) / Gk[1])
            ang = kf[0] * ki[1] - kf[1] * ki[0]
            s = (ki[0] + kf[0], ki[1] + kf[1])
            ang += s[0] * q[1] - s[1] * q[0]
            out.append(F * cmath.exp(1j * math.pi * float(ang)))
    return out
