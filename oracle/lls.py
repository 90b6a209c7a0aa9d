"""Lowest-Landau-level interaction on the Haldane sphere — TEST INFRASTRUCTURE ONLY (see
oracle.py header).

Restatement of the reference example's SphereLandauInteraction callable (examples/landau
level sphere/Landau level sphere.jl:190-307) for model Haldane pseudopotentials (mix = 1):

    <m_f1 m_f2| V |m_i1 m_i2> = sum_L V_{2S-L} <S m_i1; S m_i2|L M> <S m_f1; S m_f2|L M>

with orbitals labelled by shifted angular momentum m~ = 0..2S (physical L_z = m~ - S) and zero
unless m~_f1 + m~_f2 = m~_i1 + m~_i2. The Clebsch-Gordan coefficients come from
WignerSymbols.clebschgordan in the reference (absent here); they are computed by Racah's
closed formula (A. R. Edmonds, Angular Momentum in Quantum Mechanics, eq. 3.6.11) in exact
integer arithmetic under one square root, and pinned against textbook values and
orthogonality. The Coulomb pseudopotentials (:99-172, Legendre coefficients and 6j symbols)
are not restated.
"""
from __future__ import annotations

import math
from fractions import Fraction


def _fact(x: Fraction) -> int:
    assert x.denominator == 1 and x >= 0, x
    return math.factorial(int(x))


def clebsch_gordan(j1, m1, j2, m2, J, M) -> float:
    """<j1 m1; j2 m2 | J M> for integer or half-integer arguments (Fractions or ints)."""
    j1, m1, j2, m2, J, M = (Fraction(v) for v in (j1, m1, j2, m2, J, M))
    if m1 + m2 != M or abs(m1) > j1 or abs(m2) > j2 or abs(M) > J:
        return 0.0
    if J < abs(j1 - j2) or J > j1 + j2 or (j1 + j2 + J).denominator != 1:
        return 0.0
    for j, m in ((j1, m1), (j2, m2), (J, M)):
        if (j - m).denominator != 1:
            return 0.0
    pre = Fraction((2 * J + 1) * _fact(J + j1 - j2) * _fact(J - j1 + j2) * _fact(j1 + j2 - J),
                   _fact(j1 + j2 + J + 1))
    pre *= (_fact(J + M) * _fact(J - M) * _fact(j1 - m1) * _fact(j1 + m1) * _fact(j2 - m2)
            * _fact(j2 + m2))
    s = Fraction(0)
    k = 0
    while True:
        args = (Fraction(k), j1 + j2 - J - k, j1 - m1 - k, j2 + m2 - k, J - j2 + m1 + k,
                J - j1 - m2 + k)
        if args[1] < 0 or args[2] < 0 or args[3] < 0:
            break
        if args[4] >= 0 and args[5] >= 0:
            den = 1
            for a in args:
                den *= _fact(a)
            s += Fraction((-1) ** k, den)
        k += 1
    # sqrt(pre) * s with the sign of s; pre is an exact rational
    return math.copysign(math.sqrt(float(pre * s * s)), float(s)) if s != 0 else 0.0


class SphereHaldaneInteraction:
    """Landau level sphere.jl:190-307 with mix = 1: model pseudopotentials V_intra[m_rel]."""

    def __init__(self, N_phi: int, V_intra):
        assert N_phi >= 2
        self.N_phi = N_phi
        self.V_intra = list(V_intra)

    def __call__(self, kf1, kf2, ki2, ki1, cf1=1, cf2=1, ci2=1, ci1=1) -> complex:
        mf1, mf2, mi1, mi2 = (int(k[0]) for k in (kf1, kf2, ki1, ki2))
        if mf1 + mf2 != mi1 + mi2:
            return 0j
        tS = self.N_phi - 1
        S = Fraction(tS, 2)
        a1, a2, b1, b2 = mi1 - S, mi2 - S, mf1 - S, mf2 - S
        M = a1 + a2
        total = 0.0
        for L in range(tS + 1):
            if abs(M) > L:
                continue
            m_rel = tS - L
            V = self.V_intra[m_rel] if m_rel < len(self.V_intra) else 0.0
            if V == 0.0:
                continue
            cgi = clebsch_gordan(S, a1, S, a2, L, M)
            if cgi == 0.0:
                continue
            cgf = clebsch_gordan(S, b1, S, b2, L, M)
            total += V * cgi * cgf
        return complex(total, 0.0)
