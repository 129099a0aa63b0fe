"""Cartesian multipole term tables for the far-field Hartree sum.

The far-field branch of ``sum_up_whole_potential`` (fast_rap_src_dfpt_gpu_eval/
sum_up_whole_potential.cpp:91-187,345-378) never evaluates spherical harmonics.  It expands each
irregular solid harmonic Y_lm(r^) / r^(l+1) into Cartesian terms

    x^i y^j z^k * Fp(nn) * cc(n) * M_lm,      Fp(nn) = (-1)^nn (2nn-1)!! / r^(2nn+1),

driven by four caller-supplied tables (index_cc, cc, n_cc_lm_ijk, index_ijk_max_cc).  In
FHI-aims these tables come from the Fortran host; here they are generated from the mathematics:

* real solid harmonics r^l Y_lm = N_lm * A_m(x, y) * r^(l-|m|) P_l^(|m|)(z/r), with A_m the real
  or imaginary part of (x + i y)^|m| and P_l^(|m|) the |m|-th derivative of the Legendre
  polynomial;
* x^2 eliminated through x^2 = r^2 - y^2 - z^2 until the x exponent is 0 or 1;
* each surviving monomial x^i y^j z^k r^(2s) of an (l, m) becomes one row
  (l, m, i, j, k, nn = l - s), rows of one (l, m) sorted by (j, k), (l, m) in the order
  l = 0, 1, ...; m = -l..l.

That rule reproduces the 176 (i, j, k, nn) rows and n_cc_lm_ijk = 1,4,12,28,57,104,176 that the
reference's AOT dot product holds for l <= 6 (sum_up_far_dist_potential_dot.hpp:12-210).

The sign/normalisation convention of Y_lm is the one of SHEval (fast_rap_src_dfpt_gpu_eval/
sheval.hpp:9-329, index l^2 + l + m):

    m = 0:  K_l0 P_l(cos t)
    m > 0:  (-1)^m sqrt(2) K_lm P_l^m(cos t) cos(m phi)      (P_l^m without Condon-Shortley phase)
    m < 0:  sqrt(2) K_l|m| P_l^|m|(cos t) sin(|m| phi)
    K_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!)

``cc`` is chosen so that the Cartesian sum equals sum_lm M_lm Y_lm(r^) / r^(l+1) exactly.
"""
from fractions import Fraction
from math import comb, factorial, pi, sqrt

import numpy as np


def _legendre_derivative(l, m):
    """d^m/dt^m P_l(t) as [(power, Fraction)]."""
    out = []
    for k in range(l // 2 + 1):
        p = l - 2 * k
        if p < m:
            continue
        c = Fraction((-1) ** k * comb(l, k) * comb(2 * l - 2 * k, l), 2 ** l)
        c *= Fraction(factorial(p), factorial(p - m))
        out.append((p - m, c))
    return out


def _xy_part(m, imag):
    """Real (imag=False) or imaginary (imag=True) part of (x + i y)^m as {(a, j): int}."""
    res = {}
    for q in range(m + 1):
        if (q % 2 == 1) != imag:
            continue
        sign = 1 if (q % 4) in (0, 1) else -1
        res[(m - q, q)] = res.get((m - q, q), 0) + sign * comb(m, q)
    return res


def sh_norm(l, m):
    """Normalisation N_lm of the SHEval-convention real harmonic (see module docstring)."""
    am = abs(m)
    k = sqrt((2 * l + 1) / (4 * pi) * factorial(l - am) / factorial(l + am))
    if m == 0:
        return k
    if m > 0:
        return (-1) ** m * sqrt(2.0) * k
    return sqrt(2.0) * k


def solid_harmonic_terms(l, m):
    """Exact expansion of r^l Y_lm / N_lm as {(i, j, k, s): Fraction}, i in {0, 1}."""
    am = abs(m)
    xy = _xy_part(am, imag=(m < 0))
    poly = {}
    for (a, j), cxy in xy.items():
        for p, cz in _legendre_derivative(l, am):
            s = (l - am - p) // 2
            key = (a, j, p, s)
            poly[key] = poly.get(key, 0) + cxy * cz
    # eliminate x^2 = r^2 - y^2 - z^2
    changed = True
    while changed:
        changed = False
        nxt = {}
        for (a, j, k, s), c in poly.items():
            if c == 0:
                continue
            if a >= 2:
                changed = True
                for key, sg in (((a - 2, j, k, s + 1), 1), ((a - 2, j + 2, k, s), -1), ((a - 2, j, k + 2, s), -1)):
                    nxt[key] = nxt.get(key, 0) + sg * c
            else:
                nxt[(a, j, k, s)] = nxt.get((a, j, k, s), 0) + c
        poly = nxt
    return {key: c for key, c in poly.items() if c != 0}


def _double_factorial_odd(n):
    """(2n-1)!! with (-1)!! = 1."""
    r = 1
    for v in range(2 * n - 1, 0, -2):
        r *= v
    return r


def cartesian_terms(l_max):
    """Rows (l, m, i, j, k, nn, cc) for l = 0..l_max in table order."""
    rows = []
    for l in range(l_max + 1):
        for m in range(-l, l + 1):
            terms = solid_harmonic_terms(l, m)
            nrm = sh_norm(l, m)
            for (i, j, k, s) in sorted(terms, key=lambda t: (t[1], t[2])):
                nn = l - s
                fp = (-1) ** nn * _double_factorial_odd(nn)
                cc = float(terms[(i, j, k, s)] / fp) * nrm
                rows.append((l, m, i, j, k, nn, cc))
    return rows


def far_tables(l_max_analytic_multipole):
    """Tables in the reference's caller layout (column-major, as passed through the Fortran ABI).

    Returns dict with
      index_cc          int32 (index_cc_dim_0, 6)  columns l, m, i, j, k, nn   (Fortran order)
      cc                float64 (index_cc_dim_0,)
      n_cc_lm_ijk       int32 (l_max_analytic_multipole + 1,)
      index_ijk_max_cc  int32 (3, l_max_analytic_multipole + 1)  (Fortran order)
      index_cc_dim_0    int
    """
    rows = cartesian_terms(l_max_analytic_multipole)
    dim0 = len(rows)
    index_cc = np.zeros((dim0, 6), dtype=np.int32, order="F")
    cc = np.zeros(dim0, dtype=np.float64)
    for n, (l, m, i, j, k, nn, c) in enumerate(rows):
        index_cc[n] = (l, m, i, j, k, nn)
        cc[n] = c
    L = l_max_analytic_multipole
    n_cc = np.array([sum(1 for r in rows if r[0] <= l) for l in range(L + 1)], dtype=np.int32)
    ijk_max = np.zeros((3, L + 1), dtype=np.int32, order="F")
    for l in range(L + 1):
        sel = index_cc[: n_cc[l]]
        ijk_max[:, l] = sel[:, 2:5].max(axis=0)
    return {"index_cc": index_cc, "cc": cc, "n_cc_lm_ijk": n_cc, "index_ijk_max_cc": ijk_max,
            "index_cc_dim_0": dim0}


def real_sph_harm(l_max, xyz):
    """SHEval-convention real Y_lm at unit-vector rows xyz, column index l^2 + l + m.

    Evaluated from the exact solid-harmonic expansion (r = 1), for tests and generators.
    """
    xyz = np.asarray(xyz, dtype=np.float64)
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    out = np.zeros(xyz.shape[:-1] + ((l_max + 1) ** 2,))
    for l in range(l_max + 1):
        for m in range(-l, l + 1):
            v = 0.0
            for (i, j, k, s), c in solid_harmonic_terms(l, m).items():
                v = v + float(c) * x ** i * y ** j * z ** k
            out[..., l * l + l + m] = sh_norm(l, m) * v
    return out


def horner_fma_count(L):
    """FMAs of the device's nested-Horner evaluation of one far-field pair at l = L.

    Mirrors the kernel's structure: sum_q u^q Q_q(X, Y, Z), Q_q nested in X (degree <= 1), Y, Z
    (fast_rpa_eval_amd/csrc/farfield.hip, hz/hy/hq/hu).  Equals the merged-coefficient count plus
    the Horner joins.
    """
    def hz(q, i, j):
        return L - 2 * q - i - j

    def hy(q, i, j):
        jm = L - 2 * q - i
        return hz(q, i, j) if j == jm else 1 + hy(q, i, j + 1) + hz(q, i, j)

    def hq(q):
        return hy(q, 0, 0) if L - 2 * q == 0 else 1 + hy(q, 1, 0) + hy(q, 0, 0)

    def hu(q):
        return hq(q) if 2 * (q + 1) > L else 1 + hu(q + 1) + hq(q)

    return hu(0)


def horner_ncoef(L):
    """Merged (i, j, k, q) coefficients at l = L: sum_q (L - 2q + 1)^2."""
    return sum((L - 2 * q + 1) ** 2 for q in range(L // 2 + 1))
