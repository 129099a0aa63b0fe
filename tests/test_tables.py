"""Far-field term tables and the real spherical-harmonic convention (CPU).

Pins: the reference's AOT term table (176 rows, n_cc_lm_ijk, index_cc_dim_0 for l = 7), the
reference's SHEval coefficients for l <= 3, scipy's spherical harmonics for l <= 8, and the closed
form sum_lm M_lm Y_lm / r^(l+1) that the cc coefficients must reproduce.
"""
import json
import os

import numpy as np
import pytest

from fast_rpa_eval_amd import multipole_tables as mt

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_term_table_matches_reference_aot_rows():
    g = _load("far_terms_l6.json")
    t = mt.far_tables(7)
    assert t["index_cc_dim_0"] == g["index_cc_dim_0_l7"] == 280
    assert list(t["n_cc_lm_ijk"][:7]) == g["n_cc_lm_ijk"] == [1, 4, 12, 28, 57, 104, 176]
    assert t["index_cc"][:176, 2:6].tolist() == g["terms"]


def test_term_table_invariants_to_l10():
    rows = mt.cartesian_terms(10)
    for (l, m, i, j, k, nn, c) in rows:
        assert i in (0, 1) and 2 * nn == l + i + j + k and c != 0.0
    t = mt.far_tables(9)
    assert list(t["n_cc_lm_ijk"][7:]) == [280, 425, 620]
    ijk = t["index_ijk_max_cc"]
    for l in range(10):
        assert tuple(ijk[:, l]) == (min(l, 1), l, l)


def test_sph_harm_matches_reference_sheval_l3():
    rng = np.random.default_rng(0)
    v = rng.normal(size=(64, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    Y = mt.real_sph_harm(3, v)
    for e in _load("sheval_l3.json")["entries"]:
        ref = sum(c * v[:, 0] ** i * v[:, 1] ** j * v[:, 2] ** k for i, j, k, c in e["poly"])
        np.testing.assert_allclose(Y[:, e["index"]], ref, rtol=0, atol=2e-15, err_msg=str(e))


def test_sph_harm_matches_scipy_to_l8():
    sp = pytest.importorskip("scipy.special")
    rng = np.random.default_rng(1)
    v = rng.normal(size=(40, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    Y = mt.real_sph_harm(8, v)
    th, ph = np.arccos(v[:, 2]), np.arctan2(v[:, 1], v[:, 0])
    for l in range(9):
        for m in range(-l, l + 1):
            c = sp.sph_harm_y(l, abs(m), th, ph)  # complex, with the Condon-Shortley phase
            if m == 0:
                ref = c.real
            elif m > 0:
                ref = np.sqrt(2.0) * c.real
            else:
                ref = np.sqrt(2.0) * (-1) ** m * c.imag
            np.testing.assert_allclose(Y[:, l * l + l + m], ref, rtol=0, atol=1e-13)


def cartesian_sum(l_max, d, M):
    """sum_n x^i y^j z^k Fp(nn) cc(n) M(lm(n)) for one centre: the far-field sum of the reference."""
    t = mt.far_tables(l_max)
    r2 = np.sum(d * d, axis=1)
    r = np.sqrt(r2)
    Fp = [1.0 / r]
    for k in range(1, l_max + 1):
        Fp.append(Fp[-1] * (1 - 2 * k) / r2)
    out = np.zeros(len(d))
    for n in range(t["n_cc_lm_ijk"][l_max]):
        l, m, i, j, k, nn = t["index_cc"][n]
        out += d[:, 0] ** i * d[:, 1] ** j * d[:, 2] ** k * Fp[nn] * t["cc"][n] * M[l * l + l + m]
    return out


@pytest.mark.parametrize("l_max", [0, 2, 4, 6, 8])
def test_cc_reproduces_closed_form(l_max):
    rng = np.random.default_rng(l_max)
    d = rng.normal(size=(200, 3)) * 12.0
    M = rng.standard_normal((l_max + 1) ** 2)
    r = np.linalg.norm(d, axis=1)
    Y = mt.real_sph_harm(l_max, d / r[:, None])
    ref = sum(M[l * l + l + m] * Y[:, l * l + l + m] / r ** (l + 1) for l in range(l_max + 1) for m in range(-l, l + 1))
    got = cartesian_sum(l_max, d, M)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_cc_single_multipole_is_one_harmonic():
    """Unit M_lm for one (l, m): the sum is exactly Y_lm(r^) / r^(l+1); a dropped term or wrong sign fails."""
    rng = np.random.default_rng(7)
    d = rng.normal(size=(100, 3)) * 9.0
    r = np.linalg.norm(d, axis=1)
    Y = mt.real_sph_harm(6, d / r[:, None])
    for l in range(7):
        for m in range(-l, l + 1):
            M = np.zeros(49)
            M[l * l + l + m] = 1.0
            got = cartesian_sum(6, d, M)
            np.testing.assert_allclose(got, Y[:, l * l + l + m] / r ** (l + 1), rtol=1e-11, atol=1e-16)


def test_generated_term_header_is_current():
    from fast_rpa_eval_amd import build
    path = os.path.join(build.CSRC, "far_terms.inc")
    with open(path) as f:
        assert f.read() == build.far_terms_header()
