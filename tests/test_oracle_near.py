"""CPU oracle of the near-field branch of sum_up_whole_potential (oracle/nearfield.c).

Pins:
  * oracle_ylm against scipy's spherical harmonics (l <= 10, poles included) and the reference's
    SHEval polynomials (l <= 3 fixture);
  * the whole near sum against its closed form for spline tiles cubic in the grid index -- the
    cubic Hermite interpolant reproduces a cubic exactly, also where the interval index is clamped
    (r below r_grid_min, r beyond the last grid point) -- with two species, T = 2 and l reduced by
    outer_potential_radius;
  * node values: at r on a grid point the sum is sum_lm Y_lm c0(lm, node) for any tile (placement of
    the value and derivative planes and of the interval offset);
  * the near/far split: the near sum takes exactly the pairs with d^2 < multipole_radius_sq.
"""
import json
import os

import numpy as np
import pytest

from fast_rpa_eval_amd import multipole_tables as mt
from fast_rpa_eval_amd import synth

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _dirs(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v[0], v[1], v[2] = [0, 0, 1], [0, 0, -1], [1e-40, 0, 1]
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_ylm_matches_scipy_to_l10(oracle):
    sp = pytest.importorskip("scipy.special")
    v = _dirs(60, 3)
    Y = oracle.ylm(10, v)
    th, ph = np.arccos(np.clip(v[:, 2], -1, 1)), np.arctan2(v[:, 1], v[:, 0])
    for l in range(11):
        for m in range(-l, l + 1):
            c = sp.sph_harm_y(l, abs(m), th, ph)  # complex, with the Condon-Shortley phase
            ref = c.real if m == 0 else (np.sqrt(2.0) * c.real if m > 0 else np.sqrt(2.0) * (-1) ** m * c.imag)
            np.testing.assert_allclose(Y[:, l * l + l + m], ref, rtol=0, atol=1e-13, err_msg=f"l={l} m={m}")


def test_ylm_matches_reference_sheval_l3(oracle):
    v = _dirs(40, 4)
    Y = oracle.ylm(3, v)
    with open(os.path.join(GOLD, "sheval_l3.json")) as f:
        entries = json.load(f)["entries"]
    for e in entries:
        ref = sum(c * v[:, 0] ** i * v[:, 1] ** j * v[:, 2] ** k for i, j, k, c in e["poly"])
        np.testing.assert_allclose(Y[:, e["index"]], ref, rtol=0, atol=2e-15, err_msg=str(e))


def _problem(kind, r_max=40.0, l_max=6, n_tile=2, opr_small=False, seed=0):
    fp = synth.make_far_problem(10, 11.0, 3000, l_max, config_k=4, n_tile=n_tile, outer_radius="decreasing",
                                all_far=False)
    if opr_small:
        # l lowered inside the near region: outer_potential_radius(l) = (2 + l)^2 bohr^2
        for l in range(fp.l_pot_max + 1):
            fp.outer_potential_radius[l] = (2.0 + l) ** 2
    synth.add_near_field(fp, 50, kind, n_species=2, r_max_bohr=r_max, seed=seed)
    return fp


def _closed_form(fp):
    """sum over near pairs of sum_lm f_lm(i_r_log) Y_lm(u) for the cubic tiles (numpy, scipy-pinned Y)."""
    nf = fp.near
    T = fp.n_tile
    out = np.zeros((fp.n_full_points, T))
    pts = fp.coord_points[:, fp.i_valid_point_2_i_full_points_map].T
    for t in range(T):
        for ic, c1 in enumerate(fp.centers_hartree_potential):
            cur = c1 - 1
            a = fp.center_to_atom[cur] - 1
            d = pts - fp.coords_center[:, cur][None, :]
            d2 = np.sum(d * d, axis=1)
            near = d2 < fp.multipole_radius_sq[a, t]
            if not near.any():
                continue
            dn, d2n = d[near], d2[near]
            lh = fp.l_hartree[fp.species[a] - 1]
            lvec = np.full(len(d2n), lh)
            for _ in range(lh):  # lower l while outer_potential_radius(l) < d^2
                lower = (fp.outer_potential_radius[lvec, a, t] < d2n) & (lvec > 0)
                lvec = lvec - lower
            s = nf["species_center"][cur] - 1
            r = np.sqrt(d2n)
            irl = 1.0 + np.log(r / nf["r_grid_min"][s]) / nf["log_r_grid_inc"][s]
            Y = mt.real_sph_harm(fp.l_pot_max, dn / r[:, None])
            cf = nf["coef"][:, :, a, t]
            f = cf[:, 0][None, :] + irl[:, None] * (cf[:, 1][None, :] + irl[:, None] * (
                cf[:, 2][None, :] + irl[:, None] * cf[:, 3][None, :]))
            lm_l = np.array([l for l in range(fp.l_pot_max + 1) for _ in range(2 * l + 1)])
            mask = lm_l[None, :] <= lvec[:, None]
            idx = fp.i_valid_point_2_i_full_points_map[near]
            np.add.at(out[:, t], idx, np.sum(np.where(mask, Y * f, 0.0), axis=1))
    return out


@pytest.mark.parametrize("r_max,opr_small", [(40.0, False), (5.0, True)])
def test_near_cubic_tile_closed_form(oracle, r_max, opr_small):
    fp = _problem("cubic", r_max=r_max, opr_small=opr_small)
    # two points closer to atom 0 than r_grid_min (interval clamped to 1, t < 0)
    a0 = fp.coords_center[:, 0]
    extra = np.stack([a0 + [3e-5, 0, 0], a0 + [0, -2e-5, 4e-5]], axis=1)
    fp.coord_points = np.asfortranarray(np.concatenate([fp.coord_points, extra], axis=1))
    n_full = fp.coord_points.shape[1]
    fp.i_valid_point_2_i_full_points_map = np.concatenate(
        [fp.i_valid_point_2_i_full_points_map, [n_full - 2, n_full - 1]]).astype(np.int32)
    fp.n_full_points, fp.n_valid_points = n_full, len(fp.i_valid_point_2_i_full_points_map)
    got = oracle.near(fp)
    ref = _closed_form(fp)
    nf = fp.near
    r_all = []
    pts = fp.coord_points[:, fp.i_valid_point_2_i_full_points_map].T
    for c in range(fp.n_atoms):
        d2 = np.sum((pts - fp.coords_center[:, c]) ** 2, axis=1)
        r_all.append(np.sqrt(d2[d2 < fp.multipole_radius_sq[c, 0]]))
    r_all = np.concatenate(r_all)
    assert (r_all < nf["r_grid_min"].min()).any()          # clamped below
    if r_max < 8.0:
        assert (r_all > r_max).any()                        # clamped above
    assert np.count_nonzero(ref) > 100
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"[near closed form r_max {r_max}] rel {err:.2e}")
    assert err <= 1e-12


def test_near_node_values(oracle):
    """Points on grid nodes k of atom 0's species: the sum is sum_lm Y_lm c0(lm, k)."""
    fp = _problem("smooth", n_tile=1)
    nf = fp.near
    fp.multipole_radius_sq[:] = 1e-30
    fp.multipole_radius_sq[0, 0] = 64.0  # only centre 0 is near anything
    s = nf["species_center"][0] - 1
    rng = np.random.default_rng(7)
    ks = np.array([3, 10, 25, nf["n_grid"][s] - 5])
    u = rng.normal(size=(len(ks), 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = nf["r_grid_min"][s] * np.exp(ks * nf["log_r_grid_inc"][s])
    keep = r * r < 64.0
    ks, u, r = ks[keep], u[keep], r[keep]
    pts = fp.coords_center[:, 0][None, :] + u * r[:, None]
    fp.coord_points = np.asfortranarray(pts.T)
    fp.n_full_points = fp.n_valid_points = len(pts)
    fp.i_valid_point_2_i_full_points_map = np.arange(len(pts), dtype=np.int32)
    got = oracle.near(fp)[:, 0]
    L = fp.l_hartree[fp.species[0] - 1]
    Y = mt.real_sph_harm(fp.l_pot_max, u)[:, :(L + 1) ** 2]
    ref = np.sum(Y * nf["spl"][:(L + 1) ** 2, 0, ks, 0, 0].T, axis=1)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_near_far_split(oracle):
    fp = _problem("smooth", n_tile=1)
    base = oracle.near(fp)
    assert np.count_nonzero(base) > 0
    fp.multipole_radius_sq[:] = 0.0
    assert np.all(oracle.near(fp) == 0.0)
    # every pair near: the far branch adds nothing and the near branch takes all
    fp.multipole_radius_sq[:] = 1e12
    assert np.all(oracle.far(fp) == 0.0)
    assert np.count_nonzero(oracle.near(fp)) == fp.n_valid_points
