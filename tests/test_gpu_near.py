"""Near-field Hartree sum (spline branch of sum_up_whole_potential) on the GPU, through the C-ABI,
against the CPU FP64 oracle.

Bar: within 1e-12 of max|dV| (FP64 both sides).  Cases: l_pot_max 0..10 (the reference's device
code stops at 6), two species with different grids and l_hartree, T = 2 full and short tiles,
accumulation into a nonzero delta_v_hartree, batch-ordered points with ~20 near centres each (many
pair-queue flushes per wave), interval clamping at both ends of the grid, run-to-run determinism and
the host-side argument checks.
"""
import numpy as np
import pytest
import torch

from fast_rpa_eval_amd import api, synth

pytestmark = pytest.mark.gpu


def _gpu_near(fp, dv0=None, n_tile=None):
    plan = api.FarPlan.from_problem(fp)
    inp = api.FarInputs(fp)
    dv = api.to_device(dv0 if dv0 is not None else np.zeros((fp.n_full_points, fp.n_tile), order="F"))
    plan.near_forward(dv, *inp.near_args(), T=fp.n_tile, n_tile=n_tile)
    torch.cuda.synchronize()
    return dv.cpu().numpy().reshape((fp.n_full_points, fp.n_tile), order="F")


def _check(got, ref, base, what, rel=1e-12):
    scale = np.abs(ref - base).max()
    err = np.abs(got - ref).max()
    print(f"[{what}] max|err| {err:.3e}, max|dV| {scale:.3e}, rel {err / scale:.3e}", flush=True)
    assert np.all(np.isfinite(got))
    assert scale > 0 and err <= rel * scale, what


@pytest.mark.parametrize("l_max", [0, 2, 6, 8, 10])
def test_near_uniform_points_two_species(cuda, oracle, l_max):
    """10 atoms, points uniform around them, T = 2 (both perturbations), decreasing outer radius,
    accumulated into a random delta_v_hartree."""
    print(f"[near uniform l_max={l_max}]", flush=True)
    fp = synth.make_far_problem(10, 11.0, 3000, l_max, config_k=4, all_far=False, n_tile=2,
                                outer_radius="decreasing")
    synth.add_near_field(fp, 60, "smooth", n_species=2, seed=l_max)
    dv0 = np.asfortranarray(np.random.default_rng(5).uniform(-1, 1, size=(fp.n_full_points, 2)))
    ref = oracle.near(fp, delta_v=dv0.copy(order="F"))
    got = _gpu_near(fp, dv0)
    _check(got, ref, dv0, f"near l={l_max}")


def _batch_problem(l_max, n_tile, mr_bohr, seed):
    prob = synth.make_rho1_problem(64, 12.0, 200, (40, 160), (20, 60), config_k=6, fill_wave=False)
    fp = synth.make_far_problem(64, 12.0, 0, l_max, config_k=6, n_tile=n_tile, xyz=prob.atom_xyz,
                                min_dist_bohr=mr_bohr, grid=(prob.coord_points, prob.valid_map()))
    synth.add_near_field(fp, 90, "smooth", n_species=2, seed=seed)
    return fp


def test_near_batch_points_many_centres(cuda, oracle):
    """Integration-grid points in spatially ordered batches, multipole radius 10 bohr in a dense
    64-atom cluster: about 20 near centres per point, so every wave flushes its pair queue many times."""
    fp = _batch_problem(6, 1, 10.0, 1)
    ref = oracle.near(fp)
    pts = fp.coord_points[:, fp.i_valid_point_2_i_full_points_map].T
    n_pairs = sum(int(np.sum(np.sum((pts - fp.coords_center[:, c]) ** 2, axis=1) < 100.0)) for c in range(fp.n_atoms))
    print(f"[near batches] {fp.n_valid_points} points, {n_pairs} near pairs", flush=True)
    assert n_pairs > 10 * fp.n_valid_points
    got = _gpu_near(fp)
    _check(got, ref, 0.0, "near batches")
    again = _gpu_near(fp)
    assert np.array_equal(got, again), "near field not deterministic"


def test_near_short_tile_l8(cuda, oracle):
    """T = 2 with a short tile (n_tile = 1): the second perturbation's column stays untouched."""
    fp = _batch_problem(8, 2, 8.0, 2)
    ref = oracle.near(fp, j_atom_begin=5, j_atom_end=5)
    got = _gpu_near(fp, n_tile=1)
    assert np.all(got[:, 1] == 0.0)
    _check(got, ref, 0.0, "near short tile l=8")


def test_near_clamped_intervals_cubic(cuda, oracle):
    """Cubic tiles on a grid ending at 5 bohr (points beyond take the last interval) and two points
    closer to an atom than r_grid_min (first interval, t < 0)."""
    fp = synth.make_far_problem(10, 11.0, 3000, 6, config_k=4, all_far=False, n_tile=1)
    synth.add_near_field(fp, 50, "cubic", n_species=2, r_max_bohr=5.0, seed=3)
    a0 = fp.coords_center[:, 0]
    extra = np.stack([a0 + [3e-5, 0, 0], a0 + [0, -2e-5, 4e-5]], axis=1)
    fp.coord_points = np.asfortranarray(np.concatenate([fp.coord_points, extra], axis=1))
    n_full = fp.coord_points.shape[1]
    fp.i_valid_point_2_i_full_points_map = np.concatenate(
        [fp.i_valid_point_2_i_full_points_map, [n_full - 2, n_full - 1]]).astype(np.int32)
    fp.n_full_points, fp.n_valid_points = n_full, len(fp.i_valid_point_2_i_full_points_map)
    ref = oracle.near(fp)
    got = _gpu_near(fp)
    _check(got, ref, 0.0, "near clamped cubic")


def test_near_argument_checks(cuda):
    fp = synth.make_far_problem(6, 10.0, 500, 4, config_k=4, all_far=False)
    plan = api.FarPlan.from_problem(fp)  # no near tables
    synth.add_near_field(fp, 40, "smooth")
    inp = api.FarInputs(fp)
    dv = torch.zeros(fp.n_full_points, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="species grids not set"):
        plan.near_forward(dv, *inp.near_args())
    nf = fp.near
    plan.set_near(nf["species_center"], nf["n_grid"], nf["r_grid_min"], nf["log_r_grid_inc"])
    args = list(inp.near_args())
    args[-1] = int(nf["n_grid"].max()) - 1  # n_hartree_grid below a species' n_grid
    with pytest.raises(RuntimeError, match="n_hartree_grid"):
        plan.near_forward(dv, *args)
    with pytest.raises(RuntimeError, match="n_grid"):
        plan.set_near(nf["species_center"], np.array([1, 40], np.int32), nf["r_grid_min"], nf["log_r_grid_inc"])
    torch.cuda.synchronize()
    assert torch.all(dv == 0)
