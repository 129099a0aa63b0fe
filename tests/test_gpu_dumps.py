"""Replay of a pair of dumps through the reference-named entry points (the reference's direct-test
runner, direct_test/evaluate_first_order_rho_run.cpp): dumps written from synthetic problems, read
back, value arrays filled at random as the runner does (wave padding rows included), then
init_dfpt_device_data, the rho1 entry point (GGA gradient, bias, ATOM_TILE_SIZE = 2; its default
precision, the INT8 split with 5 digits, within 1e-10) and the whole Hartree sum (FP64, within
1e-12), against the CPU oracle on the same read arrays."""
import numpy as np
import pytest

from fast_rpa_eval_amd import dumps
from tests.test_dumps import _problems, rho_view, sumup_view

pytestmark = pytest.mark.gpu


def test_replay_dumps_through_drop_in(cuda, oracle, tmp_path):
    T = 2
    prob, fp = _problems(T)
    pr, ps = tmp_path / "first_order_rho.meta.test.nproc_1.bin", tmp_path / "sumup.test.nproc_1.bin"
    dumps.write(pr, dumps.from_rho1_problem(prob))
    dumps.write(ps, dumps.from_far_problem(fp))
    rd, sd = dumps.read(pr, "rho"), dumps.read(ps, "sumup")
    vals = dumps.random_values(rd, sd, T, seed=7)
    print(f"[dump replay] {rd.n_my_batches_work} batches, {sd.n_valid_points} Hartree points, T = {T}", flush=True)
    got = dumps.replay_drop_in(rd, sd, vals, T)

    rv = rho_view(rd, vals["wave"])
    ref, gref = oracle.rho1(rv, vals["first_order_density_matrix"], T=T, bias=vals["first_order_rho_bias_part2"],
                            gwave=vals["gradient_wave"], gbias=vals["first_order_gradient_rho_bias_batches_atoms"])
    rel = np.abs(got["first_order_rho"] - ref).max() / np.abs(ref).max()
    grel = np.abs(got["first_order_gradient_rho"] - gref).max() / np.abs(gref).max()
    sv = sumup_view(sd, T, vals["current_delta_v_hart_part_spl_tile"])
    vref = oracle.near(sv, delta_v=oracle.far(sv, delta_v=vals["delta_v_hartree"].copy(order="F")))
    vrel = np.abs(got["delta_v_hartree"] - vref).max() / np.abs(vref - vals["delta_v_hartree"]).max()
    print(f"[dump replay] rho1 rel {rel:.2e}, gradient rel {grel:.2e}, delta_v rel {vrel:.2e}", flush=True)
    assert rel <= 1e-10 and grel <= 1e-10 and vrel <= 1e-12
