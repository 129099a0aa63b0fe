"""frpa_far_forward_noalloc (the far-field forward whose coefficient table lives in a fixed device
array) against frpa_far_forward, both called through the C-ABI: bit-identical delta_v_hartree for
both kernels on two calls in a row, the second with a smaller problem (fewer atoms, lower l, fewer
points, smaller tile), both also within 1e-12 of max|dV| of the CPU oracle
(sum_up_whole_potential.cpp:382-503 far branch); a call whose table would not fit the array is
refused before anything runs and leaves delta_v_hartree untouched.
"""
import ctypes

import numpy as np
import pytest
import torch

from fast_rpa_eval_amd import _lib, api, synth

pytestmark = pytest.mark.gpu

NOALLOC_DOUBLES = 4194304  # FRPA_FAR_NOALLOC_SCRATCH_DOUBLES


def _bind():
    L = _lib.lib()
    p, i = ctypes.c_void_p, ctypes.c_int
    L.frpa_far_forward_noalloc.argtypes = [p, i, i, i, i, i, p, p, p, p, p, p, p, p]
    return L


def _call(fn, plan, inp, T, n_tile, variant, dv):
    args = inp.args()
    return fn(plan.handle, T, n_tile, variant, inp.n_full_points, args[1].numel(), api._dp(args[0]),
              api._dp(args[1]), api._dp(args[2]), api._dp(args[3]), api._dp(args[4]), api._dp(args[5]),
              api._dp(dv), api._stream())


def _case(L, oracle, fp, n_tile, variant, seed):
    plan = api.FarPlan.from_problem(fp)
    inp = api.FarInputs(fp)
    T = fp.n_tile
    dv0 = np.asfortranarray(np.random.default_rng(seed).uniform(-1, 1, size=(fp.n_full_points, T)))
    dv_ref = api.to_device(dv0)
    dv_new = api.to_device(dv0)
    _lib.check(_call(L.frpa_far_forward, plan, inp, T, n_tile, variant, dv_ref), "frpa_far_forward")
    _lib.check(_call(L.frpa_far_forward_noalloc, plan, inp, T, n_tile, variant, dv_new), "frpa_far_forward_noalloc")
    torch.cuda.synchronize()
    a = dv_ref.cpu().numpy()
    b = dv_new.cpu().numpy()
    assert np.array_equal(a, b), f"noalloc differs from frpa_far_forward: max {np.abs(a - b).max():.3e}"
    ref = oracle.far(fp, delta_v=dv0.copy(order="F"), j_atom_begin=1, j_atom_end=n_tile).ravel(order="F")
    err = np.abs(b - ref).max() / np.abs(ref).max()
    print(f"[noalloc] atoms {len(fp.species)} l {fp.l_pot_max} T {T} n_tile {n_tile} variant {variant} "
          f"points {inp.map.numel()}: bit-identical, rel err vs oracle {err:.3e}", flush=True)
    assert err <= 1e-12
    return plan, inp


@pytest.mark.parametrize("variant", [api.FAR_HORNER, api.FAR_TERMS])
def test_noalloc_matches_allocating_entry(cuda, oracle, variant):
    L = _bind()
    big = synth.make_far_problem(48, 16.0, 6000, 8, config_k=7, all_far=False, n_tile=2,
                                 outer_radius="decreasing")
    small = synth.make_far_problem(12, 10.0, 1500, 6, config_k=4, all_far=False, n_tile=1)
    _case(L, oracle, big, 2, variant, 11)
    plan, inp = _case(L, oracle, small, 1, variant, 12)

    # a table beyond the fixed array: refused before any launch, delta_v untouched
    T_big = NOALLOC_DOUBLES // len(small.species) + 1  # > the array at any per-atom count >= 1
    dv = torch.full((small.n_full_points,), 3.0, dtype=torch.float64, device="cuda")
    rc = _call(L.frpa_far_forward_noalloc, plan, inp, T_big, 1, variant, dv)
    msg = L.frpa_last_error().decode()
    print(f"[noalloc] T = {T_big}: status {rc}: {msg}", flush=True)
    assert rc == 3 and "scratch" in msg  # FRPA_ERR_UNSUPPORTED
    torch.cuda.synchronize()
    assert bool(torch.all(dv == 3.0))
