"""Adjoint of rho1 w.r.t. P1 on the GPU (the LDA first-order H integration) against the CPU oracle
of evaluate_first_order_H (fast_rap_src_dfpt_gpu_eval/evaluate_first_order_H.cpp:13-222) and the
adjoint identity <g, rho1(P)> = <H1(g), P>_F."""
import numpy as np
import pytest
import torch

from fast_rpa_eval_amd import api, synth

pytestmark = pytest.mark.gpu


def _oracle_h1(oracle, prob, g_full, T):
    """LDA H1 through the reference's argument list: partition_tab = 1, v_hartree_gradient = g."""
    nmb, nb, nf = prob.n_max_batch_size, prob.n_batches, prob.n_full_points
    z = lambda *s: np.zeros(s, order="F")
    a = dict(vrho=z(nmb * nb), vsigma=z(3 * nmb * nb), gradient_rho=z(3, nf),
             gwave=np.zeros(3 * int(prob.wave_offset[-1])), hwave=np.zeros(6 * int(prob.wave_offset[-1])),
             H_prefactor1=z(nmb, nb, T), grad_rho1=z(3, nmb, nb, T), v_hartree_gradient=np.asfortranarray(g_full),
             first_order_rho=z(nf, T), v2rho2=z(3 * nmb * nb), v2rhosigma=z(6 * nmb * nb), v2sigma2=z(6 * nmb * nb),
             partition_tab=np.ones(nf))
    H = np.zeros((prob.n_basis, prob.n_basis, T), order="F")
    return oracle.h1(prob, T, 1, T, 1, H, a)


@pytest.mark.parametrize("mode", ["atomics", "tiles"])
@pytest.mark.parametrize("k,args", [(1, (8, 10.0, 64, 100, None)), (3, (40, 16.0, 120, (1, 220), (1, 260)))])
def test_h1_lda_against_oracle(cuda, oracle, k, args, mode, monkeypatch):
    """Both adjoint modes: per-batch tiles + FP64 atomics (default) and deterministic tile owners."""
    if mode == "tiles":
        monkeypatch.setenv("FRPA_ADJOINT", "tiles")
    else:
        monkeypatch.delenv("FRPA_ADJOINT", raising=False)
    prob = synth.make_rho1_problem(*args, config_k=k, zero_nc=(2,) if k == 3 else ())
    T = 2
    rng = np.random.default_rng(k)
    g = np.asfortranarray(rng.standard_normal((prob.n_full_points, T)))
    ref = _oracle_h1(oracle, prob, g, T)
    plan = api.Rho1Plan.from_problem(prob)
    H0 = rng.standard_normal(prob.n_basis * prob.n_basis * T)
    Hd = api.to_device(H0)  # accumulated into
    plan.adjoint(api.to_device(prob.wave), api.to_device(g), T=T, H=Hd)
    got = Hd.cpu().numpy().reshape((prob.n_basis, prob.n_basis, T), order="F") - H0.reshape(
        (prob.n_basis, prob.n_basis, T), order="F")
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"[h1 lda config {k} {mode}] rel {err:.3e}", flush=True)
    assert err <= 1e-12


def test_h1_adjoint_identity_with_gpu_forward(cuda):
    prob = synth.make_rho1_problem(30, 15.0, 100, (20, 200), (20, 300), config_k=11)
    rng = np.random.default_rng(0)
    plan = api.Rho1Plan.from_problem(prob)
    wave = api.to_device(prob.wave)
    P = rng.standard_normal((prob.n_basis, prob.n_basis))
    g = rng.standard_normal(prob.n_full_points)
    rho = plan.forward(wave, api.to_device(np.asfortranarray(P[:, :, None]))).cpu().numpy()
    H = plan.adjoint(wave, api.to_device(g)).cpu().numpy().reshape((prob.n_basis, prob.n_basis), order="F")
    lhs, rhs = float(np.dot(g, rho)), float(np.sum(H * P))
    print(f"[adjoint identity] {lhs:.15e} {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs))


@pytest.mark.parametrize("precision,tol", [(api.OZAKI, 1e-10), (api.OZAKI4, 1e-8)])
@pytest.mark.parametrize("k,args", [(1, (8, 10.0, 64, 100, None)), (3, (40, 16.0, 120, (1, 300), (1, 260))),
                                    (4, (16, 12.0, 24, (330, 700), (50, 200)))])
def test_h1_int8_split_against_oracle(cuda, oracle, k, args, precision, tol):
    """INT8-split adjoint (per-row digit scales, point index contracted on INT8 MFMA): max|dH| / max|H|
    within the stated bound of the FP64 oracle; T = 2, batches of one point pass (at most 320 points)
    and, in the third case, of two and three passes; nc = 0."""
    prob = synth.make_rho1_problem(*args, config_k=k, zero_nc=(2,) if k == 3 else ())
    T = 2
    rng = np.random.default_rng(10 + k)
    g = np.asfortranarray(rng.standard_normal((prob.n_full_points, T)))
    ref = _oracle_h1(oracle, prob, g, T)
    plan = api.Rho1Plan.from_problem(prob)
    H0 = rng.standard_normal(prob.n_basis * prob.n_basis * T)
    Hd = api.to_device(H0)
    wave = api.to_device(prob.wave)
    plan.adjoint(wave, api.to_device(g), T=T, H=Hd, precision=precision)
    got = Hd.cpu().numpy().reshape((prob.n_basis, prob.n_basis, T), order="F") - H0.reshape(
        (prob.n_basis, prob.n_basis, T), order="F")
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"[h1 int8 split config {k} precision {precision}] rel {err:.3e}", flush=True)
    assert err <= tol
    # a second call with another g reuses the phi digits
    g2 = np.asfortranarray(rng.standard_normal((prob.n_full_points, T)))
    H2 = plan.adjoint(wave, api.to_device(g2), T=T, precision=precision).cpu().numpy().reshape(
        (prob.n_basis, prob.n_basis, T), order="F")
    ref2 = _oracle_h1(oracle, prob, g2, T)
    assert np.abs(H2 - ref2).max() <= tol * np.abs(ref2).max()


def test_h1_int8_split_more_than_1280_rows(cuda, oracle):
    """Batches of 1316 to 1444 compute rows (21 to 23 row tiles, 88 to 92 producer units in the
    first row tile) against the oracle."""
    prob = synth.make_rho1_problem(110, 22.0, 6, (30, 90), (1290, 1450), config_k=8)
    assert prob.n_compute_c_batches.min() > 1280
    rng = np.random.default_rng(29)
    g = np.asfortranarray(rng.standard_normal((prob.n_full_points, 1)))
    ref = _oracle_h1(oracle, prob, g, 1)
    plan = api.Rho1Plan.from_problem(prob)
    H = plan.adjoint(api.to_device(prob.wave), api.to_device(g), T=1, precision=api.OZAKI).cpu().numpy().reshape(
        (prob.n_basis, prob.n_basis, 1), order="F")
    err = np.abs(H - ref).max() / np.abs(ref).max()
    print(f"[h1 int8 split nc > 1280] rel {err:.3e}", flush=True)
    assert err <= 1e-10


@pytest.mark.parametrize("precision", [api.OZAKI, api.OZAKI4])
def test_h1_int8_split_is_bit_reproducible(cuda, precision):
    """The INT8-split adjoint accumulates H in 64-bit fixed point with integer atomics, so repeated
    calls give the same bits whatever order the atomics land in (T = 2, many batches sharing each
    H element); a g scaled by 2 gives exactly 2 H (the scale follows the bound by a power of two)."""
    prob = synth.make_rho1_problem(40, 16.0, 160, (1, 260), (1, 260), config_k=5)
    T = 2
    g = api.to_device(np.asfortranarray(np.random.default_rng(3).standard_normal((prob.n_full_points, T))))
    plan = api.Rho1Plan.from_problem(prob)
    wave = api.to_device(prob.wave)
    H = [plan.adjoint(wave, g, T=T, precision=precision).clone() for _ in range(3)]
    assert torch.count_nonzero(H[0]) > 0
    assert torch.equal(H[0], H[1]) and torch.equal(H[0], H[2])
    H2 = plan.adjoint(wave, 2.0 * g, T=T, precision=precision)
    assert torch.equal(H2, 2.0 * H[0])


def test_h1_int8_split_producer_consumer_matches_direct(cuda, monkeypatch):
    """The default INT8 adjoint hands each tile-pair block from the multiplying waves to atomic waves
    through LDS slots; with FRPA_ADJOINT_PC=0 every wave issues its own atomics.  Both add the same
    int64 values, so H is the same bits; every slot wait completes (each call checks the count)."""
    prob = synth.make_rho1_problem(40, 16.0, 160, (1, 600), (1, 300), config_k=6, zero_nc=(3,))
    T = 2
    g = api.to_device(np.asfortranarray(np.random.default_rng(4).standard_normal((prob.n_full_points, T))))
    plan = api.Rho1Plan.from_problem(prob)
    wave = api.to_device(prob.wave)
    H_pc = plan.adjoint(wave, g, T=T, precision=api.OZAKI).clone()
    monkeypatch.setenv("FRPA_ADJOINT_PC", "0")
    H_direct = plan.adjoint(wave, g, T=T, precision=api.OZAKI)
    assert torch.count_nonzero(H_pc) > 0
    assert torch.equal(H_pc, H_direct)


def test_h1_int8_split_merged_groups_against_oracle(cuda, oracle, monkeypatch):
    """The INT8 adjoint merges consecutive batches with contiguous points into one work item over the
    union of their compute sets (FRPA_ADJOINT_MERGE: 0 = every batch alone, 0.8 = the default,
    4 = whenever the points fit one 256-point pass).  Every threshold agrees with the FP64 oracle; a
    batch with n_compute_c = 0 (its points break the contiguity) and one with no points are in the
    list.  Merged groups have other digit scales, so H differs in the last bits from the unmerged one,
    which shows the groups were formed."""
    prob = synth.make_rho1_problem(40, 16.0, 200, (20, 120), (40, 300), config_k=7, zero_nc=(5,), zero_np=(9,))
    T = 2
    g_host = np.asfortranarray(np.random.default_rng(8).standard_normal((prob.n_full_points, T)))
    ref = _oracle_h1(oracle, prob, g_host, T)
    plan = api.Rho1Plan.from_problem(prob)
    wave, g = api.to_device(prob.wave), api.to_device(g_host)
    out = {}
    for m in ("0", "0.8", "4"):
        monkeypatch.setenv("FRPA_ADJOINT_MERGE", m)
        H = plan.adjoint(wave, g, T=T, precision=api.OZAKI).clone()
        got = H.cpu().numpy().reshape((prob.n_basis, prob.n_basis, T), order="F")
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"[h1 int8 split merge {m}] rel {err:.3e}", flush=True)
        assert err <= 1e-10
        out[m] = H
    assert not torch.equal(out["0"], out["4"])


def test_h1_int8_split_wait_failures_are_reported(cuda, oracle):
    """Slot waits of the INT8 adjoint that give up (forced from the host with a bound of one poll,
    so that the atomic waves stop waiting for blocks still being multiplied) are counted on the
    device: the call returns an error and leaves H untouched, every kernel ends normally, and with the
    default bound restored the same plan gives the oracle's H again.  The same count under graph
    capture is read by frpa_rho1_adjoint_status."""
    prob = synth.make_rho1_problem(40, 16.0, 160, (1, 300), (100, 400), config_k=6)
    T = 1
    g_host = np.asfortranarray(np.random.default_rng(5).standard_normal((prob.n_full_points, T)))
    ref = _oracle_h1(oracle, prob, g_host, T)
    plan = api.Rho1Plan.from_problem(prob)
    wave, g = api.to_device(prob.wave), api.to_device(g_host)
    plan.set_adjoint_wait_limit(1)
    H = torch.full((prob.n_basis * prob.n_basis * T,), 7.0, dtype=torch.float64, device="cuda")
    with pytest.raises(api.FrpaError, match="slot waits gave up"):
        plan.adjoint(wave, g, T=T, precision=api.OZAKI, H=H)
    torch.cuda.synchronize()
    assert bool(torch.all(H == 7.0)), "H changed by a call that reported failed waits"
    plan.adjoint_status()  # the failing call has read and cleared the count
    # under capture the call cannot check (its digits and groups were built by the call above): the
    # graph's replay leaves the count for adjoint_status
    stream = torch.cuda.Stream()
    Hg = torch.zeros(prob.n_basis * prob.n_basis * T, dtype=torch.float64, device="cuda")
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        plan.adjoint(wave, g, T=T, precision=api.OZAKI, H=Hg, stream=torch.cuda.current_stream())
    graph.replay()
    torch.cuda.synchronize()
    with pytest.raises(api.FrpaError, match="slot waits gave up"):
        plan.adjoint_status()
    plan.adjoint_status()  # cleared by the read
    plan.set_adjoint_wait_limit(0)
    H = plan.adjoint(wave, g, T=T, precision=api.OZAKI)
    got = H.cpu().numpy().reshape((prob.n_basis, prob.n_basis, T), order="F")
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"[h1 int8 split after failed waits] rel {err:.3e}", flush=True)
    assert err <= 1e-10
    plan.adjoint_status()


def test_h1_int8_split_wave_changed_in_place(cuda, oracle):
    """The INT8 adjoint keeps digits of the wave it first saw.  After the wave tensor is changed in
    place, frpa_rho1_prepare_wave rebuilds them (the C-ABI contract), and the host mirror does so by
    itself (it follows the tensor's version counter): both give the oracle's H for the new wave."""
    prob = synth.make_rho1_problem(30, 15.0, 80, (20, 150), (30, 200), config_k=9)
    T = 1
    rng = np.random.default_rng(6)
    g_host = np.asfortranarray(rng.standard_normal((prob.n_full_points, T)))
    plan = api.Rho1Plan.from_problem(prob)
    wave, g = api.to_device(prob.wave), api.to_device(g_host)
    plan.adjoint(wave, g, T=T, precision=api.OZAKI)  # digits of the first wave
    prob2 = synth.make_rho1_problem(30, 15.0, 80, (20, 150), (30, 200), config_k=9)
    prob2.wave = synth.random_like_wave(prob2, 1, 77)
    ref2 = _oracle_h1(oracle, prob2, g_host, T)
    # C-ABI: change in place, then frpa_rho1_prepare_wave (bypassing the mirror's own tracking)
    wave.copy_(api.to_device(prob2.wave))
    plan.prepare_wave(wave, api.OZAKI)
    H = torch.zeros(prob.n_basis * prob.n_basis * T, dtype=torch.float64, device="cuda")
    api._lib.check(api.lib().frpa_rho1_adjoint(plan.handle, T, api.OZAKI, api._dp(wave), api._dp(g), api._dp(H),
                                               api._stream()), "frpa_rho1_adjoint")
    got = H.cpu().numpy().reshape((prob.n_basis, prob.n_basis, T), order="F")
    err = np.abs(got - ref2).max() / np.abs(ref2).max()
    print(f"[h1 int8 split, wave changed + prepare_wave] rel {err:.3e}", flush=True)
    assert err <= 1e-10
    # host mirror: change in place again (back to the first wave), no explicit call
    ref1 = _oracle_h1(oracle, prob, g_host, T)
    wave.copy_(api.to_device(prob.wave))
    got = plan.adjoint(wave, g, T=T, precision=api.OZAKI).cpu().numpy().reshape(
        (prob.n_basis, prob.n_basis, T), order="F")
    err = np.abs(got - ref1).max() / np.abs(ref1).max()
    print(f"[h1 int8 split, wave changed, mirror refresh] rel {err:.3e}", flush=True)
    assert err <= 1e-10
