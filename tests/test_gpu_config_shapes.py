"""rho1 and its adjoint against the CPU FP64 oracle at the batch shapes of the two large workloads
(evaluate_first_order_rho.cpp:113-209 and evaluate_first_order_H.cpp:138-215 restated in oracle/):

* 1024 atoms (14 336 basis functions), batches with n_compute ~ U{100..1200} (up to 19 row tiles of
  64) and n_points ~ U{50..200}, P1 = N(0,1) exp(-d/2 bohr) cut at 1e-8; 384 batches centred around
  eight neighbouring atoms (the density of the 60 000-batch grid, tens of batches per atom, so that the
  INT8 adjoint merges neighbouring batches into groups).  The wave is placed 2^31 + 4096 doubles into
  a device buffer, so every batch offset needs 64 bits.  FP64 forward and adjoint within 1e-12 of
  max|.|, the INT8 split (5 digits) within 1e-10, the INT8 adjoint with merged groups (threshold 0.8)
  and with every batch alone; the FP32-MFMA / BF16x3 forwards within 1e-6 / 5e-5, the FP64
  symmetric-half forward within 1e-12.
* 512 atoms (7 168 basis functions), n_compute ~ U{80..800}: every precision of the forward against
  the oracle, with the maximum, median and condition-normalised errors printed (the last divides
  |d rho1(p)| by sum_ij |phi_i(p)| |P1_ij| |phi_j(p)|, the oracle run on |phi| and |P1|).
"""
import numpy as np
import pytest
import torch

from fast_rpa_eval_amd import api, synth

pytestmark = pytest.mark.gpu

OFFSET_BASE = 2 ** 31 + 4096  # doubles in front of the first batch block


def _device_p1(prob, seed, cut):
    """Symmetric N(0,1) exp(-d_ij / 2 bohr), entries below `cut` zeroed, generated on the device
    (column-major flat tensor) and its host copy (n_basis, n_basis, 1) in Fortran order."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    nb = prob.n_basis
    A = torch.randn(nb, nb, generator=g, device="cuda", dtype=torch.float64)
    A = torch.triu(A) + torch.triu(A, 1).T
    bx = torch.from_numpy(prob.atom_xyz[prob.basis.atom]).cuda()
    A *= torch.exp(-torch.cdist(bx, bx) / 2.0)
    if cut > 0:
        A[A.abs() < cut] = 0.0
    P1 = A.T.contiguous().view(-1)  # element (r, c) at r + nb c
    host = np.asfortranarray(A.cpu().numpy())[:, :, None]
    return P1, host


def _problem(n_atoms, box_ang, n_batches, np_range, nc_range, k, n_hosts, base):
    """Batches around the n_hosts atoms nearest the geometry's centre; wave generated on the device
    with every batch block `base` doubles further into the buffer; returns (prob with its host wave,
    plan, device wave buffer)."""
    xyz = synth.make_geometry(n_atoms, box_ang * synth.ANG, 1.5 * synth.ANG, synth.seeds(k)["geometry"])
    hosts = np.argsort(np.sum((xyz - xyz.mean(axis=0)) ** 2, axis=1), kind="stable")[:n_hosts]
    prob = synth.make_rho1_problem(n_atoms, box_ang, n_batches, np_range, nc_range, config_k=k, fill_wave=False,
                                   host_atoms=hosts)
    plan = api.Rho1Plan(prob.n_basis, prob.n_full_points, prob.n_max_batch_size, prob.n_point_batches,
                        prob.n_point_batches_prefix_sum, prob.i_valid_point_2_i_full_points_map,
                        prob.n_compute_c_batches, prob.ibasis_ptr, prob.ibasis, prob.wave_offset + base)
    wave = plan.synth_wave(prob.atom_xyz, prob.basis, api.to_device(prob.coord_points),
                           base + int(prob.wave_offset[-1]))
    prob.wave = wave[base:].cpu().numpy()
    print(f"[shape] {n_atoms} atoms, {prob.n_batches} batches, n_compute {prob.n_compute_c_batches.min()}.."
          f"{prob.n_compute_c_batches.max()}, points {int(prob.n_point_batches.sum())}, wave at +{base} doubles "
          f"({(base + int(prob.wave_offset[-1])) * 8 / 1e9:.1f} GB buffer)", flush=True)
    return prob, plan, wave


def _oracle_h1(oracle, prob, g_full):
    nmb, nb, nf = prob.n_max_batch_size, prob.n_batches, prob.n_full_points
    z = lambda *s: np.zeros(s, order="F")
    a = dict(vrho=z(nmb * nb), vsigma=z(3 * nmb * nb), gradient_rho=z(3, nf),
             gwave=np.zeros(3 * int(prob.wave_offset[-1])), hwave=np.zeros(6 * int(prob.wave_offset[-1])),
             H_prefactor1=z(nmb, nb, 1), grad_rho1=z(3, nmb, nb, 1), v_hartree_gradient=np.asfortranarray(g_full),
             first_order_rho=z(nf, 1), v2rho2=z(3 * nmb * nb), v2rhosigma=z(6 * nmb * nb), v2sigma2=z(6 * nmb * nb),
             partition_tab=np.ones(nf))
    return oracle.h1(prob, 1, 1, 1, 1, np.zeros((prob.n_basis, prob.n_basis, 1), order="F"), a)[:, :, 0]


@pytest.fixture(scope="module")
def shape1024(cuda, oracle):
    oracle.set_threads(16)
    prob, plan, wave = _problem(1024, 50.0, 384, (50, 200), (100, 1200), 5, 8, OFFSET_BASE)
    P1, P1h = _device_p1(prob, synth.seeds(5)["p1"], 1e-8)
    rho_ref = oracle.rho1(prob, P1h)[0][:, 0]
    g_host = np.random.default_rng(synth.seeds(5)["v1"]).standard_normal((prob.n_full_points, 1))
    H_ref = _oracle_h1(oracle, prob, g_host)
    yield dict(prob=prob, plan=plan, wave=wave, P1=P1, rho_ref=rho_ref, g=api.to_device(g_host), H_ref=H_ref)
    del plan, wave
    torch.cuda.empty_cache()


@pytest.mark.parametrize("precision,tol", [(api.FP64, 1e-12), (api.OZAKI, 1e-10), (api.FP32, 1e-6),
                                           (api.BF16X3, 5e-5), (api.FP64_SYM, 1e-12)])
def test_rho1_1024atom_shapes_64bit_offsets(shape1024, precision, tol):
    D = shape1024
    got = D["plan"].forward(D["wave"], D["P1"], precision=precision).cpu().numpy()
    ref = D["rho_ref"]
    valid = D["prob"].valid_map()
    err = np.abs(got[valid] - ref[valid]).max() / np.abs(ref).max()
    print(f"[rho1 1024 atoms precision {precision}] max|err|/max|rho1| {err:.3e}", flush=True)
    assert np.all(np.isfinite(got[valid])) and err <= tol


def test_rho1_gradient_1024atom_shapes(shape1024, oracle):
    """The GGA gradient at the same shapes: FP64 within 1e-12, the INT8 split (every tile pair of P1)
    within 1e-10, against the oracle; the wave buffer is passed from its first batch block (32-bit
    offsets) so that the gradient wave (three times the wave) stays a few GB."""
    D = shape1024
    prob = D["prob"]
    gwave_h = synth.random_like_wave(prob, 3, 77)
    rho_ref, g_ref = oracle.rho1(prob, np.asfortranarray(D["P1"].view(prob.n_basis, prob.n_basis).cpu().numpy().T)
                                 [:, :, None], gwave=gwave_h)
    plan = api.Rho1Plan.from_problem(prob)
    wave = D["wave"][OFFSET_BASE:]
    gw = api.to_device(gwave_h)
    gb = torch.zeros(3 * prob.n_max_batch_size * prob.n_batches, dtype=torch.float64, device="cuda")
    for precision, tol in ((api.FP64, 1e-12), (api.OZAKI, 1e-10)):
        rho, grad = plan.forward(wave, D["P1"], precision=precision, gradient_wave=gw, grad_bias=gb)
        rho = rho.cpu().numpy()
        grad = grad.cpu().numpy().reshape(g_ref.shape, order="F")
        e1 = np.abs(rho - rho_ref[:, 0]).max() / np.abs(rho_ref).max()
        e2 = np.abs(grad - g_ref).max() / np.abs(g_ref).max()
        print(f"[rho1 + gradient 1024 atoms precision {precision}] rho1 rel {e1:.3e}, gradient rel {e2:.3e}",
              flush=True)
        assert e1 <= tol and e2 <= tol


@pytest.mark.parametrize("precision,merge,tol", [(api.FP64, None, 1e-12), (api.OZAKI, "0.8", 1e-10),
                                                 (api.OZAKI, "0", 1e-10)])
def test_adjoint_1024atom_shapes_64bit_offsets(shape1024, precision, merge, tol, monkeypatch):
    D = shape1024
    if merge is not None:
        monkeypatch.setenv("FRPA_ADJOINT_MERGE", merge)
    H = D["plan"].adjoint(D["wave"], D["g"], precision=precision)  # raises if a slot wait gave up
    nb = D["prob"].n_basis
    got = H.cpu().numpy().reshape((nb, nb), order="F")
    err = np.abs(got - D["H_ref"]).max() / np.abs(D["H_ref"]).max()
    print(f"[adjoint 1024 atoms precision {precision} merge {merge}] max|err|/max|H1| {err:.3e}", flush=True)
    assert err <= tol
    if precision == api.OZAKI:
        D["plan"].adjoint_status()


def test_rho1_512atom_shapes_all_precisions(cuda, oracle):
    oracle.set_threads(16)
    prob, plan, wave = _problem(512, 40.0, 256, (50, 200), (80, 800), 3, 8, 0)
    P1, P1h = _device_p1(prob, synth.seeds(3)["p1"], 0.0)
    ref = oracle.rho1(prob, P1h)[0][:, 0]
    absprob = synth.Rho1Problem(**{**prob.__dict__})
    absprob.wave = np.abs(prob.wave)
    cond = oracle.rho1(absprob, np.abs(P1h))[0][:, 0]
    valid = prob.valid_map()
    big = np.abs(ref[valid]) > 1e-3 * np.abs(ref).max()
    for name, prec, tol in (("fp64", api.FP64, 1e-12), ("int8_split_s5", api.OZAKI, 1e-10),
                            ("int8_split_s4", api.OZAKI4, 1e-8), ("fp32_mfma", api.FP32, 1e-5),
                            ("bf16x3", api.BF16X3, 1e-3)):
        got = plan.forward(wave, P1, precision=prec).cpu().numpy()
        d = np.abs(got[valid] - ref[valid])
        mx = d.max() / np.abs(ref).max()
        med = np.median(d[big] / np.abs(ref[valid][big]))
        cn = (d / cond[valid]).max()
        print(f"[rho1 512 atoms {name}] max rel {mx:.3e}, median pointwise rel {med:.3e}, "
              f"max condition-normalised {cn:.3e}", flush=True)
        assert np.all(np.isfinite(got[valid])) and mx <= tol


def test_rho1_1024atom_row_tile_items(shape1024):
    """Launch order 3 (row-tile ranges of point passes as work items, up to 19 row tiles per batch,
    per-tile sums finished by a second kernel) gives the bits of the default launch, with and
    without the gradient, at the 1024-atom shapes."""
    D = shape1024
    prob, plan = D["prob"], D["plan"]
    wave = D["wave"]
    gw = api.to_device(synth.random_like_wave(prob, 3, 78))
    gb = torch.zeros(3 * prob.n_max_batch_size * prob.n_batches, dtype=torch.float64, device="cuda")
    w0 = wave[OFFSET_BASE:]
    plan0 = api.Rho1Plan.from_problem(prob)
    ref = plan0.forward(w0, D["P1"], precision=api.OZAKI).clone()
    gref = [x.clone() for x in plan0.forward(w0, D["P1"], precision=api.OZAKI, gradient_wave=gw, grad_bias=gb)]
    for persistent, npg in ((0, 2), (1, 2), (1, 1)):
        plan0.set_launch(api.OZAKI, 3, persistent, npg)
        got = plan0.forward(w0, D["P1"], precision=api.OZAKI)
        r2, g2 = plan0.forward(w0, D["P1"], precision=api.OZAKI, gradient_wave=gw, grad_bias=gb)
        print(f"[rho1 1024 atoms order 3 persistent {persistent} npg {npg}] identical {torch.equal(got, ref)}, "
              f"gradient identical {torch.equal(r2, gref[0]) and torch.equal(g2, gref[1])}", flush=True)
        assert torch.equal(got, ref)
        assert torch.equal(r2, gref[0]) and torch.equal(g2, gref[1])
    valid = prob.valid_map()
    err = np.abs(ref.cpu().numpy()[valid] - D["rho_ref"][valid]).max() / np.abs(D["rho_ref"]).max()
    assert err <= 1e-10
