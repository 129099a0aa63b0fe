#!/usr/bin/env python
"""Benchmark: rho1 + far-field Hartree grid points/s on one MI355X, fast path vs plain FP64.

A step is one pass of the hot path over the synthetic workload, with every input resident in HBM:
  rho1 forward over all batches (evaluate_first_order_rho), its adjoint (the LDA first-order H
  integration, H1 += sum_p g phi_i phi_j), and the far-field Hartree sum over the far-field points.
Two paths are timed on the same inputs:
  fast  : rho1/adjoint with the INT8-split GEMM (FP64 recombination), far field by merged-coefficient
          Horner (FP64);
  fp64  : rho1/adjoint with FP64 MFMA, far field term by term (FP64) -- the plain FP64 GPU kernels.
value = (rho1 points + far-field points) / fast step time.  The JSON line also carries the FP64
figure, the fast/FP64 ratio, the fast path's error against the FP64 path, the dominant kernel's
roofline (timed with HIP events on its stream) and a CPU baseline (the CPU FP64 oracle on a bounded
sample of the same inputs).

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu]
"""
import argparse
import copy
import json
import re
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # 1024 atoms (14336 basis functions), 60k batches (50-200 points, n_compute 100-1200), rho1 forward
    # + adjoint, far field at l_max = 8 over 6M points; P1 cut at 1e-8.
    "combined_1024atoms_60kbatches_6Mpts_l8": dict(
        n_atoms=1024, box_ang=50.0, n_batches=60000, np_range=(50, 200), nc_range=(100, 1200),
        n_far=6_000_000, l_far=8, config_k=5, p1_cut=1e-8),
    # 512 atoms (7168 basis functions), 20k batches (50-200 points, n_compute 80-800); far field at
    # l_max = 6 over 200k points.
    "rho1_512atoms_20kbatches_far200k_l6": dict(
        n_atoms=512, box_ang=40.0, n_batches=20000, np_range=(50, 200), nc_range=(80, 800),
        n_far=200_000, l_far=6, config_k=3, p1_cut=0.0),
    # small smoke-size workload
    "small_64atoms_2kbatches_far50k_l6": dict(
        n_atoms=64, box_ang=20.0, n_batches=2000, np_range=(50, 200), nc_range=(80, 400),
        n_far=50_000, l_far=6, config_k=4, p1_cut=0.0),
}
DEFAULT = "combined_1024atoms_60kbatches_6Mpts_l8"

PEAK_FP64 = 78.6      # TFLOP/s, dense FP64 MFMA / vector
PEAK_FP32 = 157.3     # TFLOP/s, dense FP32 MFMA
PEAK_BF16 = 2500.0    # TFLOP/s, dense BF16 MFMA
PEAK_INT8 = 5000.0    # TOP/s, dense INT8 MFMA
PEAK_HBM = 8000.0     # GB/s, HBM3E


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def make_inputs(w):
    import torch

    from fast_rpa_eval_amd import api, synth
    t0 = time.time()
    prob = synth.make_rho1_problem(w["n_atoms"], w["box_ang"], w["n_batches"], w["np_range"], w["nc_range"],
                                   config_k=w["config_k"], fill_wave=False)
    log(f"batches built: {prob.n_batches} batches, {prob.n_basis} basis, {int(prob.n_point_batches.sum())} points, "
        f"wave {prob.wave_offset[-1] * 8 / 1e9:.1f} GB ({time.time() - t0:.1f}s)")
    plan = api.Rho1Plan.from_problem(prob)
    cp = api.to_device(prob.coord_points)
    wave = plan.synth_wave(prob.atom_xyz, prob.basis, cp, int(prob.wave_offset[-1]))
    # P1 symmetric N(0,1) exp(-d_ij / 2 bohr), generated on the device
    g = torch.Generator(device="cuda").manual_seed(synth.seeds(w["config_k"])["p1"])
    nb = prob.n_basis
    A = torch.randn(nb, nb, generator=g, device="cuda", dtype=torch.float64)
    A = torch.triu(A) + torch.triu(A, 1).T
    bx = torch.from_numpy(prob.atom_xyz[prob.basis.atom]).cuda()
    A *= torch.exp(-torch.cdist(bx, bx) / 2.0)
    if w["p1_cut"] > 0:
        A[A.abs() < w["p1_cut"]] = 0.0
    P1 = A.contiguous().view(-1)
    del A, bx
    rng = np.random.default_rng(synth.seeds(w["config_k"])["v1"])
    gfull = api.to_device(rng.standard_normal(prob.n_full_points))  # partition_tab * v1 with w = 1
    log(f"device inputs ready ({time.time() - t0:.1f}s)")
    fp = synth.make_far_problem(w["n_atoms"], w["box_ang"], w["n_far"], w["l_far"], config_k=w["config_k"],
                                all_far=False, xyz=prob.atom_xyz)
    fplan = api.FarPlan.from_problem(fp)
    finp = api.FarInputs(fp)
    log(f"far-field inputs ready: {fp.n_valid_points} points, {fp.n_atoms} centres, l={fp.l_pot_max} "
        f"({time.time() - t0:.1f}s)")
    torch.cuda.synchronize()
    return dict(prob=prob, plan=plan, wave=wave, P1=P1, g=gfull, fp=fp, fplan=fplan, finp=finp)


NEAR_GRID = 200        # log-grid points per species of the near-field spline tile
NEAR_RADIUS = 8.0      # multipole radius (bohr): pairs closer than this take the spline branch


def make_near_inputs(D, w):
    """Near-field branch of sum_up_whole_potential over the rho1 integration grid (its batch-ordered
    points), two species with their own log grids, l_hartree = l_far / l_far - 2, smooth radial tiles."""
    from fast_rpa_eval_amd import api, synth
    t0 = time.time()
    prob = D["prob"]
    fpn = synth.make_far_problem(w["n_atoms"], w["box_ang"], 0, w["l_far"], config_k=w["config_k"],
                                 xyz=prob.atom_xyz, min_dist_bohr=NEAR_RADIUS,
                                 grid=(prob.coord_points, prob.valid_map()))
    synth.add_near_field(fpn, NEAR_GRID, "smooth", n_species=2, seed=w["config_k"])
    plan = api.FarPlan.from_problem(fpn)
    inp = api.FarInputs(fpn)
    log(f"near-field inputs ready: {fpn.n_valid_points} points, tile "
        f"{fpn.near['spl'].nbytes / 1e9:.2f} GB ({time.time() - t0:.1f}s)")
    return dict(fp=fpn, plan=plan, inp=inp)


def near_pairs(fpn):
    """(number of near pairs, sum over them of (l+1)^2) for the near-field inputs (device count)."""
    import torch
    pts = torch.from_numpy(fpn.coord_points[:, fpn.i_valid_point_2_i_full_points_map].T.copy()).cuda()
    cen = torch.from_numpy(fpn.coords_center.T.copy()).cuda()
    lh = torch.from_numpy(fpn.l_hartree[fpn.species - 1].astype(np.int64)).cuda()
    opr = torch.from_numpy(np.ascontiguousarray(fpn.outer_potential_radius[:, :, 0].T)).cuda()  # (atoms, l)
    mr = torch.from_numpy(fpn.multipole_radius_sq[:, 0].copy()).cuda()
    n, nlm = 0, 0.0
    for s in range(0, pts.shape[0], 200000):
        d2 = torch.cdist(pts[s:s + 200000], cen) ** 2
        near = d2 < mr[None, :]
        l = lh[None, :].expand_as(d2).clone()
        for _ in range(int(lh.max())):
            low = (torch.gather(opr, 1, l.T).T < d2) & (l > 0)
            l = l - low.long()
        n += int(near.sum())
        nlm += float((((l + 1) ** 2) * near).sum())
    return n, nlm


def bench_bias2(D, steps, warmup, no_cpu, threads, T=4, jc=1):
    """generate_first_order_rho_bias_part2 for the tile of atoms 1..T (the reference's fixed tile of
    4), j_coord = 1, over all batches: random gradient wave (3x the wave, on the device), P1 as the
    density matrix.  HIP events on torch's current stream."""
    import torch
    prob, plan = D["prob"], D["plan"]
    wl = int(prob.wave_offset[-1])
    g = torch.Generator(device="cuda").manual_seed(17)
    gwave = torch.empty(3 * wl, dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0, generator=g)
    avn_h = prob.atom_valid_n_compute_c_batches()
    avn = torch.from_numpy(np.ascontiguousarray(avn_h.ravel(order="F"))).cuda()
    out = torch.zeros(prob.n_max_batch_size * prob.n_batches * T, dtype=torch.float64, device="cuda")

    def run():
        plan.bias_part2(D["wave"], gwave, D["P1"], avn, prob.n_atoms, 1, n_tile=T, j_coord=jc, T=T, out=out)
    for _ in range(max(warmup, 1)):
        run()
    torch.cuda.synchronize()
    ev = []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        run()
        e1.record()
        ev.append((e0, e1))
    torch.cuda.synchronize()
    ms = float(np.mean([x.elapsed_time(y) for x, y in ev]))
    # algorithmic work of the batches with rows of the tile's atoms
    nc = prob.n_compute_c_batches.astype(np.float64)
    npt = prob.n_point_batches.astype(np.float64)
    nR = (avn_h[T, :] - avn_h[0, :]).astype(np.float64)
    sel = (nc > 0) & (npt > 0) & (nR > 0)
    nbytes = float(np.sum((8 * nc * npt + 16 * nR * nc + 8 * nR * npt)[sel]) + np.sum(8 * T * npt[(nc > 0)]))
    flops = float(np.sum((2 * nR * nc * npt)[sel]))
    ach = nbytes / (ms / 1e3) / 1e9
    res = {"workload": f"tile of atoms 1-{T}, j_coord {jc}, all {prob.n_batches} batches "
                       f"({int(sel.sum())} with rows of the tile's atoms)",
           "ms": ms, "batches_with_rows": int(sel.sum()),
           "roofline": {"kernel": "frpa::bias2_kernel (with its per-call DM + DM^T table)", "bound": "hbm",
                        "achieved": ach, "peak": PEAK_HBM,
                        "unit": "GB/s", "frac": ach / PEAK_HBM, "algorithmic_bytes": nbytes,
                        "algorithmic_flops": flops, "fp64_frac": flops / (ms / 1e3) / 1e12 / PEAK_FP64}}
    if not no_cpu:
        from oracle import oracle
        rng = np.random.default_rng(5)
        pick = np.sort(rng.choice(np.nonzero(sel)[0], size=min(300, int(sel.sum())), replace=False))
        sub = subproblem(prob, pick, D["wave"])
        # gradient wave of the sub-problem: the same batches' blocks
        gw_h = np.concatenate([gwave[3 * int(prob.wave_offset[b]):3 * int(prob.wave_offset[b + 1])].cpu().numpy()
                               for b in pick])
        DM = D["P1"].view(prob.n_basis, prob.n_basis).T.cpu().numpy()  # column-major view of the same matrix
        oracle.set_threads(threads)
        t0 = time.time()
        ref = oracle.bias2(sub, np.asfortranarray(DM), gw_h, T, 1, j_coord=jc)
        cpu_s = time.time() - t0
        got = out.view(T, prob.n_batches, prob.n_max_batch_size).cpu().numpy()
        err = 0.0
        for q, b in enumerate(pick):
            n = prob.n_point_batches[b]
            err = max(err, float(np.abs(got[:, b, :n].T - ref[:n, q, :]).max()))
        res["vs_oracle_rel"] = err / float(np.abs(ref).max())
        res["cpu_baseline"] = {"value": len(pick) / cpu_s, "unit": "batches/s", "cores": threads, "kind": "port",
                               "sample": f"{len(pick)} random batches with rows of the tile's atoms"}
    del gwave, avn, out
    return res


H1_WORKLOAD = "rho1_512atoms_20kbatches_far200k_l6"


def bench_h1_gga(steps, warmup, no_cpu, threads, jb=1, jc=1):
    """Full first-order H (LDA + GGA terms of evaluate_first_order_H) for atom jb, coordinate jc, on the
    512-atom, 20k-batch workload: wave synthesised on the device, random gradient and hessian waves and
    XC arrays.  HIP events on torch's current stream around the whole call (Hs fill, kernel, unrelabel)."""
    import torch

    from fast_rpa_eval_amd import api, synth
    w3 = WORKLOADS[H1_WORKLOAD]
    t0 = time.time()
    prob = synth.make_rho1_problem(w3["n_atoms"], w3["box_ang"], w3["n_batches"], w3["np_range"], w3["nc_range"],
                                   config_k=w3["config_k"], fill_wave=False)
    plan = api.Rho1Plan.from_problem(prob)
    wl = int(prob.wave_offset[-1])
    wave = plan.synth_wave(prob.atom_xyz, prob.basis, api.to_device(prob.coord_points), wl)
    g = torch.Generator(device="cuda").manual_seed(29)
    U = lambda n, lo=-1.0, hi=1.0: torch.empty(n, dtype=torch.float64, device="cuda").uniform_(lo, hi, generator=g)
    nmb, nb, nf = prob.n_max_batch_size, prob.n_batches, prob.n_full_points
    gwave, hwave = U(3 * wl), U(6 * wl)
    xc = dict(vsigma=U(3 * nmb * nb), v2rho2=U(3 * nmb * nb), v2rhosigma=U(6 * nmb * nb), v2sigma2=U(6 * nmb * nb),
              gradient_rho=U(3 * nf), H_prefactor1=U(nmb * nb), grad_rho1=U(3 * nmb * nb),
              v_hartree_gradient=U(nf), first_order_rho=U(nf), partition_tab=U(nf, 0.1, 1.0))
    avn_h = prob.atom_valid_n_compute_c_batches()
    avn = torch.from_numpy(np.ascontiguousarray(avn_h.ravel(order="F"))).cuda()
    H = torch.zeros(prob.n_basis * prob.n_basis, dtype=torch.float64, device="cuda")
    log(f"H1 GGA inputs ready: {nb} batches, {prob.n_basis} basis, waves {(1 + 3 + 6) * wl * 8 / 1e9:.1f} GB "
        f"({time.time() - t0:.1f}s)")

    def run():
        plan.h1(wave, gwave, hwave, xc, avn, prob.n_atoms, jb, n_tile=1, j_coord=jc, H=H)
    for _ in range(max(warmup, 1)):
        run()
    torch.cuda.synchronize()
    ev = []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        run()
        e1.record()
        ev.append((e0, e1))
    torch.cuda.synchronize()
    ms = float(np.mean([x.elapsed_time(y) for x, y in ev]))
    # algorithmic work per batch: the two products C phi^T + phi C^T on the upper triangle (4 nc(nc+1)/2 np flop)
    # plus C (8 flop per (row, point), 6 more on the atom's rows); bytes: phi, dphi (3), the atom's rows of the
    # hessian (3 of 6 components), ~100 B of XC data per point, the atomics' read-modify-write of the triangle
    nc, npt = _nc_np(prob)
    sel = (prob.n_compute_c_batches > 0) & (prob.n_point_batches > 0)
    nv = (avn_h[jb, :] - avn_h[jb - 1, :])[sel].astype(np.float64)
    tri = nc * (nc + 1) / 2
    flops = float(np.sum(4 * tri * npt + 8 * nc * npt + 6 * nv * npt))
    nbytes = float(np.sum(8 * nc * npt + 24 * nc * npt + 24 * nv * npt + 100 * npt + 16 * tri))
    ach = flops / (ms / 1e3) / 1e12
    res = {"workload": f"{H1_WORKLOAD}: {nb} batches, {prob.n_basis} basis functions, atom {jb}, j_coord {jc}",
           "ms": ms, "points_per_s": float(npt.sum()) / (ms / 1e3),
           "roofline": {"kernel": "frpa::h1_gga_kernel", "bound": "mfma", "achieved": ach, "peak": PEAK_FP64,
                        "unit": "TFLOP/s", "frac": ach / PEAK_FP64, "algorithmic_flops": flops,
                        "algorithmic_bytes": nbytes, "hbm_frac": nbytes / (ms / 1e3) / 1e9 / PEAK_HBM}}
    if not no_cpu:
        from oracle import oracle
        rng = np.random.default_rng(31)
        pick = np.sort(rng.choice(np.nonzero(sel)[0], size=min(200, int(sel.sum())), replace=False))
        sub = subproblem(prob, pick, wave)
        nmb2 = sub.n_max_batch_size
        blocks = lambda a, k: np.concatenate([a[k * int(prob.wave_offset[b]):k * int(prob.wave_offset[b + 1])]
                                              .cpu().numpy() for b in pick])
        per_batch = lambda a, k: np.concatenate([a[k * nmb * b:k * nmb * (b + 1)].cpu().numpy()
                                                 .reshape(k, nmb, order="F")[:, :nmb2].ravel(order="F") for b in pick])
        arr = dict(vrho=np.zeros(nmb2 * len(pick)), vsigma=per_batch(xc["vsigma"], 3),
                   v2rho2=per_batch(xc["v2rho2"], 3), v2rhosigma=per_batch(xc["v2rhosigma"], 6),
                   v2sigma2=per_batch(xc["v2sigma2"], 6),
                   gradient_rho=xc["gradient_rho"].cpu().numpy().reshape(3, nf, order="F"),
                   H_prefactor1=per_batch(xc["H_prefactor1"], 1).reshape(nmb2, len(pick), 1, order="F"),
                   grad_rho1=per_batch(xc["grad_rho1"], 3).reshape(3, nmb2, len(pick), 1, order="F"),
                   v_hartree_gradient=xc["v_hartree_gradient"].cpu().numpy().reshape(nf, 1),
                   first_order_rho=xc["first_order_rho"].cpu().numpy().reshape(nf, 1),
                   partition_tab=xc["partition_tab"].cpu().numpy(), gwave=blocks(gwave, 3), hwave=blocks(hwave, 6))
        oracle.set_threads(threads)
        t1 = time.time()
        ref = oracle.h1(sub, 1, jb, jb, jc, np.zeros((prob.n_basis, prob.n_basis, 1), order="F"), arr)
        cpu_s = time.time() - t1
        splan = api.Rho1Plan.from_problem(sub)
        dxc = {k: api.to_device(v) for k, v in arr.items() if k not in ("vrho", "gwave", "hwave")}
        Hs = splan.h1(api.to_device(sub.wave), api.to_device(arr["gwave"]), api.to_device(arr["hwave"]), dxc,
                      api.to_device(sub.atom_valid_n_compute_c_batches()), prob.n_atoms, jb, j_coord=jc)
        got = Hs.cpu().numpy().reshape(prob.n_basis, prob.n_basis, order="F")
        res["vs_oracle_rel"] = float(np.abs(got - ref[:, :, 0]).max() / max(np.abs(ref).max(), 1e-300))
        res["cpu_baseline"] = {"value": float(sub.n_point_batches.sum()) / cpu_s, "unit": "points/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{len(pick)} random batches ({int(sub.n_point_batches.sum())} points)"}

    # bias part 2 of the first-order H (evaluate_first_order_H_bias_part2) for the same atom: random
    # H_times_psi and vrho; HBM-bound (H_times_psi, phi and dphi read once, the atom's rows of dphi, the
    # point data, two FP64 atomics per (row, atom row) element)
    hpsi, vrho = U(wl), U(nmb * nb)

    def run_b2():
        plan.h1_bias_part2(wave, gwave, hpsi, vrho, xc["vsigma"], xc["gradient_rho"], xc["partition_tab"], avn,
                           prob.n_atoms, jb, j_coord=jc, H=H)
    for _ in range(max(warmup, 1)):
        run_b2()
    torch.cuda.synchronize()
    ev = []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        run_b2()
        e1.record()
        ev.append((e0, e1))
    torch.cuda.synchronize()
    ms2 = float(np.mean([x.elapsed_time(y) for x, y in ev]))
    b2_bytes = float(np.sum((5 * 8 * nc * npt + 40 * npt + 32 * nc * nv)[nv > 0]))
    b2_flops = float(np.sum((2 * nc * nv * npt + 12 * nc * npt)[nv > 0]))
    res["bias_part2"] = {"ms": ms2, "algorithmic_bytes": b2_bytes, "algorithmic_flops": b2_flops,
                         "roofline": {"kernel": "frpa::h1_bias2_kernel", "bound": "hbm",
                                      "achieved": b2_bytes / (ms2 / 1e3) / 1e9, "peak": PEAK_HBM, "unit": "GB/s",
                                      "frac": b2_bytes / (ms2 / 1e3) / 1e9 / PEAK_HBM}}
    if not no_cpu:
        hp_s = blocks(hpsi, 1)
        vr_s = per_batch(vrho, 1).reshape(nmb2, len(pick), order="F")
        vs_s = per_batch(xc["vsigma"], 3).reshape(3, nmb2, len(pick), order="F")
        ref2 = oracle.h1_bias2(sub, 1, jb, jb, jc, np.zeros((prob.n_basis, prob.n_basis, 1), order="F"), vr_s, vs_s,
                               arr["gradient_rho"], hp_s, arr["gwave"], arr["partition_tab"])
        Hb = splan.h1_bias_part2(api.to_device(sub.wave), api.to_device(arr["gwave"]), api.to_device(hp_s),
                                 api.to_device(vr_s), api.to_device(vs_s), api.to_device(arr["gradient_rho"]),
                                 api.to_device(arr["partition_tab"]), api.to_device(sub.atom_valid_n_compute_c_batches()),
                                 prob.n_atoms, jb, j_coord=jc)
        got2 = Hb.cpu().numpy().reshape(prob.n_basis, prob.n_basis, order="F")
        res["bias_part2"]["vs_oracle_rel"] = float(np.abs(got2 - ref2[:, :, 0]).max() / max(np.abs(ref2).max(), 1e-300))
    log(f"H1 bias part 2: {res['bias_part2']}")
    return res


def time_near(N, steps, warmup):
    """Near-field launch timed with HIP events on torch's current stream (the stream it runs on)."""
    import torch
    fpn = N["fp"]
    dv = torch.zeros(fpn.n_full_points, dtype=torch.float64, device="cuda")

    def run():
        dv.zero_()
        N["plan"].near_forward(dv, *N["inp"].near_args())
    for _ in range(max(warmup, 1)):
        run()
    torch.cuda.synchronize()
    t = []
    for _ in range(steps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        dv.zero_()
        a.record()
        N["plan"].near_forward(dv, *N["inp"].near_args())
        b.record()
        t.append((a, b))
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(b) for a, b in t])), dv


def step_fns(D, prec, adj_prec, far_variant, have_adj):
    import torch

    from fast_rpa_eval_amd import api
    prob, plan = D["prob"], D["plan"]
    rho = torch.empty(prob.n_full_points, dtype=torch.float64, device="cuda")
    H = torch.zeros(prob.n_basis * prob.n_basis, dtype=torch.float64, device="cuda")
    dv = torch.zeros(D["fp"].n_full_points, dtype=torch.float64, device="cuda")

    def rho1():
        plan.forward(D["wave"], D["P1"], precision=prec, out=rho)

    def adj():
        if have_adj:
            H.zero_()
            plan.adjoint(D["wave"], D["g"], precision=adj_prec, H=H)

    def far():
        dv.zero_()
        D["fplan"].forward(dv, *D["finp"].args(), variant=far_variant)

    return dict(rho1=rho1, adjoint=adj, far=far), dict(rho=rho, H=H, dv=dv)


def time_graph(fns, names, steps, warmup):
    """The step captured once in a HIP graph (torch.cuda.CUDAGraph over the C-ABI launches) and
    replayed: ms per replay over `steps` replays."""
    import torch
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(max(warmup, 1)):
            for n in names:
                fns[n]()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for n in names:
            fns[n]()
    g.replay()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(steps):
        g.replay()
    t1.record()
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps, g


def time_it(fns, names, steps, warmup):
    import torch
    for _ in range(warmup):
        for n in names:
            fns[n]()
    torch.cuda.synchronize()
    ev = {n: [] for n in names}
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record()
    for _ in range(steps):
        for n in names:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fns[n]()
            b.record()
            ev[n].append((a, b))
    t1.record()
    torch.cuda.synchronize()
    total = t0.elapsed_time(t1) / steps
    per = {n: float(np.mean([a.elapsed_time(b) for a, b in ev[n]])) for n in names}
    return total, per


def _nc_np(prob):
    nc = prob.n_compute_c_batches.astype(np.float64)
    npt = prob.n_point_batches.astype(np.float64)
    sel = (nc > 0) & (npt > 0)
    return nc[sel], npt[sel]


def work_model(prob, kernel, prec, S=None):
    """Algorithmic work of one launch: (flops or int8 ops, HBM bytes), summed over the batches.

    rho1 FP64 (full GEMM as in the reference): 2 nc^2 np + 2 nc np flop;
        bytes 8 nc^2 (P1 gather) + 8 nc np (phi) + 4 nc + 12 np (basis list, map, output).
    rho1 INT8 split (upper tile pairs of the symmetric part): S(S+1)/2 * 2 * nc(nc+1)/2 * np int8 ops;
        bytes 8 nc(nc+1)/2 (packed P' digits) + S nc np (phi digits; phi(i, p) for the ddot is rebuilt
        from them, the FP64 wave is not read) + 4 nc + 12 np.
    adjoint FP64 (upper triangle of the symmetric local H): 2 nc(nc+1)/2 np + nc np flop;
        bytes 8 nc np (phi) + 12 np + 16 nc(nc+1)/2 (read-modify-write of the scatter into H).
    adjoint INT8 split: S(S+1)/2 * 2 * nc(nc+1)/2 * np int8 ops; bytes S nc np (phi digits, made once per
        wave) + 12 np + 16 nc(nc+1)/2.
    """
    nc, npt = _nc_np(prob)
    tri = nc * (nc + 1) / 2
    if kernel == "rho1" and S is None:
        return float(np.sum(2 * nc * nc * npt + 2 * nc * npt)), float(np.sum(8 * nc * nc + 8 * nc * npt + 4 * nc + 12 * npt))
    if kernel == "rho1":
        return (float(np.sum(S * (S + 1) / 2 * 2 * tri * npt)),
                float(np.sum(8 * tri + S * nc * npt + 4 * nc + 12 * npt)))
    if kernel == "adjoint" and S is None:
        return float(np.sum(2 * tri * npt + nc * npt)), float(np.sum(8 * nc * npt + 12 * npt + 16 * tri))
    if kernel == "adjoint":
        return (float(np.sum(S * (S + 1) / 2 * 2 * tri * npt)),
                float(np.sum(S * nc * npt + 12 * npt + 16 * tri)))
    raise ValueError(kernel)


PROFILED_KERNEL = {  # bench kernel (path) -> kernel name prefix in the rocprofv3 summaries
    ("rho1", "fast"): "frpa::rho1_ozaki_kernel", ("rho1", "fp64"): "frpa::rho1_fp64_kernel",
    ("adjoint", "fast"): "frpa::h1_ozaki_kernel", ("adjoint", "fp64"): "frpa::h1_push_kernel",
    ("far", "fast"): "frpa::far_horner_kernel", ("far", "fp64"): "frpa::far_forward_kernel<0,",
}


def read_counter_table(path, counters=("FETCH_SIZE", "WRITE_SIZE")):
    """{(kernel, counter): mean per dispatch} from a summarize_profile counter table."""
    out = {}
    for ln in open(path):
        parts = ln.split()
        if len(parts) >= 5 and parts[-4] in counters:
            out[(" ".join(parts[:-4]), parts[-4])] = float(parts[-1])
    return out


F64_COUNTERS = ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_TRANS_F64")


def load_f64_counters(workload, kernel):
    """Executed FP64 VALU work of one launch of `kernel` from profiles/counters_fp64_<workload>.txt:
    64 lanes x (2 per v_fma_f64 + 1 per add / mul / transcendental) per wave instruction (lanes masked
    off still count), and the instruction counts; None if the table is absent."""
    path = os.path.join(ROOT, "profiles", f"counters_fp64_{workload}.txt")
    if not os.path.exists(path):
        return None
    c = {k[1]: v for k, v in read_counter_table(path, F64_COUNTERS).items() if k[0].startswith(kernel)}
    if "SQ_INSTS_VALU_FMA_F64" not in c:
        return None
    flop = 64.0 * (2.0 * c["SQ_INSTS_VALU_FMA_F64"] + c.get("SQ_INSTS_VALU_ADD_F64", 0.0) +
                   c.get("SQ_INSTS_VALU_MUL_F64", 0.0) + c.get("SQ_INSTS_VALU_TRANS_F64", 0.0))
    return {"executed_flop": flop, "wave_instructions": c}


def traffic_table(fetch_txt, write_txt, cf_txt, cw_txt, copy_bytes=2 ** 30):
    """Markdown table of HBM bytes per launch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of each
    frpa kernel, each scaled by copy_bytes / (the counter's reading for a plain copy of copy_bytes)."""
    f, w = read_counter_table(fetch_txt), read_counter_table(write_txt)
    cf, cw = read_counter_table(cf_txt), read_counter_table(cw_txt)
    sf = copy_bytes / (cf[("__amd_rocclr_copyBuffer", "FETCH_SIZE")] * 1024)
    sw = copy_bytes / (cw[("__amd_rocclr_copyBuffer", "WRITE_SIZE")] * 1024)
    rows = [f"copy calibration: {copy_bytes} B copy reads FETCH_SIZE "
            f"{cf[('__amd_rocclr_copyBuffer', 'FETCH_SIZE')]:.1f} KiB -> fetch scale {sf:.4f}; "
            f"WRITE_SIZE {cw[('__amd_rocclr_copyBuffer', 'WRITE_SIZE')]:.1f} KiB -> write scale {sw:.4f}", "",
            "| kernel | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | HBM bytes/launch (scaled) |",
            "|---|---|---|---|"]
    for (k, c), v in sorted(f.items()):
        if not k.startswith("frpa::") or c != "FETCH_SIZE":
            continue
        wv = w.get((k, "WRITE_SIZE"), 0.0)
        rows.append(f"| {k} | {v:.1f} | {wv:.1f} | {v * 1024 * sf + wv * 1024 * sw:.4e} |")
    return "\n".join(rows)


def load_traffic(workload):
    """HBM bytes per launch by kernel name from profiles/hbm_traffic_<workload>.md (None if absent)."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    out = {}
    for ln in open(path):
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0].startswith("frpa::"):
            out[cells[0]] = float(cells[3])
    return out


def compare_precisions(D, rho_fp64, steps):
    """rho1 of every precision on the same inputs: ms per call and the error against the FP64 GPU path,
    max over points of |d rho1(p)| / max|rho1| and the median over points with |rho1(p)| > 1e-3 max of
    the pointwise |d rho1(p)| / |rho1(p)|.  fp64_symmetric is FP64 on the INT8 split's algorithm (the
    upper tile pairs of the symmetric part of P1, the same launch)."""
    import torch

    from fast_rpa_eval_amd import api
    from fast_rpa_eval_amd._lib import FrpaError
    prob = D["prob"]
    valid = torch.from_numpy(prob.i_valid_point_2_i_full_points_map[:prob.n_point_batches_prefix_sum[-1]]).cuda().long()
    ref = rho_fp64[valid]
    big = ref.abs() > 1e-3 * ref.abs().max()
    out = torch.empty_like(rho_fp64)
    res = {}
    for name, prec in (("fp64_symmetric", api.FP64_SYM), ("int8_split_s5", api.OZAKI), ("int8_split_s4", api.OZAKI4),
                       ("fp32_mfma", api.FP32), ("bf16x3", api.BF16X3)):
        try:
            D["plan"].forward(D["wave"], D["P1"], precision=prec, out=out)
        except FrpaError as e:
            res[name] = {"error": str(e)}
            continue
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(steps):
            D["plan"].forward(D["wave"], D["P1"], precision=prec, out=out)
        t1.record()
        torch.cuda.synchronize()
        d = (out[valid] - ref).abs()
        res[name] = {"ms": t0.elapsed_time(t1) / steps, "max_rel_err": float(d.max() / ref.abs().max()),
                     "median_pointwise_rel_err": float((d[big] / ref[big].abs()).median())}
        log(f"precision {name}: {res[name]}")
    return res


def host_cores():
    """Threads the CPU baseline uses: OMP_NUM_THREADS if set (the GPU host's share), else every CPU in
    this process's affinity mask; and the size of that mask."""
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else aff
    return min(n, aff), aff


def cpu_baseline(D, rho_s=2.0, far_s=1.0, runs=5):
    """The CPU FP64 oracle (the reference's loops restated in C: its dgemm as a plain loop nest, no
    BLAS) on bounded samples of the same inputs: random batches (rho1) and the first far-field points.
    Two legs -- all host cores (OpenMP across batches / points) and one thread (the reference's own
    build links the sequential MKL, task1_atom_tiling/version3_rho_with_bias/Makefile:17) -- each sized
    by a pilot run (also the warm-up) to about rho_s + far_s seconds, then the median of `runs` runs."""
    from fast_rpa_eval_amd import synth
    from oracle import oracle
    prob, fp = D["prob"], D["fp"]
    rng = np.random.default_rng(0)
    # the device P1 is column-major: element (r, c) at r + n_basis * c
    P1 = np.asfortranarray(D["P1"].view(prob.n_basis, prob.n_basis).cpu().numpy().T)[:, :, None]
    perm = rng.permutation(prob.n_batches)
    cores, aff = host_cores()

    def rho_sample(nb):
        sel = np.sort(perm[:nb])
        return sel, subproblem(prob, sel, D["wave"])

    def far_sample(n):
        fsub = synth.FarProblem(**{**fp.__dict__})
        fsub.i_valid_point_2_i_full_points_map = fp.i_valid_point_2_i_full_points_map[:n].copy()
        fsub.n_valid_points = n
        return fsub

    def timed(f):
        t0 = time.time()
        out = f()
        return out, time.time() - t0

    legs = {}
    for name, threads in (("all_cores", cores), ("single_thread", 1)):
        oracle.set_threads(threads)
        nb = min(16 * threads, prob.n_batches)
        sel, sub = rho_sample(nb)
        _, t = timed(lambda: oracle.rho1(sub, P1))
        nb = int(min(prob.n_batches, max(1, nb * rho_s / max(t, 1e-3))))
        sel, sub = rho_sample(nb)
        nf = min(fp.n_valid_points, 200 * threads)
        _, t = timed(lambda: oracle.far(far_sample(nf)))
        nf = int(min(fp.n_valid_points, max(64, nf * far_s / max(t, 1e-3))))
        fsub = far_sample(nf)
        t_rho, t_far = [], []
        for _ in range(runs):
            (rho_cpu, _), tr = timed(lambda: oracle.rho1(sub, P1))
            dv_cpu, tf = timed(lambda: oracle.far(fsub))
            t_rho.append(tr)
            t_far.append(tf)
        pts_rho = int(sub.n_point_batches.sum())
        mr, mf = float(np.median(t_rho)), float(np.median(t_far))
        legs[name] = dict(threads=threads, rho_pts=pts_rho, far_pts=nf, rho_s=mr, far_s=mf,
                          value=(pts_rho + nf) / (mr + mf), rho_runs_s=t_rho, far_runs_s=t_far,
                          sel=sel, sub=sub, rho_cpu=rho_cpu, dv_cpu=dv_cpu, nfs=nf)
        log(f"CPU oracle, {threads} thread(s): rho1 {nb} batches {mr:.2f}s, far {nf} points {mf:.2f}s "
            f"(median of {runs}) -> {legs[name]['value']:.0f} points/s")
    legs["affinity_cpus"] = aff
    return legs


def subproblem(prob, sel, wave_dev):
    """Problem restricted to the batches `sel`, with their wave blocks copied to the host."""
    import copy
    sub = copy.copy(prob)
    nps = prob.n_point_batches[sel]
    ncs = prob.n_compute_c_batches[sel]
    ncp = (ncs.astype(np.int64) + 3) // 4 * 4
    sub.n_point_batches = nps.astype(np.int32)
    sub.n_compute_c_batches = ncs.astype(np.int32)
    sub.n_point_batches_prefix_sum = np.concatenate([[0], np.cumsum(nps)]).astype(np.int32)
    maps = [prob.i_valid_point_2_i_full_points_map[prob.n_point_batches_prefix_sum[b]:prob.n_point_batches_prefix_sum[b + 1]]
            for b in sel]
    sub.i_valid_point_2_i_full_points_map = np.concatenate(maps).astype(np.int32)
    sub.batch_sizes = prob.batch_sizes[sel]
    sub.ibasis = np.concatenate([prob.ibasis[prob.ibasis_ptr[b]:prob.ibasis_ptr[b + 1]] for b in sel])
    sub.ibasis_ptr = np.concatenate([[0], np.cumsum(ncs)]).astype(np.int64)
    sub.wave_offset = np.concatenate([[0], np.cumsum(ncp * nps)]).astype(np.int64)
    sub.n_max_batch_size = int(max(nps.max(), 1))
    sub.n_max_compute_ham = int(ncs.max())
    sub.wave = np.concatenate([wave_dev[int(prob.wave_offset[b]):int(prob.wave_offset[b + 1])].cpu().numpy()
                               for b in sel])
    return sub


def rooflines(r, path, D, w, tuned, workload):
    """{kernel: roofline record} for one path's per-kernel times r["per"] (see the roofline of main)."""
    from fast_rpa_eval_amd import api
    prob, fp = D["prob"], D["fp"]
    traffic = load_traffic(workload) or {}
    rl = {}
    for name, tms in r["per"].items():
        t = tms / 1e3
        prof = PROFILED_KERNEL[(name, path)]
        if name == "rho1" and path == "fast":
            # the INT8 kernel's template arguments follow the chosen launch: npg 2 = four-wave workgroups
            # with register-staged phi chunks, npg 1 = four-wave workgroups with direct-to-LDS staging
            m = re.search(r"npg=(\d+)", tuned.get(path) or "")
            npg = int(m.group(1)) if m else 0
            prof += {1: "<5,4,1,false,true>", 2: "<5,4,1,false,false>"}.get(npg, "<5,4,2,false,false>")
        tr = next((v for k, v in traffic.items() if k.startswith(prof)), None)
        if name == "far":
            from fast_rpa_eval_amd.multipole_tables import horner_fma_count
            pairs = far_pairs(D)
            per_pair = (2.0 * horner_fma_count(w["l_far"]) + 25) if r["var"] == api.FAR_HORNER else \
                (2.0 * fp.tables["n_cc_lm_ijk"][w["l_far"]] + 25)
            ach = pairs * per_pair / t / 1e12
            rl[name] = {"kernel": prof, "bound": "mfma", "achieved": ach, "peak": PEAK_FP64, "unit": "TFLOP/s",
                        "frac": ach / PEAK_FP64, "traffic": tr, "pairs": pairs, "flop_per_pair": per_pair,
                        "note": "FP64 VALU (v_fma_f64); the FP64 vector and matrix peaks are equal on MI355X"}
            ex = load_f64_counters(workload, prof)
            if ex:
                rl[name]["executed"] = {
                    "flop_per_launch": ex["executed_flop"], "achieved": ex["executed_flop"] / t / 1e12,
                    "frac": ex["executed_flop"] / t / 1e12 / PEAK_FP64,
                    "wave_instructions": ex["wave_instructions"],
                    "note": f"executed FP64 VALU instructions per launch (rocprofv3 counters, profiles/"
                            f"counters_fp64_{workload}.txt) over this run's kernel time"}
            continue
        prec = r["prec"] if name == "rho1" else r["adj_prec"]
        S = {api.OZAKI: 5, api.OZAKI4: 4}.get(prec)
        ops, nbytes = work_model(prob, name, prec, S)
        peak_ops = PEAK_INT8 if S else PEAK_FP64
        t_ops = ops / (peak_ops * 1e12)
        t_mem = nbytes / (PEAK_HBM * 1e9)
        if t_mem > t_ops:
            ach = nbytes / t / 1e9
            rl[name] = {"kernel": prof, "bound": "hbm", "achieved": ach, "peak": PEAK_HBM, "unit": "GB/s",
                        "frac": ach / PEAK_HBM, "traffic": tr, "algorithmic_bytes": nbytes}
        else:
            ach = ops / t / 1e12
            rl[name] = {"kernel": prof, "bound": "mfma", "achieved": ach, "peak": peak_ops,
                        "unit": "TOP/s" if S else "TFLOP/s", "frac": ach / peak_ops, "traffic": tr,
                        "algorithmic_bytes": nbytes}
        rl[name]["algorithmic_ops"] = ops
        rl[name]["other_bound_frac"] = (ops / t / 1e12 / peak_ops) if t_mem > t_ops else (nbytes / t / 1e9 / PEAK_HBM)
    return rl


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU oracle in the side measurements (0 = OMP_NUM_THREADS, else "
                         "every CPU of the affinity mask)")
    ap.add_argument("--paths", default="fast,fp64")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default rho1 launch configuration")
    ap.add_argument("--compare-precisions", type=int, default=1,
                    help="1: also time rho1 in every precision (INT8 split S = 5 / 4, FP32 MFMA, BF16x3) and state "
                         "its error against the FP64 path")
    ap.add_argument("--launch-fast", metavar="ORDER,PERSISTENT,NPG",
                    help="fixed rho1 launch configuration of the fast path instead of the autotuner (profiles)")
    ap.add_argument("--launch-fp64", metavar="ORDER,PERSISTENT,NPG",
                    help="fixed rho1 launch configuration of the FP64 path instead of the autotuner")
    ap.add_argument("--graph", type=int, default=1,
                    help="1: the timed step is one replay of the step captured in a HIP graph (per-kernel "
                         "times from an eager pass); 0: eager launches")
    ap.add_argument("--near", type=int, default=1,
                    help="1: also time the near-field (spline) branch of the Hartree sum over the rho1 grid "
                         "(reported as near_field, outside the step)")
    ap.add_argument("--bias2", type=int, default=1,
                    help="1: also time the bias generator (generate_first_order_rho_bias_part2) for the tile of "
                         "atoms 1-4 over all batches (reported as bias_part2, outside the step)")
    ap.add_argument("--h1-gga", type=int, default=1,
                    help="1: also time the full first-order H (LDA + GGA terms) on the 512-atom, 20k-batch workload "
                         "(reported as h1_gga, outside the step; run last, after the step's inputs are released)")
    ap.add_argument("--kernels", default="rho1,adjoint,far",
                    help="subset of the step to time (a partial step is not a bench line)")
    ap.add_argument("--summarize-profile", metavar="DIR",
                    help="only summarise a rocprofv3 --output-format csv directory (kernel stats and/or "
                         "per-kernel counter sums) as a text table and exit")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run a plain device-to-device copy of GIB GiB five times (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiler run) and exit")
    ap.add_argument("--traffic-table", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="only print the HBM traffic table from four summaries made by --summarize-profile "
                         "(kernel FETCH_SIZE, kernel WRITE_SIZE, 1 GiB copy FETCH_SIZE, copy WRITE_SIZE)")
    ap.add_argument("--config3", type=int, default=1,
                    help="1: also run the 512-atom, 20k-batch rho1 workload (n_compute 80-800) in every precision, "
                         "with and without the GGA gradient: ms and errors against the FP64 path (reported as "
                         "rho1_512atoms_precisions, outside the step; run last)")
    a = ap.parse_args()
    if a.cpu_threads <= 0:
        a.cpu_threads = host_cores()[0]
    if a.traffic_table:
        print(traffic_table(*a.traffic_table))
        return
    if a.summarize_profile:
        summarize_profile(a.summarize_profile)
        return
    if a.copy_calibration > 0:
        import torch
        n = int(a.copy_calibration * 2 ** 30) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n * 8, "launches": 5}))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU; the rho1 and far-field paths would shard by batches/points "
              "with no exchange (H1 needs one all-reduce), but multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from fast_rpa_eval_amd import api
    from fast_rpa_eval_amd._lib import FrpaError
    w = WORKLOADS[a.workload]
    D = make_inputs(w)
    prob, fp = D["prob"], D["fp"]

    # capability probes (small, on the real inputs)
    have_adj = True
    try:
        D["plan"].adjoint(D["wave"], D["g"], precision=api.FP64)
    except FrpaError as e:
        have_adj = False
        log(f"adjoint unavailable: {e}")
    have_ozaki = True
    try:
        D["plan"].forward(D["wave"], D["P1"], precision=api.OZAKI)
    except FrpaError as e:
        have_ozaki = False
        log(f"INT8 split unavailable: {e}")
    have_ozaki_adj = have_adj and have_ozaki
    if have_ozaki_adj:
        try:
            D["plan"].adjoint(D["wave"], D["g"], precision=api.OZAKI)
        except FrpaError as e:
            have_ozaki_adj = False
            log(f"INT8-split adjoint unavailable (fast path uses the FP64 adjoint): {e}")
    torch.cuda.synchronize()

    names = ["rho1", "adjoint", "far"] if have_adj else ["rho1", "far"]
    names = [n for n in names if n in a.kernels.split(",")]
    res = {}
    outs = {}
    tuned = {}
    for path in a.paths.split(","):
        prec = api.OZAKI if (path == "fast" and have_ozaki) else api.FP64
        adj_prec = api.OZAKI if (path == "fast" and have_ozaki_adj) else api.FP64
        var = api.FAR_HORNER if path == "fast" else api.FAR_TERMS
        fixed = a.launch_fast if path == "fast" else a.launch_fp64
        if fixed:
            o, pers, npg = (int(x) for x in fixed.split(","))
            D["plan"].set_launch(prec, o, pers, npg)
            tuned[path] = f"fixed order={o} persistent={pers} npg={npg}"
        elif not a.no_autotune and "rho1" in names:
            t0 = time.time()
            tuned[path] = D["plan"].autotune(D["wave"], D["P1"], precision=prec, reps=2)
            log(f"autotune {path} ({time.time() - t0:.1f}s): {tuned[path]}")
        fns, o = step_fns(D, prec, adj_prec, var, have_adj)
        eager, per = time_it(fns, names, a.steps, a.warmup)
        total, graph_ms = eager, None
        if a.graph:
            graph_ms, _ = time_graph(fns, names, a.steps, a.warmup)
            total = graph_ms
        res[path] = dict(ms=total, eager_ms=eager, graph_ms=graph_ms, per=per, prec=prec, adj_prec=adj_prec, var=var)
        outs[path] = {k: v.clone() for k, v in o.items()}
        log(f"path {path}: {total:.3f} ms/step (eager {eager:.3f}, graph {graph_ms})  " +
            "  ".join(f"{n} {per[n]:.3f} ms" for n in names))

    pts_rho = int(prob.n_point_batches[prob.n_compute_c_batches > 0].sum())
    pts = pts_rho + fp.n_valid_points
    fast = res.get("fast", res.get("fp64"))
    value = pts / (fast["ms"] / 1e3)
    line = {
        "metric": "rho1+Hartree grid points/s (fast path vs plain FP64, max rel err <= 1e-9)",
        "value": value, "unit": "points/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": fast["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8-split+f64" if (have_ozaki and "fast" in res) else "f64", "data": "synthetic",
        "config": {"workload": a.workload, "n_atoms": w["n_atoms"], "n_basis": prob.n_basis,
                   "n_batches": prob.n_batches, "rho1_points": pts_rho, "far_points": fp.n_valid_points,
                   "l_max_far": w["l_far"], "adjoint": have_adj, "int8_split_rho1": have_ozaki,
                   "int8_split_adjoint": have_ozaki_adj},
    }
    if "fp64" in res:
        line["fp64_value"] = pts / (res["fp64"]["ms"] / 1e3)
        line["fp64_ms_per_step"] = res["fp64"]["ms"]
        line["speedup_vs_fp64"] = res["fp64"]["ms"] / fast["ms"]
    if "fp64" in res and "fast" in res and set(res["fp64"]["per"]) == set(res["fast"]["per"]) and "far" in names:
        # the FP64 baseline with the same far-field kernel as the fast path (FP64 Horner), so that the
        # ratio measures the precision change of rho1 and its adjoint alone; per-kernel eager times
        f, d = res["fast"]["per"], res["fp64"]["per"]
        t_f = sum(f.values())
        t_d = f["far"] + sum(v for k, v in d.items() if k != "far")
        line["speedup_vs_fp64_same_far"] = {
            "value": t_d / t_f, "fast_ms": t_f, "fp64_ms": t_d,
            "note": "FP64 rho1 + FP64 adjoint + the FP64 Horner far field against the fast path, from the "
                    "per-kernel eager times; speedup_vs_fp64 also counts the far field's change of formulation "
                    "(term by term in the plain FP64 path, Horner in both paths here).  The FP64 rho1 kernel "
                    "multiplies the full P1 (as the reference); the INT8 split only the upper tile pairs of its "
                    "symmetric part"}
    if "fast" in res and have_ozaki:
        # the one-off conversion of the wave into the INT8 digits of the forward and of the adjoint
        # (frpa_rho1_prepare_wave rebuilds both once the adjoint has made its own), kept while the wave
        # is unchanged; outside the timed step
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        nbytes = D["plan"].prepare_wave(D["wave"], api.OZAKI)
        e1.record()
        torch.cuda.synchronize()
        one_ms = e0.elapsed_time(e1)
        line["one_off"] = {"ms": one_ms, "bytes": nbytes, "wave_bytes": int(prob.wave_offset[-1]) * 8,
                           "ms_per_step_with_one_off": fast["ms"] + one_ms,
                           "note": "phi_digits_kernel (forward) + h1_digits_kernel and h1_pointmax_kernel (adjoint) "
                                   "and their host set-up, once per wave (the wave is fixed across the CPSCF "
                                   "iterations and perturbations that reuse it); bytes = device memory the digits "
                                   "hold"}
        log(f"one-off digit conversion: {one_ms:.1f} ms, {nbytes / 1e9:.1f} GB")
    if "fp64" in res and "fast" in res and all(k in res["fast"]["per"] for k in ("rho1", "far")):
        f, d = res["fast"]["per"], res["fp64"]["per"]
        t_f, t_d = f["rho1"] + f["far"], d["rho1"] + d["far"]
        line["forward_only"] = {
            "value": pts / (t_f / 1e3), "fp64_value": pts / (t_d / 1e3), "speedup_vs_fp64": t_d / t_f,
            "ms": t_f, "fp64_ms": t_d,
            "note": "rho1 forward + far field without the adjoint: (points_rho1 + points_far) / (t_rho1 + t_far) "
                    "from the per-kernel eager times; the headline value times the whole step (adjoint included)"}
    line["ms_per_kernel"] = {p: r["per"] for p, r in res.items()}
    line["step_timing"] = {p: {"graph_ms": r["graph_ms"], "eager_ms": r["eager_ms"]} for p, r in res.items()}
    if tuned:
        line["autotune"] = tuned
    if len(names) < (3 if have_adj else 2):
        line["partial_step"] = names  # timing experiment: value does not cover the whole step
    if "fast" in outs and "fp64" in outs:
        r_f, r_d = outs["fast"]["rho"], outs["fp64"]["rho"]
        line["rho1_rel_err_vs_fp64"] = float((r_f - r_d).abs().max() / r_d.abs().max())
        v_f, v_d = outs["fast"]["dv"], outs["fp64"]["dv"]
        line["far_rel_err_vs_fp64"] = float((v_f - v_d).abs().max() / v_d.abs().max())
        if have_adj:
            h_f, h_d = outs["fast"]["H"], outs["fp64"]["H"]
            line["h1_rel_err_vs_fp64"] = float((h_f - h_d).abs().max() / h_d.abs().max())

    if a.compare_precisions and "fp64" in outs and "rho1" in names:
        line["rho1_precision_comparison"] = compare_precisions(D, outs["fp64"]["rho"], a.steps)
        sym = line["rho1_precision_comparison"].get("fp64_symmetric", {})
        if "speedup_vs_fp64_same_far" in line and "ms" in sym and "adjoint" in res["fp64"]["per"]:
            # FP64 on the fast path's algorithms: rho1 on the symmetric half (the same launch as the INT8
            # split), the FP64 adjoint (upper triangle, as the INT8 one), the FP64 Horner far field
            f, d = res["fast"]["per"], res["fp64"]["per"]
            t_f = sum(f.values())
            t_d = sym["ms"] + d["adjoint"] + f["far"]
            line["speedup_vs_fp64_same_algorithm"] = {
                "value": t_d / t_f, "fast_ms": t_f, "fp64_ms": t_d,
                "note": "FP64 symmetric-half rho1 (rho1_precision_comparison.fp64_symmetric) + FP64 adjoint + FP64 "
                        "Horner far field against the fast path, per-kernel times: the precision change alone for "
                        "rho1 and the far field; the FP64 adjoint scatters every batch on its own, the INT8 one "
                        "merges batch groups (fewer H updates)"}
            log(f"speedup vs FP64 on the same algorithms: {t_d:.1f} / {t_f:.1f} ms = {t_d / t_f:.2f}")

    # roofline of every kernel of each path: algorithmic work / measured time against the peak of
    # whichever resource binds the kernel at its algorithmic work; the fast path's dominant kernel is
    # "roofline"
    rls = {p: rooflines(res[p], p, D, w, tuned, a.workload) for p in res}
    rl = rls["fast" if "fast" in res else "fp64"]
    dom = max(fast["per"], key=fast["per"].get)
    line["roofline"] = rl[dom]
    line["roofline_all"] = rl
    if "fp64" in rls and "fast" in rls:
        line["roofline_fp64_path"] = rls["fp64"]

    if a.near:
        N = make_near_inputs(D, w)
        ms, dvn = time_near(N, a.steps, a.warmup)
        fpn = N["fp"]
        npairs, nlm = near_pairs(fpn)
        L = w["l_far"]
        # algorithmic work per near pair: 4 FMAs per lm for the Hermite dot products (8 flop), ~3 flop per
        # lm for Y_lm by its recurrence, ~40 for distance, log and weights; (l+1)^2 per pair after the
        # outer-radius reduction
        flops = 11.0 * nlm + 40.0 * npairs
        coef_bytes = 32.0 * nlm
        ach = flops / (ms / 1e3) / 1e12
        near = {"workload": f"near field over the rho1 grid: {fpn.n_valid_points} points, {fpn.n_atoms} centres, "
                            f"l_hartree {L}/{max(L - 2, 0)} (2 species), multipole radius {NEAR_RADIUS} bohr, "
                            f"{NEAR_GRID}-point log grids",
                "ms": ms, "points_per_s": fpn.n_valid_points / (ms / 1e3), "near_pairs": npairs,
                "pairs_per_s": npairs / (ms / 1e3),
                "roofline": {"kernel": "frpa::near_kernel", "bound": "mfma", "achieved": ach, "peak": PEAK_FP64,
                             "unit": "TFLOP/s", "frac": ach / PEAK_FP64, "flop_per_pair_lm": 11,
                             "coef_bytes_per_s": coef_bytes / (ms / 1e3),
                             "note": "FP64 VALU at the nominal count; the binding resource is the per-pair "
                                     "gather of the interval's 4 (l+1)^2 coefficients (coef_bytes_per_s), served "
                                     "from L2 (the tiles, 0.27 GB, are read from HBM about once)"}}
        if not a.no_cpu:
            from oracle import oracle
            sel = np.sort(np.random.default_rng(11).choice(fpn.n_valid_points, size=min(20000, fpn.n_valid_points),
                                                           replace=False))
            sub = copy.copy(fpn)
            sub.i_valid_point_2_i_full_points_map = fpn.i_valid_point_2_i_full_points_map[sel]
            sub.n_valid_points = len(sel)
            oracle.set_threads(a.cpu_threads)
            t0 = time.time()
            ref = oracle.near(sub)[:, 0]
            cpu_s = time.time() - t0
            mp = sub.i_valid_point_2_i_full_points_map
            got = dvn.cpu().numpy()[mp]
            near["vs_oracle_rel"] = float(np.abs(got - ref[mp]).max() / np.abs(ref[mp]).max())
            near["cpu_baseline"] = {"value": len(sel) / cpu_s, "unit": "points/s", "cores": a.cpu_threads,
                                    "kind": "port", "sample": f"{len(sel)} random points of the grid"}
        tr = load_traffic(a.workload) or {}
        near["roofline"]["traffic"] = next((v for k, v in tr.items() if k.startswith("frpa::near_kernel")), None)
        line["near_field"] = near
        log(f"near field: {ms:.3f} ms, {npairs} pairs, {near['pairs_per_s'] / 1e9:.2f} G pairs/s, "
            f"{ach:.1f} TF/s ({100 * ach / PEAK_FP64:.0f}% FP64)")

    if a.bias2:
        try:
            line["bias_part2"] = bench_bias2(D, a.steps, a.warmup, a.no_cpu, a.cpu_threads)
            b2 = line["bias_part2"]
            # the call's kernels: the DM + DM^T table (b2_mark / b2_scan / b2_table) and the product
            tr = load_traffic(a.workload) or {}
            parts = [v for k, v in tr.items() if k.startswith("frpa::bias2_kernel") or k.startswith("frpa::b2_")]
            b2["roofline"]["traffic"] = float(sum(parts)) if parts else None
            log(f"bias part 2: {b2['ms']:.3f} ms, {b2['roofline']['achieved']:.0f} GB/s "
                f"({100 * b2['roofline']['frac']:.0f}% HBM)")
        except torch.cuda.OutOfMemoryError as e:
            log(f"bias part 2 skipped: {e}")
        torch.cuda.empty_cache()

    if not a.no_cpu:
        legs = cpu_baseline(D)
        cb, c1 = legs["all_cores"], legs["single_thread"]
        line["cpu_baseline"] = {
            "value": cb["value"], "unit": "points/s", "cores": cb["threads"], "kind": "port",
            "statistic": "median of 5 runs after a pilot run",
            "sample": f"{len(cb['sel'])} random batches ({cb['rho_pts']} rho1 points, {cb['rho_s']:.2f}s) + "
                      f"{cb['far_pts']} far-field points x {fp.n_atoms} centres ({cb['far_s']:.2f}s), CPU FP64 "
                      f"oracle (the reference's dgemm as a plain loop nest, no BLAS), OpenMP {cb['threads']} "
                      f"threads of {legs['affinity_cpus']} CPUs in the affinity mask",
            "single_thread": {"value": c1["value"], "cores": 1,
                              "sample": f"{len(c1['sel'])} random batches ({c1['rho_pts']} rho1 points, "
                                        f"{c1['rho_s']:.2f}s) + {c1['far_pts']} far-field points "
                                        f"({c1['far_s']:.2f}s), median of 5"},
            "runs_s": {"all_cores": {"rho1": cb["rho_runs_s"], "far": cb["far_runs_s"]},
                       "single_thread": {"rho1": c1["rho_runs_s"], "far": c1["far_runs_s"]}}}
        # the GPU paths against the oracle on the sample
        if "fp64" in outs or "fast" in outs:
            ref = cb["rho_cpu"][:, 0]
            sub = cb["sub"]
            mp_sub = sub.i_valid_point_2_i_full_points_map[:sub.n_point_batches_prefix_sum[-1]]
            for p in outs:
                got = outs[p]["rho"].cpu().numpy()
                line[f"rho1_{p}_vs_oracle_rel"] = float(np.abs(got[mp_sub] - ref[mp_sub]).max() /
                                                        np.abs(ref[mp_sub]).max())
            mpf = fp.i_valid_point_2_i_full_points_map[:cb["nfs"]]
            dref = cb["dv_cpu"][mpf, 0]
            for p in outs:
                dg = outs[p]["dv"].cpu().numpy()[mpf]
                line[f"far_{p}_vs_oracle_rel"] = float(np.abs(dg - dref).max() / np.abs(dref).max())
    if a.h1_gga:
        # the 60k-batch inputs (and the step closures holding their plans) are released first: the
        # gradient and hessian waves are 9x the wave
        fns = None
        D.clear()
        import gc
        gc.collect()
        torch.cuda.empty_cache()
        try:
            line["h1_gga"] = bench_h1_gga(a.steps, a.warmup, a.no_cpu, a.cpu_threads)
            log(f"H1 GGA: {line['h1_gga']['ms']:.3f} ms, {line['h1_gga']['roofline']['achieved']:.1f} TF/s")
        except torch.cuda.OutOfMemoryError as e:
            log(f"H1 GGA skipped: {e}")
    if a.config3:
        D.clear()
        import gc
        gc.collect()
        torch.cuda.empty_cache()
        try:
            line["rho1_512atoms_precisions"] = bench_config3(a.steps, a.warmup, a.no_cpu, a.cpu_threads)
        except torch.cuda.OutOfMemoryError as e:
            log(f"512-atom precision leg skipped: {e}")
    leftover = child_processes()
    if leftover:
        line["child_processes_at_exit"] = leftover
    print(json.dumps(line), flush=True)


def child_processes():
    """Child processes of this one still alive at the end (their command lines; the pool of the
    oracle's OpenMP runtime is threads, not processes).  Each is ended and waited for, so that none
    outlives the bench."""
    try:
        import psutil
    except ImportError:
        return []
    kids = psutil.Process().children(recursive=True)
    out = []
    for k in kids:
        try:
            out.append(" ".join(k.cmdline())[:200])
            k.terminate()
        except psutil.Error:
            pass
    psutil.wait_procs(kids, timeout=5)
    if out:
        log(f"child processes at exit (ended): {out}")
    return out


C3_WORKLOAD = "rho1_512atoms_20kbatches_far200k_l6"


def bench_config3(steps, warmup, no_cpu, threads):
    """rho1 on the 512-atom, 20k-batch workload (points 50-200, n_compute 80-800 nearest basis
    functions, P1 = N(0,1) exp(-d/2 bohr) symmetric) in every precision: ms per call (the call captured
    in a HIP graph and replayed, HIP events on torch's current stream; each precision's launch
    autotuned), and per precision the error against the FP64 GPU path over all points: max |d rho1| / max
    |rho1|, the median over points with |rho1| > 1e-3 max of |d rho1(p)| / |rho1(p)|, and the maximum
    of |d rho1(p)| / cond(p) with cond(p) = sum_ij |phi_i(p)| |P1_ij| |phi_j(p)| (the FP64 path run on
    |phi| and |P1|).  With the GGA gradient (random gradient wave): FP64 against the INT8 split.  The
    FP64 path is checked against the CPU oracle on a sample of batches."""
    import torch

    from fast_rpa_eval_amd import api, synth
    from fast_rpa_eval_amd._lib import FrpaError
    w3 = WORKLOADS[C3_WORKLOAD]
    t0 = time.time()
    prob = synth.make_rho1_problem(w3["n_atoms"], w3["box_ang"], w3["n_batches"], w3["np_range"], w3["nc_range"],
                                   config_k=w3["config_k"], fill_wave=False)
    plan = api.Rho1Plan.from_problem(prob)
    wl = int(prob.wave_offset[-1])
    wave = plan.synth_wave(prob.atom_xyz, prob.basis, api.to_device(prob.coord_points), wl)
    g = torch.Generator(device="cuda").manual_seed(synth.seeds(w3["config_k"])["p1"])
    nb = prob.n_basis
    A = torch.randn(nb, nb, generator=g, device="cuda", dtype=torch.float64)
    A = torch.triu(A) + torch.triu(A, 1).T
    bx = torch.from_numpy(prob.atom_xyz[prob.basis.atom]).cuda()
    A *= torch.exp(-torch.cdist(bx, bx) / 2.0)
    P1 = A.contiguous().view(-1)
    del A, bx
    pts = int(prob.n_point_batches[prob.n_compute_c_batches > 0].sum())
    log(f"512-atom workload ready: {prob.n_batches} batches, {nb} basis, {pts} points, wave {wl * 8 / 1e9:.1f} GB "
        f"({time.time() - t0:.1f}s)")
    valid = torch.from_numpy(prob.valid_map().astype(np.int64)).cuda()
    out = torch.empty(prob.n_full_points, dtype=torch.float64, device="cuda")

    def timed(fn):
        # warm-up calls (conversions, work lists and scratch made there), then the call captured once
        # in a HIP graph and replayed: ms per replay, one persistent launch per call
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(max(warmup, 1)):
                fn()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            fn()
        graph.replay()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            graph.replay()
        e1.record()
        torch.cuda.synchronize()
        del graph
        return e0.elapsed_time(e1) / steps

    tuned = {}
    for prec in (api.FP64, api.FP64_SYM, api.OZAKI, api.FP32, api.BF16X3):
        tuned[prec] = plan.autotune(wave, P1, precision=prec, reps=2).split(";")[0]
    ref = plan.forward(wave, P1, precision=api.FP64)[valid].clone()
    cond = plan.forward(wave.abs(), P1.abs(), precision=api.FP64)[valid].clone()
    big = ref.abs() > 1e-3 * ref.abs().max()
    res = {"workload": f"{C3_WORKLOAD}: {prob.n_batches} batches, {nb} basis functions, n_compute 80-800, "
                       f"{pts} points", "points": pts,
           "launch": {"fp64": tuned[api.FP64], "fp64_symmetric": tuned[api.FP64_SYM], "int8_split_s5": tuned[api.OZAKI],
                      "fp32_mfma": tuned[api.FP32], "bf16x3": tuned[api.BF16X3]}, "precisions": {}}
    for name, prec in (("fp64", api.FP64), ("fp64_symmetric", api.FP64_SYM), ("int8_split_s5", api.OZAKI),
                       ("int8_split_s4", api.OZAKI4), ("fp32_mfma", api.FP32), ("bf16x3", api.BF16X3)):
        try:
            ms = timed(lambda: plan.forward(wave, P1, precision=prec, out=out))
        except FrpaError as e:
            res["precisions"][name] = {"error": str(e)}
            continue
        d = (out[valid] - ref).abs()
        res["precisions"][name] = {
            "ms": ms, "points_per_s": pts / (ms / 1e3), "max_rel_err": float(d.max() / ref.abs().max()),
            "median_pointwise_rel_err": float((d[big] / ref[big].abs()).median()),
            "max_cond_normalised_err": float((d / cond).max())}
        log(f"512 atoms, {name}: {res['precisions'][name]}")
    fp64_ms = res["precisions"]["fp64"]["ms"]
    for v in res["precisions"].values():
        if "ms" in v:
            v["speedup_vs_fp64"] = fp64_ms / v["ms"]
    # matrix-core rate of each precision's algorithmic GEMM work against that precision's dense peak
    # (FP64: the full product as the reference; the others multiply the upper tile pairs of the
    # symmetric part: FP64 symmetric one product, INT8 split S(S+1)/2 digit products, BF16x3 three
    # products, FP32 one)
    nc, npt = _nc_np(prob)
    tri = float(np.sum(nc * (nc + 1) / 2 * npt))
    mfma = {"fp64": (float(np.sum(2 * nc * nc * npt)), PEAK_FP64, "TFLOP/s"),
            "fp64_symmetric": (2 * tri, PEAK_FP64, "TFLOP/s"),
            "int8_split_s5": (15 * 2 * tri, PEAK_INT8, "TOP/s"), "int8_split_s4": (10 * 2 * tri, PEAK_INT8, "TOP/s"),
            "fp32_mfma": (2 * tri, PEAK_FP32, "TFLOP/s"), "bf16x3": (3 * 2 * tri, PEAK_BF16, "TFLOP/s")}
    for name, v in res["precisions"].items():
        if "ms" in v and name in mfma:
            ops, peak, unit = mfma[name]
            ach = ops / (v["ms"] / 1e3) / 1e12
            v["mfma_rate"] = {"achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak, "ops": ops}
    # GGA gradient: FP64 and the INT8 split (every tile pair of P1) on a random gradient wave
    gw = torch.empty(3 * wl, dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0, generator=g)
    gb = torch.zeros(3 * prob.n_max_batch_size * prob.n_batches, dtype=torch.float64, device="cuda")
    gout = torch.empty_like(gb)
    gga = {}
    for name, prec in (("fp64", api.FP64), ("int8_split_s5", api.OZAKI)):
        ms = timed(lambda: plan.forward(wave, P1, precision=prec, out=out, gradient_wave=gw, grad_bias=gb,
                                        grad_out=gout))
        gga[name] = {"ms": ms, "rho": out[valid].clone(), "grad": gout.clone()}
    r_ref, g_ref = gga["fp64"]["rho"], gga["fp64"]["grad"]
    res["with_gradient"] = {
        n: {"ms": v["ms"], "points_per_s": pts / (v["ms"] / 1e3), "speedup_vs_fp64": gga["fp64"]["ms"] / v["ms"],
            "rho1_max_rel_err": float((v["rho"] - r_ref).abs().max() / r_ref.abs().max()),
            "grad_max_rel_err": float((v["grad"] - g_ref).abs().max() / g_ref.abs().max())}
        for n, v in gga.items()}
    log(f"512 atoms with the gradient: {res['with_gradient']}")
    del gw, gb, gout, gga
    if not no_cpu:
        from oracle import oracle
        pick = np.sort(np.random.default_rng(3).choice(prob.n_batches, size=64, replace=False))
        sub = subproblem(prob, pick, wave)
        oracle.set_threads(threads)
        P1h = np.asfortranarray(P1.view(nb, nb).cpu().numpy().T)[:, :, None]
        rc, _ = oracle.rho1(sub, P1h)
        got = plan.forward(wave, P1, precision=api.FP64).cpu().numpy()
        mp = sub.valid_map()
        res["fp64_vs_oracle_rel"] = float(np.abs(got[mp] - rc[mp, 0]).max() / np.abs(rc[mp, 0]).max())
    return res


def summarize_profile(d):
    """Per-kernel table of a rocprofv3 CSV output directory: kernel stats (calls, average duration) and
    counter sums / per-dispatch means for each counter collected."""
    import csv
    import glob
    stats = sorted(glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True))
    ctrs = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))
    def short(n):
        # drop the namespace decoration and the argument list, keep template arguments
        n = n.replace("(anonymous namespace)::", "").replace("void ", "", 1)
        return n.split("(")[0].replace(" ", "")
    for f in stats:
        print(f"# kernel stats: {os.path.relpath(f, d)}")
        print(f"{'kernel':60s} {'calls':>6s} {'avg_ms':>10s} {'total_ms':>10s} {'pct':>6s}")
        for r in csv.DictReader(open(f)):
            print(f"{short(r['Name'])[:60]:60s} {int(r['Calls']):6d} {float(r['AverageNs']) / 1e6:10.3f} "
                  f"{float(r['TotalDurationNs']) / 1e6:10.3f} {float(r['Percentage']):6.2f}")
    for f in ctrs:
        print(f"# counters: {os.path.relpath(f, d)}")
        agg = {}
        for r in csv.DictReader(open(f)):
            k = (short(r["Kernel_Name"]), r["Counter_Name"])
            v = agg.setdefault(k, [0.0, 0])
            v[0] += float(r["Counter_Value"])
            v[1] += 1
        print(f"{'kernel':60s} {'counter':>12s} {'dispatches':>10s} {'sum':>16s} {'mean/dispatch':>16s}")
        for (kn, cn), (sm, n) in sorted(agg.items()):
            print(f"{kn[:60]:60s} {cn:>12s} {n:10d} {sm:16.1f} {sm / n:16.1f}")


def far_pairs(D):
    """Number of (point, centre) pairs in the far-field branch (host count on the inputs)."""
    import torch
    fp = D["fp"]
    pts = torch.from_numpy(fp.coord_points[:, fp.i_valid_point_2_i_full_points_map].T.copy()).cuda()
    cen = torch.from_numpy(fp.coords_center.T.copy()).cuda()
    n = 0
    for s in range(0, pts.shape[0], 200000):
        d2 = torch.cdist(pts[s:s + 200000], cen) ** 2
        n += int(((d2 >= 64.0) & (d2 < 1e8)).sum())
    return n


if __name__ == "__main__":
    main()
