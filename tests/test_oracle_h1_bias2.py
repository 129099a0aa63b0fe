"""CPU oracle of evaluate_first_order_h_bias_part2 (oracle/h1bias2.c).

Pins:
* With vrho = vsigma = 0 the operand is C(i, p) = w(p) H_times_psi(i, p), and the term added to H is
  L + L^T with L(i, k) = d/dR_j sum_p C(i, p) phi_k(p) for the functions k of atom j (they move with
  it: d phi_k / d R_j = -grad phi_k) -- checked by central differences of that sum under a
  displacement of atom j, with the synthetic basis' true values and gradients, every coordinate, a
  tile of three atoms (rows of atom j, sign, coordinate and the symmetric scatter).
* The vrho and vsigma prefactors: with H_times_psi = 0 and vrho = df/drho, vsigma(0, .) = df/dsigma of
  a model functional f(rho, sigma) = a rho^2 + b rho sigma + c sigma^2 (sigma = |grad rho|^2), the
  result equals the one with H_times_psi = d/de f(rho + e phi_i, |grad rho + e grad phi_i|^2) at e = 0
  (central differences in e) and vrho = vsigma = 0.
"""
import numpy as np

from fast_rpa_eval_amd import synth
from tests.test_oracle_bias2 import _true_wave_and_gradient


def _zeros_H(prob, T):
    return np.zeros((prob.n_basis, prob.n_basis, T), order="F")


def _fd_expected(prob, w_full, hpsi, j_atoms, j_coord, h=1e-5):
    """sum over batches of L + L^T, L(i, k) = central difference of sum_p w hpsi(i, p) phi_k(p) in R_j."""
    T = len(j_atoms)
    H = np.zeros((prob.n_basis, prob.n_basis, T))
    avn = prob.atom_valid_n_compute_c_batches()
    for b in range(prob.n_batches):
        nc, npt = int(prob.n_compute_c_batches[b]), int(prob.n_point_batches[b])
        if nc == 0 or npt == 0:
            continue
        ld = (nc + 3) // 4 * 4
        idx = prob.ibasis[prob.ibasis_ptr[b]:prob.ibasis_ptr[b + 1]] - 1
        mp = prob.i_valid_point_2_i_full_points_map[prob.n_point_batches_prefix_sum[b]:prob.n_point_batches_prefix_sum[b + 1]]
        pts = prob.coord_points[:, mp].T
        off = int(prob.wave_offset[b])
        C = hpsi[off:off + ld * npt].reshape(ld, npt, order="F")[:nc] * w_full[mp][None, :]
        for t, j in enumerate(j_atoms):
            v0, v1 = int(avn[j - 1, b]), int(avn[j, b])
            if v1 <= v0:
                continue
            vals = []
            for s in (1, -1):
                xyz = prob.atom_xyz.copy()
                xyz[j - 1, j_coord - 1] += s * h
                vals.append(synth.eval_basis(xyz, prob.basis, idx[v0:v1], pts))
            dphi = (vals[0] - vals[1]) / (2 * h)      # (|V|, np)
            L = C @ dphi.T                            # (nc, |V|)
            for kk, k in enumerate(range(v0, v1)):
                H[idx, idx[k], t] += L[:, kk]
                H[idx[k], idx, t] += L[:, kk]
    return H


def test_h_bias2_is_the_atom_displacement_derivative(oracle):
    prob = synth.make_rho1_problem(8, 8.0, 16, 40, (20, 80), config_k=1, fill_wave=False, zero_nc=(5,))
    wave, gwave = _true_wave_and_gradient(prob)
    prob.wave = wave
    rng = np.random.default_rng(31)
    w_full = rng.uniform(0.1, 1.0, prob.n_full_points)
    hpsi = rng.standard_normal(len(wave))
    nmb, nb = prob.n_max_batch_size, prob.n_batches
    z = np.zeros((nmb, nb))
    z3 = np.zeros((3, nmb, nb))
    T, jb = 3, 2
    for jc in (1, 2, 3):
        H = oracle.h1_bias2(prob, T, jb, jb + T - 1, jc, _zeros_H(prob, T), z, z3, np.zeros((3, prob.n_full_points)),
                            hpsi, gwave, w_full)
        ref = _fd_expected(prob, w_full, hpsi, list(range(jb, jb + T)), jc)
        err = np.abs(H - ref).max() / np.abs(ref).max()
        print(f"[H bias part 2 FD j_coord={jc}] rel {err:.2e}")
        assert err <= 1e-7
        assert np.allclose(H, np.transpose(H, (1, 0, 2)))


def test_h_bias2_xc_prefactors_are_the_functional_derivative(oracle):
    prob = synth.make_rho1_problem(10, 10.0, 20, (10, 50), (10, 60), config_k=2, zero_nc=(3,))
    rng = np.random.default_rng(5)
    gw = synth.random_like_wave(prob, 3, 6)
    nf, nmb, nb = prob.n_full_points, prob.n_max_batch_size, prob.n_batches
    rho = rng.uniform(0.5, 2.0, nf)
    grho = rng.uniform(-1, 1, (3, nf))
    sig = np.sum(grho ** 2, axis=0)
    a, bb, c = 0.7, -0.3, 0.2
    w_full = rng.uniform(0.1, 1.0, nf)
    vrho = np.zeros((nmb, nb))
    vsig = rng.uniform(-1, 1, (3, nmb, nb))  # components 1, 2 are not read
    hpsi_model = np.zeros(len(prob.wave))
    eps = 1e-6
    f = lambda r, s: a * r * r + bb * r * s + c * s * s
    for b in range(nb):
        nc, npt = int(prob.n_compute_c_batches[b]), int(prob.n_point_batches[b])
        mp = prob.i_valid_point_2_i_full_points_map[prob.n_point_batches_prefix_sum[b]:prob.n_point_batches_prefix_sum[b + 1]]
        vrho[:npt, b] = 2 * a * rho[mp] + bb * sig[mp]
        vsig[0, :npt, b] = bb * rho[mp] + 2 * c * sig[mp]
        if nc == 0:
            continue
        ld = (nc + 3) // 4 * 4
        off = int(prob.wave_offset[b])
        phi = prob.wave[off:off + ld * npt].reshape(ld, npt, order="F")
        dphi = gw[3 * off:3 * off + 3 * ld * npt].reshape(ld, 3, npt, order="F")
        g = grho[:, mp]                                              # (3, np)
        plus = f(rho[mp][None, :] + eps * phi, np.sum((g[None] + eps * dphi) ** 2, axis=1))
        minus = f(rho[mp][None, :] - eps * phi, np.sum((g[None] - eps * dphi) ** 2, axis=1))
        blk = (plus - minus) / (2 * eps)
        blk[nc:] = 0.0
        hpsi_model[off:off + ld * npt] = blk.ravel(order="F")
    T, jb, jc = 2, 3, 2
    HA = oracle.h1_bias2(prob, T, jb, jb + 1, jc, _zeros_H(prob, T), vrho, vsig, grho, np.zeros(len(prob.wave)), gw,
                         w_full)
    HB = oracle.h1_bias2(prob, T, jb, jb + 1, jc, _zeros_H(prob, T), np.zeros((nmb, nb)), np.zeros((3, nmb, nb)),
                         np.zeros((3, nf)), hpsi_model, gw, w_full)
    err = np.abs(HA - HB).max() / np.abs(HB).max()
    print(f"[H bias part 2 vrho/vsigma vs functional derivative] rel {err:.2e}")
    assert np.abs(HB).max() > 0 and err <= 1e-8
