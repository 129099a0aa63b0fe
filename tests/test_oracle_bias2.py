"""CPU oracle of generate_first_order_rho_bias_part2 (oracle/bias2.c).

Pin: the bias is the derivative of rho(p) = sum_ik phi_i(p) DM_ik phi_k(p) (the batch's compute
functions) with respect to displacing atom j along j_coord -- atom j's functions move with it, so
d phi_i / d R_j = -grad phi_i for i on atom j.  With the true basis values and gradients of the
synthetic basis, the oracle must match central differences of rho under that displacement, for a
symmetric and a nonsymmetric DM, every coordinate, a tile of three atoms.  Also: atoms without
functions in a batch give exactly 0, batches with n_compute_c = 0 are left untouched.
"""
import numpy as np
import pytest

from fast_rpa_eval_amd import synth


def _true_wave_and_gradient(prob, h=1e-5):
    """phi and grad phi (central differences in the point coordinates, h = 1e-5) in the
    wave_batches_compress layouts (nc_pad, np) and (nc_pad, 3, np); padding rows 0."""
    wl = int(prob.wave_offset[-1])
    wave = np.zeros(wl)
    gwave = np.zeros(3 * wl)
    for b in range(prob.n_batches):
        nc, npt = int(prob.n_compute_c_batches[b]), int(prob.n_point_batches[b])
        if nc == 0 or npt == 0:
            continue
        ld = (nc + 3) // 4 * 4
        idx = prob.ibasis[prob.ibasis_ptr[b]:prob.ibasis_ptr[b + 1]] - 1
        mp = prob.i_valid_point_2_i_full_points_map[prob.n_point_batches_prefix_sum[b]:prob.n_point_batches_prefix_sum[b + 1]]
        pts = prob.coord_points[:, mp].T
        off = int(prob.wave_offset[b])
        w = np.zeros((ld, npt))
        w[:nc] = synth.eval_basis(prob.atom_xyz, prob.basis, idx, pts)
        wave[off:off + ld * npt] = w.ravel(order="F")
        g = np.zeros((ld, 3, npt))
        for c in range(3):
            e = np.zeros(3)
            e[c] = h
            g[:nc, c] = (synth.eval_basis(prob.atom_xyz, prob.basis, idx, pts + e) -
                         synth.eval_basis(prob.atom_xyz, prob.basis, idx, pts - e)) / (2 * h)
        gwave[3 * off:3 * off + 3 * ld * npt] = g.ravel(order="F")
    return wave, gwave


def _fd_bias(prob, DM, j_atom, j_coord, h=1e-5):
    """Central difference of rho(p) = phi^T DM phi under R_j += h e_c, as (n_max_batch_size, n_batches)."""
    out = np.zeros((prob.n_max_batch_size, prob.n_batches))
    for b in range(prob.n_batches):
        nc, npt = int(prob.n_compute_c_batches[b]), int(prob.n_point_batches[b])
        if nc == 0 or npt == 0:
            continue
        idx = prob.ibasis[prob.ibasis_ptr[b]:prob.ibasis_ptr[b + 1]] - 1
        mp = prob.i_valid_point_2_i_full_points_map[prob.n_point_batches_prefix_sum[b]:prob.n_point_batches_prefix_sum[b + 1]]
        pts = prob.coord_points[:, mp].T
        dm = DM[np.ix_(idx, idx)]
        vals = []
        for s in (1, -1):
            xyz = prob.atom_xyz.copy()
            xyz[j_atom - 1, j_coord - 1] += s * h
            phi = synth.eval_basis(xyz, prob.basis, idx, pts)
            vals.append(np.einsum("ip,ik,kp->p", phi, dm, phi))
        out[:npt, b] = (vals[0] - vals[1]) / (2 * h)
    return out


@pytest.mark.parametrize("symmetric", [True, False])
def test_bias2_is_the_atom_displacement_derivative(oracle, symmetric):
    prob = synth.make_rho1_problem(8, 8.0, 16, 40, (20, 80), config_k=1, fill_wave=False, zero_nc=(5,))
    wave, gwave = _true_wave_and_gradient(prob)
    prob.wave = wave
    rng = np.random.default_rng(2)
    A = rng.standard_normal((prob.n_basis, prob.n_basis))
    DM = (A + A.T) / 2 if symmetric else A
    T, jb = 3, 2
    for jc in (1, 2, 3):
        init = np.full((prob.n_max_batch_size, prob.n_batches, T), 7.0, order="F")
        got = oracle.bias2(prob, DM, gwave, T, jb, j_coord=jc, bias=init)
        assert np.all(got[:, 5, :] == 7.0)  # n_compute_c = 0: untouched
        for t in range(T):
            ref = _fd_bias(prob, DM, jb + t, jc)
            sel = np.ones(prob.n_batches, bool)
            sel[5] = False
            g = got[:, sel, t].copy()
            for b in np.nonzero(sel)[0]:
                g[prob.n_point_batches[b]:, list(np.nonzero(sel)[0]).index(b)] = 0.0
            err = np.abs(g - ref[:, sel]).max() / np.abs(ref).max()
            print(f"[bias2 FD symmetric={symmetric} j={jb + t} c={jc}] rel {err:.2e}")
            assert err <= 1e-7


def test_bias2_empty_atom_rows_give_zero(oracle):
    prob = synth.make_rho1_problem(8, 12.0, 12, 30, (10, 30), config_k=1)
    avn = prob.atom_valid_n_compute_c_batches()
    gw = synth.random_like_wave(prob, 3, 5)
    DM = np.random.default_rng(4).standard_normal((prob.n_basis, prob.n_basis))
    got = oracle.bias2(prob, DM, gw, 8, 1, bias=np.full((prob.n_max_batch_size, prob.n_batches, 8), 3.0, order="F"))
    n_empty = 0
    for b in range(prob.n_batches):
        for t in range(8):
            npt = prob.n_point_batches[b]
            if avn[t + 1, b] == avn[t, b]:
                n_empty += 1
                assert np.all(got[:npt, b, t] == 0.0)
            else:
                assert np.all(got[:npt, b, t] != 0.0)
            assert np.all(got[npt:, b, t] == 3.0)
    assert n_empty > 0


def _fd_bias_one_sided(prob, DM, j_atom, j_coord, h=1e-5):
    """Central difference of sum_ik phi_i(R_j + h e_c) DM_ik phi_k(R) -- only the first factor moves."""
    out = np.zeros((prob.n_max_batch_size, prob.n_batches))
    for b in range(prob.n_batches):
        nc, npt = int(prob.n_compute_c_batches[b]), int(prob.n_point_batches[b])
        if nc == 0 or npt == 0:
            continue
        idx = prob.ibasis[prob.ibasis_ptr[b]:prob.ibasis_ptr[b + 1]] - 1
        mp = prob.i_valid_point_2_i_full_points_map[prob.n_point_batches_prefix_sum[b]:prob.n_point_batches_prefix_sum[b + 1]]
        pts = prob.coord_points[:, mp].T
        phi0 = synth.eval_basis(prob.atom_xyz, prob.basis, idx, pts)
        w = DM[np.ix_(idx, idx)] @ phi0
        vals = []
        for s in (1, -1):
            xyz = prob.atom_xyz.copy()
            xyz[j_atom - 1, j_coord - 1] += s * h
            vals.append(np.sum(synth.eval_basis(xyz, prob.basis, idx, pts) * w, axis=0))
        out[:npt, b] = (vals[0] - vals[1]) / (2 * h)
    return out


def test_inline_bias_is_the_one_sided_displacement_derivative(oracle):
    """The v6 inline bias: d/dR_j of sum_ik phi_i DM_ik phi_k with only phi_i moving (nonsymmetric DM,
    so the side matters); ADDED to the given bias; n_compute_c = 0 batches untouched."""
    prob = synth.make_rho1_problem(8, 8.0, 16, 40, (20, 80), config_k=1, fill_wave=False, zero_nc=(5,))
    wave, gwave = _true_wave_and_gradient(prob)
    prob.wave = wave
    DM = np.random.default_rng(12).standard_normal((prob.n_basis, prob.n_basis))
    T, jb = 3, 2
    for jc in (1, 3):
        init = np.full((prob.n_max_batch_size, prob.n_batches, T), 7.0, order="F")
        got = oracle.bias_inline(prob, DM, gwave, T, jb, j_coord=jc, bias=init) - 7.0
        assert np.all(got[:, 5, :] == 0.0)
        for t in range(T):
            ref = _fd_bias_one_sided(prob, DM, jb + t, jc)
            g = got[:, :, t].copy()
            for b in range(prob.n_batches):
                g[prob.n_point_batches[b]:, b] = 0.0
            err = np.abs(g - ref).max() / np.abs(ref).max()
            print(f"[inline bias FD j={jb + t} c={jc}] rel {err:.2e}")
            assert err <= 1e-7
        # the other side of DM is a different function (DM nonsymmetric)
        other = oracle.bias_inline(prob, DM.T, gwave, T, jb, j_coord=jc)
        assert np.abs(other[:, :, 0] - got[:, :, 0]).max() > 1e-3 * np.abs(got).max()


def test_inline_bias_two_sides_make_bias_part2(oracle):
    """inline(DM) + inline(DM^T) = bias part 2 of DM (the two terms of the symmetric derivative),
    with random wave and gradient values and a short tile (j_atom_end < j_atom_begin + T - 1)."""
    prob = synth.make_rho1_problem(10, 10.0, 20, (10, 50), (10, 60), config_k=2, zero_nc=(3,))
    gw = synth.random_like_wave(prob, 3, 8)
    DM = np.random.default_rng(9).standard_normal((prob.n_basis, prob.n_basis))
    T, jb, je = 4, 3, 5
    a = oracle.bias_inline(prob, DM, gw, T, jb, j_atom_end=je, j_coord=2)
    both = oracle.bias_inline(prob, DM.T, gw, T, jb, j_atom_end=je, j_coord=2, bias=a)
    ref = oracle.bias2(prob, DM, gw, T, jb, j_atom_end=je, j_coord=2)
    err = np.abs(both - ref).max() / np.abs(ref).max()
    print(f"[inline bias two sides vs part 2] rel {err:.2e}")
    assert err <= 1e-13
    assert np.all(both[:, :, je - jb + 1:] == 0.0)
