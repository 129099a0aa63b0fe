"""Host-side mirror of the reference interface for the rho1 and far-field Hartree paths.

Two levels, both calling libfrpa.so through its C-ABI:

* ``Rho1Plan`` / ``FarPlan``: device-resident metadata plus ``forward`` / ``adjoint`` methods taking
  torch tensors already in HBM (the measured path).  Argument names and layouts are those of the
  reference's Fortran-ABI arguments (fast_rap_src_dfpt_gpu_eval/evaluate_first_order_rho.h:3-35,
  sum_up_whole_potential.h:3-48); Fortran-ordered arrays are passed as flat tensors in memory order.
* The reference-named entry points (init_dfpt_device_info, init_dfpt_device_data,
  evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_,
  sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_, ...) are exported by the same
  library and called with host arrays, as a Fortran/C++ host calls them; tests/test_gpu_reference_abi.py
  drives them through ctypes.

torch provides device memory and streams only; all arithmetic runs in libfrpa's HIP kernels.
"""
import ctypes
import weakref

import numpy as np
import torch

from . import _lib
from ._lib import FrpaError, check, lib  # noqa: F401  (FrpaError re-exported)

FP64, OZAKI, OZAKI4 = _lib.FRPA_PREC_FP64, _lib.FRPA_PREC_OZAKI, _lib.FRPA_PREC_OZAKI4
FP32, BF16X3 = _lib.FRPA_PREC_FP32, _lib.FRPA_PREC_BF16X3
FP64_SYM = _lib.FRPA_PREC_FP64_SYM
FAR_TERMS, FAR_HORNER = _lib.FRPA_FAR_TERMS, _lib.FRPA_FAR_HORNER


def _hp(a, dtype):
    """Host pointer of a contiguous numpy array (kept alive by the caller)."""
    assert a.dtype == dtype, (a.dtype, dtype)
    assert a.flags.c_contiguous or a.flags.f_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def _dp(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "device arrays must be contiguous CUDA tensors"
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def to_device(a):
    """Flat device copy of a numpy array in memory order (Fortran arrays keep their layout)."""
    a = np.asarray(a)
    order = "F" if (a.flags.f_contiguous and not a.flags.c_contiguous) else "C"
    return torch.from_numpy(np.ascontiguousarray(a.ravel(order=order))).cuda()


class Rho1Plan:
    """Batch metadata of evaluate_first_order_rho, resident on the device."""

    def __init__(self, n_basis, n_full_points, n_max_batch_size, n_point_batches, n_point_batches_prefix_sum,
                 i_valid_point_2_i_full_points_map, n_compute_c_batches, ibasis_ptr, ibasis, wave_offset):
        self._keep = [np.ascontiguousarray(n_point_batches, np.int32),
                      np.ascontiguousarray(n_point_batches_prefix_sum, np.int32),
                      np.ascontiguousarray(i_valid_point_2_i_full_points_map, np.int32),
                      np.ascontiguousarray(n_compute_c_batches, np.int32),
                      np.ascontiguousarray(ibasis_ptr, np.int64), np.ascontiguousarray(ibasis, np.int32),
                      np.ascontiguousarray(wave_offset, np.int64)]
        k = self._keep
        self.n_batches = len(k[0])
        self.n_basis, self.n_full_points, self.n_max_batch_size = n_basis, n_full_points, n_max_batch_size
        # per converted wave (the INT8 digits; the FP32 planes; the BF16x3 planes): (weakref, version,
        # address) of the wave it was made from
        self._wave_seen = {}
        self.handle = ctypes.c_void_p()
        check(lib().frpa_rho1_plan_create(
            ctypes.byref(self.handle), self.n_batches, n_basis, n_full_points, n_max_batch_size,
            _hp(k[0], np.int32), _hp(k[1], np.int32), _hp(k[2], np.int32), _hp(k[3], np.int32),
            _hp(k[4], np.int64), _hp(k[5], np.int32), _hp(k[6], np.int64)), "frpa_rho1_plan_create")

    @classmethod
    def from_problem(cls, prob):
        return cls(prob.n_basis, prob.n_full_points, prob.n_max_batch_size, prob.n_point_batches,
                   prob.n_point_batches_prefix_sum, prob.i_valid_point_2_i_full_points_map,
                   prob.n_compute_c_batches, prob.ibasis_ptr, prob.ibasis, prob.wave_offset)

    def __del__(self):
        if getattr(self, "handle", None) and lib is not None:
            try:
                lib().frpa_rho1_plan_destroy(self.handle)
            except Exception:
                pass
            self.handle = None

    def _follow_wave(self, wave, precision, stream=None):
        """The library keeps INT8 digits (and the FP32 / BF16x3 planes) of a wave by its address; a
        different tensor at the same address (a freed one's memory reused) or an in-place change (a new
        version) needs them rebuilt, which frpa_rho1_prepare_wave does for the forward and the adjoint."""
        if precision in (OZAKI, OZAKI4):
            fam = "int8"  # one digit set at a time; the library converts again when S changes
        elif precision in (FP32, BF16X3):
            fam = precision  # a converted wave per precision
        else:
            return
        seen = self._wave_seen.get(fam)
        if seen is not None and (seen[0]() is not wave or seen[1] != wave._version or seen[2] != wave.data_ptr()):
            self.prepare_wave(wave, precision, stream)
        self._wave_seen[fam] = (weakref.ref(wave), wave._version, wave.data_ptr())

    def forward(self, wave, density_matrix, T=1, n_tile=None, precision=FP64, rho_bias=None,
                gradient_wave=None, grad_bias=None, out=None, grad_out=None, stream=None):
        """first_order_rho (n_full_points * T,) [and first_order_gradient_rho] on the device."""
        self._follow_wave(wave, precision, stream)
        n_tile = T if n_tile is None else n_tile
        dev = wave.device
        if out is None:
            out = torch.empty(self.n_full_points * T, dtype=torch.float64, device=dev)
        if gradient_wave is not None and grad_out is None:
            grad_out = torch.empty(3 * self.n_max_batch_size * self.n_batches * T, dtype=torch.float64, device=dev)
        check(lib().frpa_rho1_forward(self.handle, T, n_tile, precision, _dp(wave), _dp(density_matrix),
                                      _dp(rho_bias), _dp(out), _dp(gradient_wave), _dp(grad_bias),
                                      _dp(grad_out if gradient_wave is not None else None), _stream(stream)),
              "frpa_rho1_forward")
        return (out, grad_out) if gradient_wave is not None else out

    def adjoint(self, wave, g, T=1, precision=FP64, H=None, stream=None):
        """first_order_H (n_basis^2 * T,) += sum_p g(map(p), t) phi_i phi_j."""
        self._follow_wave(wave, precision, stream)
        if H is None:
            H = torch.zeros(self.n_basis * self.n_basis * T, dtype=torch.float64, device=wave.device)
        check(lib().frpa_rho1_adjoint(self.handle, T, precision, _dp(wave), _dp(g), _dp(H), _stream(stream)),
              "frpa_rho1_adjoint")
        return H

    def set_adjoint_wait_limit(self, polls=0):
        """Bound (in polls) of the INT8 adjoint's slot waits; 0 = the default (see frpa.h)."""
        check(lib().frpa_rho1_set_adjoint_wait_limit(self.handle, polls), "frpa_rho1_set_adjoint_wait_limit")

    def adjoint_status(self, stream=None):
        """Raises FrpaError if an INT8 adjoint slot wait gave up since the last check (graph replays
        included); synchronises the stream."""
        check(lib().frpa_rho1_adjoint_status(self.handle, _stream(stream)), "frpa_rho1_adjoint_status")

    def autotune(self, wave, density_matrix, precision=FP64, T=1, reps=3, out=None, stream=None):
        """Time the launch configurations of forward() for this precision on these inputs and keep the
        fastest in the plan; returns the per-candidate timings, the chosen one first."""
        if out is None:
            out = torch.empty(self.n_full_points * T, dtype=torch.float64, device=wave.device)
        buf = ctypes.create_string_buffer(2048)
        check(lib().frpa_rho1_autotune(self.handle, T, precision, _dp(wave), _dp(density_matrix), _dp(out), reps,
                                       buf, len(buf), _stream(stream)), "frpa_rho1_autotune")
        return buf.value.decode()

    def set_launch(self, precision, order=0, persistent=0, points_per_wave=0):
        """Choose the launch configuration of forward() for one precision (see frpa_rho1_set_launch)."""
        check(lib().frpa_rho1_set_launch(self.handle, precision, order, persistent, points_per_wave),
              "frpa_rho1_set_launch")

    def bias_part2(self, wave, gradient_wave, density_matrix, atom_valid_n_compute_c_batches, n_atoms,
                   j_atom_begin, n_tile=1, j_coord=1, T=None, out=None, stream=None):
        """first_order_rho_bias_part2 (n_max_batch_size, n_batches, T) for the atoms j_atom_begin ..
        j_atom_begin + n_tile - 1 (generate_first_order_rho_bias_part2.cpp:9-167); device tensors,
        atom_valid_n_compute_c_batches int32 (n_atoms + 1, n_batches).  Entries of batches with
        n_compute_c = 0 or n_points = 0 keep `out`'s values (zeros if out is None)."""
        T = n_tile if T is None else T
        if out is None:
            out = torch.zeros(self.n_max_batch_size * self.n_batches * T, dtype=torch.float64, device="cuda")
        check(lib().frpa_rho1_bias_part2(
            self.handle, T, j_atom_begin, n_tile, j_coord, n_atoms, _dp(atom_valid_n_compute_c_batches), _dp(wave),
            _dp(gradient_wave), _dp(density_matrix), _dp(out), _stream(stream)), "frpa_rho1_bias_part2")
        return out

    def bias_inline(self, wave, gradient_wave, density_matrix, atom_valid_n_compute_c_batches, n_atoms,
                    j_atom_begin, bias, n_tile=1, j_coord=1, T=None, stream=None):
        """The v6 variant's inline bias -sum_{i in V} grad phi_i sum_k DM_ik phi_k
        (version6_rho_with_bias_calculate/evaluate_first_order_rho.cpp:179-231), ADDED in place to the
        device tensor `bias` (n_max_batch_size * n_batches * T); returns bias."""
        T = n_tile if T is None else T
        check(lib().frpa_rho1_bias_inline(
            self.handle, T, j_atom_begin, n_tile, j_coord, n_atoms, _dp(atom_valid_n_compute_c_batches), _dp(wave),
            _dp(gradient_wave), _dp(density_matrix), _dp(bias), _stream(stream)), "frpa_rho1_bias_inline")
        return bias

    def h1_bias_part2(self, wave, gradient_wave, H_times_psi, vrho, vsigma, gradient_rho, partition_tab,
                      atom_valid_n_compute_c_batches, n_atoms, j_atom_begin, n_tile=1, j_coord=1, T=None, H=None,
                      stream=None):
        """Bias part 2 of the first-order H (evaluate_first_order_H_bias_part2.cpp:13-218), added into H
        (n_basis^2 * T, zeros if None); device tensors in the reference's layouts (see
        frpa_rho1_h1_bias_part2)."""
        T = n_tile if T is None else T
        if H is None:
            H = torch.zeros(self.n_basis * self.n_basis * T, dtype=torch.float64, device="cuda")
        check(lib().frpa_rho1_h1_bias_part2(
            self.handle, T, j_atom_begin, n_tile, j_coord, n_atoms, _dp(atom_valid_n_compute_c_batches), _dp(wave),
            _dp(gradient_wave), _dp(H_times_psi), _dp(vrho), _dp(vsigma), _dp(gradient_rho), _dp(partition_tab), _dp(H),
            _stream(stream)), "frpa_rho1_h1_bias_part2")
        return H

    def h1(self, wave, gradient_wave, hessian_wave, xc, atom_valid_n_compute_c_batches, n_atoms, j_atom_begin,
           n_tile=1, j_coord=1, T=None, H=None, stream=None):
        """Full first-order H (LDA + GGA terms of evaluate_first_order_H), accumulated into H
        (n_basis^2 * T, zeros if None).  xc: dict of device tensors vsigma, gradient_rho, H_prefactor1,
        grad_rho1, v_hartree_gradient, first_order_rho, v2rho2, v2rhosigma, v2sigma2, partition_tab in the
        reference's layouts (see frpa_rho1_h1)."""
        T = n_tile if T is None else T
        if H is None:
            H = torch.zeros(self.n_basis * self.n_basis * T, dtype=torch.float64, device="cuda")
        check(lib().frpa_rho1_h1(
            self.handle, T, j_atom_begin, n_tile, j_coord, n_atoms, _dp(atom_valid_n_compute_c_batches), _dp(wave),
            _dp(gradient_wave), _dp(hessian_wave), _dp(xc["vsigma"]), _dp(xc["gradient_rho"]), _dp(xc["H_prefactor1"]),
            _dp(xc["grad_rho1"]), _dp(xc["v_hartree_gradient"]), _dp(xc["first_order_rho"]), _dp(xc["v2rho2"]),
            _dp(xc["v2rhosigma"]), _dp(xc["v2sigma2"]), _dp(xc["partition_tab"]), _dp(H), _stream(stream)),
            "frpa_rho1_h1")
        return H

    def prepare_wave(self, wave, precision=OZAKI, stream=None):
        b = ctypes.c_longlong(0)
        check(lib().frpa_rho1_prepare_wave(self.handle, precision, _dp(wave), ctypes.byref(b), _stream(stream)),
              "frpa_rho1_prepare_wave")
        return b.value

    def synth_wave(self, atom_xyz, basis, coord_points_dev, n_wave, stream=None):
        """Generate wave_batches_compress on the device (synthetic basis, see synth.py)."""
        axyz = np.ascontiguousarray(atom_xyz, np.float64)
        ba = np.ascontiguousarray(basis.atom, np.int32)
        bl = np.ascontiguousarray(basis.l, np.int32)
        bm = np.ascontiguousarray(basis.m, np.int32)
        bal = np.ascontiguousarray(basis.alpha, np.float64)
        wave = torch.empty(n_wave, dtype=torch.float64, device=coord_points_dev.device)
        check(lib().frpa_synth_wave(self.handle, len(axyz), _hp(axyz, np.float64), _hp(ba, np.int32),
                                    _hp(bl, np.int32), _hp(bm, np.int32), _hp(bal, np.float64),
                                    _dp(coord_points_dev), _dp(wave), _stream(stream)), "frpa_synth_wave")
        return wave


class FarPlan:
    """Geometry and term tables of the far-field Hartree sum, resident on the device."""

    def __init__(self, l_pot_max, l_max_analytic_multipole, index_cc, cc, n_cc_lm_ijk, index_ijk_max_cc,
                 species, l_hartree, centers_hartree_potential, center_to_atom, coords_center,
                 hartree_force_l_add=1):
        k = self._keep = dict(
            species=np.ascontiguousarray(species, np.int32), l_hartree=np.ascontiguousarray(l_hartree, np.int32),
            chp=np.ascontiguousarray(centers_hartree_potential, np.int32),
            c2a=np.ascontiguousarray(center_to_atom, np.int32),
            cc_xyz=np.asfortranarray(coords_center, np.float64), index_cc=np.asfortranarray(index_cc, np.int32),
            cc=np.ascontiguousarray(cc, np.float64), ncc=np.ascontiguousarray(n_cc_lm_ijk, np.int32),
            ijk=np.asfortranarray(index_ijk_max_cc, np.int32))
        self.l_pot_max = l_pot_max
        self.n_atoms = len(k["species"])
        self.handle = ctypes.c_void_p()
        check(lib().frpa_far_plan_create(
            ctypes.byref(self.handle), l_pot_max, l_max_analytic_multipole, k["index_cc"].shape[0],
            k["cc_xyz"].shape[1], len(k["chp"]), self.n_atoms, len(k["l_hartree"]), hartree_force_l_add,
            _hp(k["species"], np.int32), _hp(k["l_hartree"], np.int32), _hp(k["chp"], np.int32),
            _hp(k["c2a"], np.int32), _hp(k["cc_xyz"], np.float64), _hp(k["index_cc"], np.int32),
            _hp(k["cc"], np.float64), _hp(k["ncc"], np.int32), _hp(k["ijk"], np.int32)), "frpa_far_plan_create")

    @classmethod
    def from_problem(cls, fp):
        t = fp.tables
        plan = cls(fp.l_pot_max, fp.l_max_analytic_multipole, t["index_cc"], t["cc"], t["n_cc_lm_ijk"],
                   t["index_ijk_max_cc"], fp.species, fp.l_hartree, fp.centers_hartree_potential, fp.center_to_atom,
                   fp.coords_center, fp.hartree_force_l_add)
        if fp.near is not None:
            nf = fp.near
            plan.set_near(nf["species_center"], nf["n_grid"], nf["r_grid_min"], nf["log_r_grid_inc"])
        return plan

    def set_near(self, species_center, n_grid, r_grid_min, log_r_grid_inc):
        """Near-field log grids: species_center (n_centers) 1-based, per-species n_grid, r_grid_min,
        log_r_grid_inc (host arrays, copied)."""
        a = [np.ascontiguousarray(species_center, np.int32), np.ascontiguousarray(n_grid, np.int32),
             np.ascontiguousarray(r_grid_min, np.float64), np.ascontiguousarray(log_r_grid_inc, np.float64)]
        check(lib().frpa_far_plan_set_near(self.handle, len(a[1]), _hp(a[0], np.int32), _hp(a[1], np.int32),
                                           _hp(a[2], np.float64), _hp(a[3], np.float64)), "frpa_far_plan_set_near")

    def near_forward(self, delta_v_hartree, coord_points, i_valid_point_2_i_full_points_map, multipole_radius_sq,
                     outer_potential_radius, spl_tile, n_full_points, n_hartree_grid, T=1, n_tile=None, stream=None):
        """delta_v_hartree (n_full_points * T,) += near-field spline sum (in place; returned).
        spl_tile: current_delta_v_hart_part_spl_tile ((l_pot_max+1)^2, 2, n_hartree_grid, n_atoms, T)."""
        n_tile = T if n_tile is None else n_tile
        check(lib().frpa_near_forward(
            self.handle, T, n_tile, n_full_points, i_valid_point_2_i_full_points_map.numel(), n_hartree_grid,
            _dp(coord_points), _dp(i_valid_point_2_i_full_points_map), _dp(multipole_radius_sq),
            _dp(outer_potential_radius), _dp(spl_tile), _dp(delta_v_hartree), _stream(stream)), "frpa_near_forward")
        return delta_v_hartree

    def __del__(self):
        if getattr(self, "handle", None) and lib is not None:
            try:
                lib().frpa_far_plan_destroy(self.handle)
            except Exception:
                pass
            self.handle = None

    def forward(self, delta_v_hartree, coord_points, i_valid_point_2_i_full_points_map, l_hartree_max_far_distance,
                multipole_radius_sq, outer_potential_radius, multipole_moments, n_full_points, T=1, n_tile=None,
                variant=FAR_HORNER, stream=None):
        """delta_v_hartree (n_full_points * T,) += far-field sum (in place; returned)."""
        n_tile = T if n_tile is None else n_tile
        check(lib().frpa_far_forward(
            self.handle, T, n_tile, variant, n_full_points, i_valid_point_2_i_full_points_map.numel(),
            _dp(coord_points), _dp(i_valid_point_2_i_full_points_map), _dp(l_hartree_max_far_distance),
            _dp(multipole_radius_sq), _dp(outer_potential_radius), _dp(multipole_moments), _dp(delta_v_hartree),
            _stream(stream)), "frpa_far_forward")
        return delta_v_hartree

    def adjoint(self, gV, coord_points, i_valid_point_2_i_full_points_map, l_hartree_max_far_distance,
                multipole_radius_sq, outer_potential_radius, multipole_moments, n_full_points, T=1, n_tile=None,
                grad_moments=None, stream=None):
        n_tile = T if n_tile is None else n_tile
        if grad_moments is None:
            grad_moments = torch.zeros_like(multipole_moments)
        check(lib().frpa_far_adjoint(
            self.handle, T, n_tile, n_full_points, i_valid_point_2_i_full_points_map.numel(), _dp(coord_points),
            _dp(i_valid_point_2_i_full_points_map), _dp(l_hartree_max_far_distance), _dp(multipole_radius_sq),
            _dp(outer_potential_radius), _dp(multipole_moments), _dp(gV), _dp(grad_moments), _stream(stream)),
            "frpa_far_adjoint")
        return grad_moments


class FarInputs:
    """Per-call far-field arrays of a synth.FarProblem, uploaded once."""

    def __init__(self, fp):
        self.n_full_points = fp.n_full_points
        self.T = fp.n_tile
        self.coord_points = to_device(fp.coord_points)
        self.map = to_device(fp.i_valid_point_2_i_full_points_map)
        self.lmfd = to_device(fp.l_hartree_max_far_distance)
        self.mrsq = to_device(fp.multipole_radius_sq)
        self.opr = to_device(fp.outer_potential_radius)
        self.moments = to_device(fp.multipole_moments)
        self.spl = to_device(fp.near["spl"]) if fp.near is not None else None
        self.n_hartree_grid = fp.near["n_hartree_grid"] if fp.near is not None else 0

    def args(self):
        return (self.coord_points, self.map, self.lmfd, self.mrsq, self.opr, self.moments, self.n_full_points)

    def near_args(self):
        return (self.coord_points, self.map, self.mrsq, self.opr, self.spl, self.n_full_points, self.n_hartree_grid)
