"""The reference's binary test dumps (first_order_rho.meta / sumup .test.nproc_<n>.bin): byte layout,
round trip, malformed files, and that a problem written and read back gives the oracle the same
results (CPU).  The reference's own dumps are Git-LFS stubs, so the layout is pinned against the
writer's field order in evaluate_first_order_rho_direct_test.hpp:27-58 and sum_up_direct_test.hpp:
60-124 restated as struct packing here."""
import struct
import types

import numpy as np
import pytest

from fast_rpa_eval_amd import dumps, synth


def _problems(T=2):
    prob = synth.make_rho1_problem(10, 11.0, 40, (20, 120), (10, 90), config_k=4)
    # both dumps of one run describe the same integration grid (init_dfpt_device_data takes one
    # n_full_points for rho1 and the Hartree sum)
    fp = synth.make_far_problem(10, 11.0, 0, 6, config_k=4, xyz=prob.atom_xyz, n_tile=T,
                                outer_radius="decreasing", grid=(prob.coord_points, prob.valid_map()))
    synth.add_near_field(fp, 40, "smooth", n_species=2, seed=1)
    return prob, fp


def test_byte_layout_of_a_tiny_rho_dump(tmp_path):
    scal = dict(zip(dumps.RHO_SCALARS, range(1, 8)))
    arrs = {name: np.arange(1, 3 + i, dtype=np.int32) for i, (name, _) in enumerate(dumps.RHO_ARRAYS)}
    p = tmp_path / "r.bin"
    dumps.write(p, dumps.Dump("rho", scal, arrs))
    expect = struct.pack("<7i", *range(1, 8))
    for i in range(len(dumps.RHO_ARRAYS)):
        n = 2 + i
        expect += struct.pack("<Q", n) + struct.pack(f"<{n}i", *range(1, n + 1))
    assert p.read_bytes() == expect


def test_byte_layout_of_a_sumup_dump_header(tmp_path):
    _, fp = _problems()
    d = dumps.from_far_problem(fp)
    p = tmp_path / "s.bin"
    dumps.write(p, d)
    b = p.read_bytes()
    hdr = struct.unpack_from("<14i", b, 0)
    assert hdr == (1, 2, fp.n_full_points, 7, fp.tables["index_cc_dim_0"], fp.n_valid_points, 6, 1, 40,
                   fp.n_centers, fp.n_atoms, fp.n_atoms, 1, 2)
    # first array: n_radial (n_species int32), second: species (n_atoms int32)
    assert struct.unpack_from("<Q", b, 56)[0] == 2
    assert struct.unpack_from("<Q", b, 56 + 8 + 8)[0] == fp.n_atoms
    assert list(struct.unpack_from(f"<{fp.n_atoms}i", b, 56 + 16 + 8)) == list(fp.species)


def test_round_trip_both_kinds(tmp_path):
    prob, fp = _problems()
    for kind, d in (("rho", dumps.from_rho1_problem(prob)), ("sumup", dumps.from_far_problem(fp))):
        p = tmp_path / f"{kind}.bin"
        dumps.write(p, d)
        r = dumps.read(p, kind)
        assert r.scalars == {k: int(v) for k, v in d.scalars.items()}
        for name, _ in dumps._layout(kind)[1]:
            a, b = np.asarray(d.arrays[name]), r.arrays[name]
            assert b.dtype in (np.int32, np.float64) and np.array_equal(a.astype(b.dtype), b), name
    assert dumps.file_names(4) == ("first_order_rho.meta.test.nproc_4.bin", "sumup.test.nproc_4.bin")


def test_malformed_files(tmp_path):
    prob, _ = _problems()
    p = tmp_path / "r.bin"
    dumps.write(p, dumps.from_rho1_problem(prob))
    b = p.read_bytes()
    (tmp_path / "short.bin").write_bytes(b[:-3])
    with pytest.raises(ValueError, match="array i_batch_2_wave_offset"):
        dumps.read(tmp_path / "short.bin", "rho")
    (tmp_path / "long.bin").write_bytes(b + b"\0")
    with pytest.raises(ValueError, match="trailing"):
        dumps.read(tmp_path / "long.bin", "rho")
    zero = b[:28] + struct.pack("<Q", 0) + b[28 + 8 + 4 * prob.n_basis:]
    (tmp_path / "zero.bin").write_bytes(zero)
    with pytest.raises(ValueError, match="basis_atom has size 0"):
        dumps.read(tmp_path / "zero.bin", "rho")
    d = dumps.from_rho1_problem(prob)
    d.arrays["basis_atom"] = np.zeros(0, np.int32)
    with pytest.raises(ValueError, match="non-empty"):
        dumps.write(tmp_path / "x.bin", d)


def rho_view(d, wave):
    """The read rho dump as the oracle's problem interface."""
    nb = d.n_my_batches_work
    return types.SimpleNamespace(
        n_batches=nb, n_full_points=d.n_full_points, n_basis=d.n_basis, n_atoms=d.n_atoms,
        n_max_compute_ham=d.n_max_compute_ham, n_centers_basis_I=d.n_centers_basis_I,
        n_max_batch_size=d.n_max_batch_size, batch_sizes=d.batch_sizes, n_point_batches=d.n_point_batches,
        n_point_batches_prefix_sum=d.n_point_batches_prefix_sum,
        i_valid_point_2_i_full_points_map=d.i_valid_point_2_i_full_points_map,
        n_compute_c_batches=d.n_compute_c_batches,
        dense_i_basis_batches=lambda: np.asfortranarray(d.i_basis_batches.reshape((d.n_centers_basis_I, nb), order="F")),
        atom_valid_n_compute_c_batches=lambda: np.asfortranarray(
            d.atom_valid_n_compute_c_batches.reshape((d.n_atoms + 1, nb), order="F")),
        wave_offset=d.i_batch_2_wave_offset.astype(np.int64), wave_offset_i32=lambda: d.i_batch_2_wave_offset,
        wave=wave)


def sumup_view(d, T, spl):
    """The read sum-up dump as the oracle's far/near problem interface."""
    F = lambda a, *s: np.asfortranarray(a.reshape(s, order="F"))
    lp1 = d.l_pot_max + 1
    dim0, lama = d.index_cc_dim_0, d.l_max_analytic_multipole
    return types.SimpleNamespace(
        n_tile=T, n_full_points=d.n_full_points, n_valid_points=d.n_valid_points, l_pot_max=d.l_pot_max,
        l_max_analytic_multipole=lama, hartree_force_l_add=d.hartree_force_l_add, n_atoms=d.n_atoms,
        n_centers=d.n_centers, species=d.species, l_hartree=d.l_hartree,
        l_hartree_max_far_distance=F(d.l_hartree_max_far_distance, d.n_atoms, T),
        i_valid_point_2_i_full_points_map=d.i_valid_point_2_i_full_points_map,
        coord_points=F(d.coord_points, 3, d.n_full_points), multipole_radius_sq=F(d.multipole_radius_sq, d.n_atoms, T),
        outer_potential_radius=F(d.outer_potential_radius, lp1, d.n_atoms, T),
        multipole_moments=F(d.multipole_moments, lp1 * lp1, d.n_atoms, T),
        centers_hartree_potential=d.centers_hartree_potential, center_to_atom=d.center_to_atom,
        coords_center=F(d.coords_center, 3, d.n_centers),
        tables=dict(index_cc=F(d.index_cc, dim0, 6), index_cc_dim_0=dim0, cc=d.cc, n_cc_lm_ijk=d.n_cc_lm_ijk,
                    index_ijk_max_cc=F(d.index_ijk_max_cc, 3, lama + 1)),
        near=dict(n_hartree_grid=d.n_hartree_grid, species_center=d.species_center, n_grid=d.n_grid,
                  r_grid_min=d.r_grid_min, log_r_grid_inc=d.log_r_grid_inc, spl=spl))


def test_read_dumps_drive_the_oracle_like_the_problems(tmp_path, oracle):
    """Written and read back, the dumps give the oracle bit-identical rho1 and Hartree sums."""
    T = 2
    prob, fp = _problems(T)
    pr, ps = tmp_path / "r.bin", tmp_path / "s.bin"
    dumps.write(pr, dumps.from_rho1_problem(prob))
    dumps.write(ps, dumps.from_far_problem(fp))
    rd, sd = dumps.read(pr, "rho"), dumps.read(ps, "sumup")
    vals = dumps.random_values(rd, sd, T, seed=3)
    wave = np.where(np.isnan(prob.wave), vals["wave"], prob.wave)
    P1 = vals["first_order_density_matrix"]
    a = oracle.rho1(prob, P1, T=T, bias=vals["first_order_rho_bias_part2"])[0]
    b = oracle.rho1(rho_view(rd, wave), P1, T=T, bias=vals["first_order_rho_bias_part2"])[0]
    assert np.array_equal(a, b)
    spl = vals["current_delta_v_hart_part_spl_tile"]
    fp.near["spl"] = spl
    v = sumup_view(sd, T, spl)
    assert np.array_equal(oracle.far(fp), oracle.far(v))
    assert np.array_equal(oracle.near(fp), oracle.near(v))


def test_writer_byte_count_matches_the_reference_dumps(tmp_path):
    """The reference's FHI-aims dumps (Git-LFS pointers: fragment_save_bin/case_18 .. case_32, 8 cases)
    record each file's byte size, and its direct-test log echoes the scalars read from each (fixture
    tests/golden/dump_sizes.json, made by make_dump_sizes.py).  A dump written with those scalars and
    arrays of the lengths they fix (dumps.array_lengths) has exactly the recorded size, for both
    formats and every case: the scalar count and type, every array's element count, element type and
    size_t count prefix agree with the reference's files."""
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dump_sizes.json")))
    assert len(fx["cases"]) == 8
    for case, e in fx["cases"].items():
        for kind in ("rho", "sumup"):
            sc = e[f"{kind}_scalars"]
            _, arrs = dumps._layout(kind)
            lens = dumps.array_lengths(kind, sc)
            arrays = {name: np.zeros(lens[name], dtype=np.int32 if t == "i" else np.float64) for name, t in arrs}
            path = tmp_path / f"{kind}_{case}.bin"
            dumps.write(path, dumps.Dump(kind, dict(sc), arrays))
            size = os.path.getsize(path)
            assert size == e[f"{kind}_bytes"], (case, kind, size, e[f"{kind}_bytes"])
            back = dumps.read(path, kind)
            assert back.scalars == sc
            os.remove(path)
