"""Reader and writer of the reference's binary test dumps, and their replay through the drop-in.

Formats (fast_rap_src_dfpt_gpu_eval/evaluate_first_order_rho_direct_test.hpp:7-128 and
sum_up_direct_test.hpp:7-219, written by the host program when built with SAVE_DFPT_DATA_TEST):
native-endian int32 scalars in a fixed order, then each array as a size_t element count (> 0)
followed by its raw elements (int32 or float64), in a fixed order.  The files are named
first_order_rho.meta.test.nproc_<n>.bin and sumup.test.nproc_<n>.bin
(direct_test/evaluate_first_order_rho_run.cpp:324-325).

The dumps hold the batch and grid metadata only; the value arrays (wave, density matrices, bias,
spline tile, delta_v) are filled with random numbers by the reference's runner
(evaluate_first_order_rho_run.cpp:13-60), and `replay` does the same with a seeded generator.
"""
import struct
from dataclasses import dataclass, field

import numpy as np

RHO_SCALARS = ("n_my_batches_work", "n_full_points", "n_basis", "n_atoms", "n_max_compute_ham",
               "n_centers_basis_I", "n_max_batch_size")
RHO_ARRAYS = (("basis_atom", "i"), ("batch_sizes", "i"), ("n_point_batches", "i"),
              ("n_point_batches_prefix_sum", "i"), ("i_valid_point_2_i_full_points_map", "i"),
              ("n_compute_c_batches", "i"), ("i_basis_batches", "i"), ("atom_valid_n_compute_c_batches", "i"),
              ("i_batch_2_wave_offset", "i"))

SUMUP_SCALARS = ("j_atom_begin", "j_atom_end", "n_full_points", "l_max_analytic_multipole", "index_cc_dim_0",
                 "n_valid_points", "l_pot_max", "n_max_radial", "n_hartree_grid", "n_centers",
                 "n_centers_hartree_potential", "n_atoms", "hartree_force_l_add", "n_species")
SUMUP_ARRAYS = (("n_radial", "i"), ("species", "i"), ("species_center", "i"), ("l_hartree_max_far_distance", "i"),
                ("i_valid_point_2_i_full_points_map", "i"), ("coord_points", "d"), ("multipole_radius_sq", "d"),
                ("outer_potential_radius", "d"), ("multipole_moments", "d"), ("l_hartree", "i"), ("n_grid", "i"),
                ("n_cc_lm_ijk", "i"), ("centers_hartree_potential", "i"), ("center_to_atom", "i"),
                ("index_cc", "i"), ("index_ijk_max_cc", "i"), ("cc", "d"), ("coords_center", "d"),
                ("r_grid_min", "d"), ("log_r_grid_inc", "d"), ("scale_radial", "d"), ("partition_tab", "d"))

_DT = {"i": np.dtype("=i4"), "d": np.dtype("=f8")}


@dataclass
class Dump:
    """One dump: `scalars` (name -> int) and `arrays` (name -> flat numpy array, column-major data)."""
    kind: str
    scalars: dict = field(default_factory=dict)
    arrays: dict = field(default_factory=dict)

    def __getattr__(self, name):
        d = self.__dict__
        if name in d.get("scalars", {}):
            return d["scalars"][name]
        if name in d.get("arrays", {}):
            return d["arrays"][name]
        raise AttributeError(name)


def _layout(kind):
    if kind == "rho":
        return RHO_SCALARS, RHO_ARRAYS
    if kind == "sumup":
        return SUMUP_SCALARS, SUMUP_ARRAYS
    raise ValueError(kind)


def write(path, dump):
    """Write a dump in the reference's format (every array must be non-empty, as its writer checks)."""
    scal, arrs = _layout(dump.kind)
    with open(path, "wb") as f:
        f.write(struct.pack("=" + "i" * len(scal), *(int(dump.scalars[k]) for k in scal)))
        for name, t in arrs:
            a = np.ascontiguousarray(np.asarray(dump.arrays[name]).ravel(order="F"), dtype=_DT[t])
            if a.size == 0:
                raise ValueError(f"{name}: the format requires a non-empty array")
            f.write(struct.pack("=Q", a.size))
            f.write(a.tobytes())


def read(path, kind):
    """Read a dump written by the reference (or by `write`); kind 'rho' or 'sumup'."""
    scal, arrs = _layout(kind)
    with open(path, "rb") as f:
        buf = f.read()
    off = 4 * len(scal)
    if len(buf) < off:
        raise ValueError(f"{path}: truncated scalar header")
    d = Dump(kind, dict(zip(scal, struct.unpack_from("=" + "i" * len(scal), buf, 0))), {})
    for name, t in arrs:
        if off + 8 > len(buf):
            raise ValueError(f"{path}: truncated before array {name}")
        (n,) = struct.unpack_from("=Q", buf, off)
        off += 8
        nbytes = n * _DT[t].itemsize
        if n == 0 or off + nbytes > len(buf):
            raise ValueError(f"{path}: array {name} has size {n} (file has {len(buf) - off} bytes left)")
        d.arrays[name] = np.frombuffer(buf, dtype=_DT[t], count=n, offset=off).copy()
        off += nbytes
    if off != len(buf):
        raise ValueError(f"{path}: {len(buf) - off} trailing bytes")
    return d


def array_lengths(kind, scalars):
    """Element count of every array of a dump, as the scalars fix it: the reference's host program
    writes the arrays with these shapes (the atom tile of the sum-up arrays is j_atom_end -
    j_atom_begin + 1; i_valid_point_2_i_full_points_map of the rho dump has one entry per full grid
    point, of the sum-up dump one per valid point)."""
    s = scalars
    if kind == "rho":
        nb, nbI = s["n_my_batches_work"], s["n_centers_basis_I"]
        return dict(basis_atom=s["n_basis"], batch_sizes=nb, n_point_batches=nb, n_point_batches_prefix_sum=nb + 1,
                    i_valid_point_2_i_full_points_map=s["n_full_points"], n_compute_c_batches=nb,
                    i_basis_batches=nbI * nb, atom_valid_n_compute_c_batches=(s["n_atoms"] + 1) * nb,
                    i_batch_2_wave_offset=nb + 1)
    if kind == "sumup":
        T = s["j_atom_end"] - s["j_atom_begin"] + 1
        na, ns, nc = s["n_atoms"], s["n_species"], s["n_centers"]
        lp1, lama1, d0 = s["l_pot_max"] + 1, s["l_max_analytic_multipole"] + 1, s["index_cc_dim_0"]
        return dict(n_radial=ns, species=na, species_center=nc, l_hartree_max_far_distance=na * T,
                    i_valid_point_2_i_full_points_map=s["n_valid_points"], coord_points=3 * s["n_full_points"],
                    multipole_radius_sq=na * T, outer_potential_radius=lp1 * na * T,
                    multipole_moments=lp1 * lp1 * na * T, l_hartree=ns, n_grid=ns, n_cc_lm_ijk=lama1,
                    centers_hartree_potential=s["n_centers_hartree_potential"], center_to_atom=nc,
                    index_cc=d0 * 6, index_ijk_max_cc=3 * lama1, cc=d0, coords_center=3 * nc, r_grid_min=ns,
                    log_r_grid_inc=ns, scale_radial=ns, partition_tab=s["n_full_points"])
    raise ValueError(kind)


def file_names(nproc):
    """(rho meta dump, sum-up dump) names for a run on nproc MPI ranks (rank 0 writes them)."""
    return f"first_order_rho.meta.test.nproc_{nproc}.bin", f"sumup.test.nproc_{nproc}.bin"


# ------------------------------------------------------------------------------------------------
# conversion from the synthetic problems (the writer's input in tests and examples)
# ------------------------------------------------------------------------------------------------

def from_rho1_problem(prob):
    """Rho1Problem -> 'rho' dump (dense 1-based i_basis_batches, int32 wave offsets)."""
    ib = prob.dense_i_basis_batches()
    off = np.asarray(prob.wave_offset, dtype=np.int64)
    if off[-1] > np.iinfo(np.int32).max:
        raise ValueError("wave offsets exceed int32: the dump format stores i_batch_2_wave_offset as int")
    scal = dict(n_my_batches_work=prob.n_batches, n_full_points=prob.n_full_points, n_basis=prob.n_basis,
                n_atoms=prob.n_atoms, n_max_compute_ham=prob.n_max_compute_ham,
                n_centers_basis_I=prob.n_centers_basis_I, n_max_batch_size=prob.n_max_batch_size)
    arrs = dict(basis_atom=prob.basis.atom + 1, batch_sizes=prob.batch_sizes, n_point_batches=prob.n_point_batches,
                n_point_batches_prefix_sum=prob.n_point_batches_prefix_sum,
                i_valid_point_2_i_full_points_map=prob.i_valid_point_2_i_full_points_map,
                n_compute_c_batches=prob.n_compute_c_batches, i_basis_batches=ib,
                atom_valid_n_compute_c_batches=prob.atom_valid_n_compute_c_batches(),
                i_batch_2_wave_offset=off.astype(np.int32))
    return Dump("rho", scal, {k: np.asarray(v).ravel(order="F") for k, v in arrs.items()})


def from_far_problem(fp, j_atom_begin=1):
    """FarProblem (with near-field inputs from synth.add_near_field) -> 'sumup' dump."""
    if fp.near is None:
        raise ValueError("the sum-up dump carries the near-field grids: call synth.add_near_field first")
    nf, tb = fp.near, fp.tables
    n_species = len(fp.l_hartree)
    scal = dict(j_atom_begin=j_atom_begin, j_atom_end=j_atom_begin + fp.n_tile - 1, n_full_points=fp.n_full_points,
                l_max_analytic_multipole=fp.l_max_analytic_multipole, index_cc_dim_0=tb["index_cc_dim_0"],
                n_valid_points=fp.n_valid_points, l_pot_max=fp.l_pot_max, n_max_radial=1,
                n_hartree_grid=nf["n_hartree_grid"], n_centers=fp.n_centers,
                n_centers_hartree_potential=len(fp.centers_hartree_potential), n_atoms=fp.n_atoms,
                hartree_force_l_add=fp.hartree_force_l_add, n_species=n_species)
    arrs = dict(n_radial=np.ones(n_species, np.int32), species=fp.species, species_center=nf["species_center"],
                l_hartree_max_far_distance=fp.l_hartree_max_far_distance,
                i_valid_point_2_i_full_points_map=fp.i_valid_point_2_i_full_points_map,
                coord_points=fp.coord_points, multipole_radius_sq=fp.multipole_radius_sq,
                outer_potential_radius=fp.outer_potential_radius, multipole_moments=fp.multipole_moments,
                l_hartree=fp.l_hartree, n_grid=nf["n_grid"], n_cc_lm_ijk=tb["n_cc_lm_ijk"],
                centers_hartree_potential=fp.centers_hartree_potential, center_to_atom=fp.center_to_atom,
                index_cc=tb["index_cc"], index_ijk_max_cc=tb["index_ijk_max_cc"], cc=tb["cc"],
                coords_center=fp.coords_center, r_grid_min=nf["r_grid_min"], log_r_grid_inc=nf["log_r_grid_inc"],
                scale_radial=np.ones(n_species), partition_tab=np.ones(fp.n_full_points))
    return Dump("sumup", scal, {k: np.asarray(v).ravel(order="F") for k, v in arrs.items()})


# ------------------------------------------------------------------------------------------------
# replay (the reference's direct-test runners)
# ------------------------------------------------------------------------------------------------

def random_values(rho, sumup, T, seed=0):
    """The value arrays the reference's runner fills at random (evaluate_first_order_rho_run.cpp:20-60),
    uniform in [-1, 1), column-major, for an atom tile of T perturbations."""
    rng = np.random.default_rng(seed)
    F = lambda *s: np.asfortranarray(rng.uniform(-1.0, 1.0, size=s))
    wl = int(rho.i_batch_2_wave_offset[rho.n_my_batches_work])
    nb, nmb, nbas = rho.n_my_batches_work, rho.n_max_batch_size, rho.n_basis
    L2 = (sumup.l_pot_max + 1) ** 2
    return dict(
        wave=rng.uniform(-1.0, 1.0, size=wl), gradient_wave=rng.uniform(-1.0, 1.0, size=3 * wl),
        first_order_density_matrix=F(nbas, nbas, T), first_order_rho_bias_part2=F(nmb, nb, T),
        first_order_gradient_rho_bias_batches_atoms=F(3, nmb, nb, T), density_matrix=F(nbas, nbas),
        current_delta_v_hart_part_spl_tile=F(L2, 2, sumup.n_hartree_grid, sumup.n_atoms, T),
        delta_v_hartree=F(sumup.n_full_points, T))


def replay_drop_in(rho, sumup, values, T, j_coord=1):
    """Run the reference-named entry points of libfrpa.so on a pair of dumps and their value arrays,
    as the reference's runner does (evaluate_first_order_rho_run.cpp:100-200): init_dfpt_device_data,
    the rho1 entry point for the atom tile j_atom = 1..T, the Hartree sum over the tile, free.
    The library's atom tile size is taken from FRPA_ATOM_TILE_SIZE (set here to T for the call).
    Both dumps must describe one integration grid (equal n_full_points), as those of one run do; the
    drop-in entry points exit the process on any inconsistency, like the reference's.
    Returns dict(first_order_rho (n_full, T), first_order_gradient_rho (3, nmb, nb, T),
    delta_v_hartree (n_full_sumup, T)), all column-major numpy arrays."""
    import ctypes
    import os

    from ._lib import lib
    if rho.n_full_points != sumup.n_full_points:
        raise ValueError(f"the dumps describe different grids: n_full_points {rho.n_full_points} (rho) vs "
                         f"{sumup.n_full_points} (sum-up)")
    L = lib()
    P = ctypes.c_void_p
    keep = []

    def ip(v):
        c = ctypes.c_int(int(v))
        keep.append(c)
        return ctypes.byref(c)

    def ptr(a, dtype):
        a = np.asfortranarray(np.asarray(a, dtype=dtype))
        keep.append(a)
        return a.ctypes.data_as(P)

    I = lambda a: ptr(a, np.int32)
    D = lambda a: ptr(a, np.float64)
    nb, nmb = rho.n_my_batches_work, rho.n_max_batch_size
    out = dict(first_order_rho=np.zeros((rho.n_full_points, T), order="F"),
               first_order_gradient_rho=np.zeros((3, nmb, nb, T), order="F"),
               delta_v_hartree=np.asfortranarray(values["delta_v_hartree"].copy()))
    rho_p, grad_p, dv_p = D(out["first_order_rho"]), D(out["first_order_gradient_rho"]), D(out["delta_v_hartree"])
    # the entry points write through these pointers: keep the arrays themselves
    out["first_order_rho"], out["first_order_gradient_rho"], out["delta_v_hartree"] = keep[-3], keep[-2], keep[-1]
    v = values
    old = os.environ.get("FRPA_ATOM_TILE_SIZE")
    os.environ["FRPA_ATOM_TILE_SIZE"] = str(T)
    try:
        L.init_dfpt_device_info()
        L.init_dfpt_device_data(
            ip(nb), ip(sumup.n_full_points), ip(sumup.n_valid_points), ip(rho.n_basis), ip(sumup.n_atoms),
            ip(rho.n_max_compute_ham), ip(rho.n_centers_basis_I), ip(nmb), ip(sumup.l_pot_max), ip(sumup.n_max_radial),
            ip(sumup.n_hartree_grid), ip(sumup.n_centers), ip(sumup.n_centers_hartree_potential),
            ip(sumup.hartree_force_l_add), ip(sumup.n_species), ip(sumup.l_max_analytic_multipole),
            ip(sumup.index_cc_dim_0),
            I(rho.basis_atom), I(rho.batch_sizes), I(rho.n_point_batches), I(rho.n_point_batches_prefix_sum),
            I(rho.i_valid_point_2_i_full_points_map), I(rho.n_compute_c_batches), I(rho.i_basis_batches),
            I(rho.atom_valid_n_compute_c_batches), I(rho.i_batch_2_wave_offset), D(v["wave"]), D(v["gradient_wave"]),
            D(v["density_matrix"]),
            I(sumup.n_radial), I(sumup.species), I(sumup.species_center), D(sumup.coord_points), I(sumup.l_hartree),
            I(sumup.n_grid), I(sumup.n_cc_lm_ijk), I(sumup.centers_hartree_potential), I(sumup.center_to_atom),
            I(sumup.index_cc), I(sumup.index_ijk_max_cc), D(sumup.cc), D(sumup.coords_center), D(sumup.r_grid_min),
            D(sumup.log_r_grid_inc), D(sumup.scale_radial), D(sumup.partition_tab),
            D(v["first_order_density_matrix"]), rho_p, D(v["first_order_rho_bias_part2"]), grad_p,
            D(v["first_order_gradient_rho_bias_batches_atoms"]),
            I(sumup.l_hartree_max_far_distance), D(sumup.multipole_radius_sq), D(sumup.outer_potential_radius),
            D(sumup.multipole_moments), D(v["current_delta_v_hart_part_spl_tile"]), dv_p)
        L.evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_(
            ip(nb), ip(1), ip(T), ip(j_coord), ip(rho.n_full_points), ip(rho.n_basis), ip(rho.n_atoms),
            ip(rho.n_max_compute_ham), ip(rho.n_centers_basis_I), ip(nmb), I(rho.batch_sizes), I(rho.n_point_batches),
            I(rho.n_point_batches_prefix_sum), I(rho.i_valid_point_2_i_full_points_map), I(rho.n_compute_c_batches),
            I(rho.i_basis_batches), I(rho.atom_valid_n_compute_c_batches), I(rho.i_batch_2_wave_offset),
            D(v["wave"]), D(v["gradient_wave"]), D(v["first_order_density_matrix"]), rho_p,
            D(v["first_order_rho_bias_part2"]), grad_p, D(v["first_order_gradient_rho_bias_batches_atoms"]),
            I(rho.basis_atom), D(np.ones(rho.n_full_points)), D(v["density_matrix"]))
        f = L.sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_
        f.argtypes = [ctypes.c_int] * 14 + [P] * 24
        f(sumup.j_atom_begin, sumup.j_atom_begin + T - 1, sumup.n_full_points, sumup.l_max_analytic_multipole,
          sumup.index_cc_dim_0, sumup.n_valid_points, sumup.l_pot_max, sumup.n_max_radial, sumup.n_hartree_grid,
          sumup.n_centers, sumup.n_centers_hartree_potential, sumup.n_atoms, sumup.hartree_force_l_add,
          sumup.n_species,
          I(sumup.n_radial), I(sumup.species), I(sumup.species_center), I(sumup.l_hartree_max_far_distance),
          I(sumup.i_valid_point_2_i_full_points_map), D(sumup.coord_points), D(sumup.multipole_radius_sq),
          D(sumup.outer_potential_radius), D(sumup.multipole_moments), D(v["current_delta_v_hart_part_spl_tile"]),
          I(sumup.l_hartree), I(sumup.n_grid), I(sumup.n_cc_lm_ijk), I(sumup.centers_hartree_potential),
          I(sumup.center_to_atom), I(sumup.index_cc), I(sumup.index_ijk_max_cc), D(sumup.cc),
          D(sumup.coords_center), D(sumup.r_grid_min), D(sumup.log_r_grid_inc), D(sumup.scale_radial),
          D(sumup.partition_tab), dv_p)
        L.free_dfpt_device_data()
        L.free_dfpt_device_info()
    finally:
        if old is None:
            os.environ.pop("FRPA_ATOM_TILE_SIZE", None)
        else:
            os.environ["FRPA_ATOM_TILE_SIZE"] = old
    return out
