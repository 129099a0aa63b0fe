"""Seeded synthetic inputs for the rho1 and far-field Hartree paths, in the reference's layouts.

The reference's drivers fill every numeric array with random values on top of index structures
dumped from FHI-aims (fast_rap_src_dfpt_gpu_eval/direct_test/evaluate_first_order_rho_run.cpp:58-65).
Those dumps are not available, so the index structures are generated here too, from a small
physical model:

* atoms uniform in a cube with a minimum separation (rejection sampling);
* 14 basis functions per atom: 3 s shells, 2 p shells, 1 d shell, each exp(-alpha r^2) Y_lm(r^),
  alpha log-uniform in [0.1, 10] bohr^-2, Y_lm in the SHEval convention;
* integration batches around atoms, points ~ centre + N(0, 1 bohr^2), listed in Morton order of
  their centres (neighbouring batches are spatial neighbours, as in a grid partition);
* P1 symmetric, P1_ij ~ N(0, 1) exp(-d_ij / 2 bohr), d_ij the distance of the host atoms.

Layouts follow the reference's Fortran-ABI arguments (evaluate_first_order_rho.h:3-35):
``i_basis_batches`` (n_centers_basis_I, n_batches) 1-based, ``wave_batches_compress`` batch b at
``i_batch_2_wave_offset[b]`` as (nc_pad, np) column-major with nc_pad = ceil(nc/4)*4, the valid ->
full point map 0-based.  Padding rows are filled with NaN so that any read of them shows.

Large configurations keep the basis-index lists compact (CSR: ``ibasis`` + ``ibasis_ptr``) and
64-bit wave offsets, since the dense (n_centers_basis_I, n_batches) table and int32 offsets of the
reference ABI do not scale to them; ``dense_i_basis_batches`` expands on demand.
"""
from dataclasses import dataclass, field

import math

import numpy as np

from .multipole_tables import far_tables, sh_norm, solid_harmonic_terms

ANG = 1.8897261  # bohr per Angstrom

# Basis recipe: (l, count of shells) -> 3 s + 2 p + 1 d = 14 functions per atom.
SHELLS = ((0, 3), (1, 2), (2, 1))


def seeds(k):
    """Seeds of configuration k (1-based): geometry, basis exponents, P1/M, grid, v1."""
    return {"geometry": 1000 + k, "basis": 2000 + k, "p1": 3000 + k, "grid": 4000 + k, "v1": 5000 + k}


def morton_order(xyz, bits=10):
    """Permutation sorting points along a Z-order (Morton) curve of their bounding box."""
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    q = ((xyz - lo) / np.maximum(hi - lo, 1e-300) * (2 ** bits - 1)).astype(np.int64)
    code = np.zeros(len(xyz), dtype=np.int64)
    for b in range(bits):
        for c in range(3):
            code |= ((q[:, c] >> b) & 1) << (3 * b + c)
    return np.argsort(code, kind="stable")


def make_geometry(n_atoms, box_bohr, min_sep_bohr, seed):
    rng = np.random.default_rng(seed)
    out = []
    tries = 0
    while len(out) < n_atoms:
        c = rng.uniform(0.0, box_bohr, size=3)
        tries += 1
        if tries > 2000 * n_atoms:
            raise RuntimeError("cannot place atoms with the requested minimum separation")
        if out and np.min(np.sum((np.asarray(out) - c) ** 2, axis=1)) < min_sep_bohr ** 2:
            continue
        out.append(c)
    return np.asarray(out)


@dataclass
class Basis:
    atom: np.ndarray   # int32 (n_basis,) 0-based host atom
    l: np.ndarray      # int32
    m: np.ndarray      # int32
    alpha: np.ndarray  # float64, bohr^-2

    @property
    def n_basis(self):
        return len(self.atom)


def make_basis(n_atoms, seed):
    rng = np.random.default_rng(seed)
    atom, ls, ms, al = [], [], [], []
    for a in range(n_atoms):
        for l, nsh in SHELLS:
            for _ in range(nsh):
                alpha = float(np.exp(rng.uniform(np.log(0.1), np.log(10.0))))
                for m in range(-l, l + 1):
                    atom.append(a); ls.append(l); ms.append(m); al.append(alpha)
    return Basis(np.asarray(atom, np.int32), np.asarray(ls, np.int32), np.asarray(ms, np.int32),
                 np.asarray(al, np.float64))


def _sh_poly_coeffs(l_max=2):
    """Unit-sphere polynomial of each SHEval Y_lm, l <= l_max: {(l, m): [(i, j, k, coef)]}."""
    out = {}
    for l in range(l_max + 1):
        for m in range(-l, l + 1):
            nrm = sh_norm(l, m)
            out[(l, m)] = [(i, j, k, float(c) * nrm) for (i, j, k, s), c in solid_harmonic_terms(l, m).items()]
    return out


_SH2 = _sh_poly_coeffs(2)


def eval_basis(atom_xyz, basis, idx, pts):
    """phi_{idx[i]}(pts[p]) as a (len(idx), len(pts)) array."""
    d = pts[None, :, :] - atom_xyz[basis.atom[idx]][:, None, :]
    r2 = np.sum(d * d, axis=2)
    r = np.sqrt(r2)
    with np.errstate(invalid="ignore", divide="ignore"):
        u = np.where(r[..., None] > 0, d / r[..., None], 0.0)
    x, y, z = u[..., 0], u[..., 1], u[..., 2]
    ang = np.zeros_like(r2)
    ls, ms = basis.l[idx], basis.m[idx]
    for (l, m), poly in _SH2.items():
        sel = (ls == l) & (ms == m)
        if not np.any(sel):
            continue
        v = 0.0
        for i, j, k, c in poly:
            v = v + c * x[sel] ** i * y[sel] ** j * z[sel] ** k
        ang[sel] = v
    return np.exp(-basis.alpha[idx][:, None] * r2) * ang


def make_p1(atom_xyz, basis, seed, n_tile=1, cut=0.0):
    """Symmetric P1 (n_basis, n_basis, n_tile), Fortran order."""
    rng = np.random.default_rng(seed)
    nb = basis.n_basis
    xyz = atom_xyz[basis.atom]
    out = np.empty((nb, nb, n_tile), order="F")
    for t in range(n_tile):
        a = rng.standard_normal((nb, nb))
        a = np.triu(a) + np.triu(a, 1).T
        d = np.sqrt(np.sum((xyz[:, None, :] - xyz[None, :, :]) ** 2, axis=2))
        p = a * np.exp(-d / 2.0)
        if cut > 0:
            p[np.abs(p) < cut] = 0.0
        out[:, :, t] = p
    return out


@dataclass
class Rho1Problem:
    """Inputs of evaluate_first_order_rho in the reference's argument layout."""
    n_atoms: int
    n_basis: int
    n_full_points: int
    n_max_batch_size: int
    n_max_compute_ham: int
    atom_xyz: np.ndarray
    basis: Basis
    batch_sizes: np.ndarray                 # int32 (nb,)
    n_point_batches: np.ndarray             # int32 (nb,)
    n_point_batches_prefix_sum: np.ndarray  # int32 (nb+1,)
    i_valid_point_2_i_full_points_map: np.ndarray  # int32 (n_full_points,), first sum(np) used
    n_compute_c_batches: np.ndarray         # int32 (nb,)
    ibasis_ptr: np.ndarray                  # int64 (nb+1,) CSR of 1-based basis indices
    ibasis: np.ndarray                      # int32 (sum nc,)
    wave_offset: np.ndarray                 # int64 (nb+1,)
    coord_points: np.ndarray                # float64 (3, n_full_points) Fortran order
    wave: np.ndarray = None                 # float64 (wave_offset[-1],) or None (generated on GPU)
    extra: dict = field(default_factory=dict)

    @property
    def n_batches(self):
        return len(self.n_point_batches)

    @property
    def n_centers_basis_I(self):
        return self.n_basis

    def dense_i_basis_batches(self):
        out = np.zeros((self.n_centers_basis_I, self.n_batches), dtype=np.int32, order="F")
        for b in range(self.n_batches):
            s, e = self.ibasis_ptr[b], self.ibasis_ptr[b + 1]
            out[: e - s, b] = self.ibasis[s:e]
        return out

    def atom_valid_n_compute_c_batches(self):
        """(n_atoms + 1, n_batches): per-atom prefix over a batch's compute rows (basis sorted by atom)."""
        out = np.zeros((self.n_atoms + 1, self.n_batches), dtype=np.int32, order="F")
        for b in range(self.n_batches):
            ib = self.ibasis[self.ibasis_ptr[b]:self.ibasis_ptr[b + 1]] - 1
            cnt = np.bincount(self.basis.atom[ib], minlength=self.n_atoms)
            out[1:, b] = np.cumsum(cnt)
        return out

    def valid_map(self):
        """i_valid_point_2_i_full_points_map restricted to the valid points (its first sum(np)
        entries; the rho1 path reads it only through n_point_batches_prefix_sum)."""
        return self.i_valid_point_2_i_full_points_map[:int(self.n_point_batches_prefix_sum[-1])]

    def wave_offset_i32(self):
        if self.wave_offset[-1] >= 2 ** 31:
            raise OverflowError("wave_batches_compress exceeds the int32 offsets of the reference ABI")
        return self.wave_offset.astype(np.int32)


def _batch_layout(nps, ncs, extra_rng):
    """Full/valid point bookkeeping: batch b owns full points [fs, fs + batch_size)."""
    nb = len(nps)
    batch_sizes = nps + extra_rng.integers(0, np.maximum(nps // 8, 1) + 1)
    full_start = np.concatenate([[0], np.cumsum(batch_sizes)])
    prefix = np.concatenate([[0], np.cumsum(nps)]).astype(np.int32)
    n_full = int(full_start[-1])
    mp = np.zeros(n_full, dtype=np.int32)
    for b in range(nb):
        sel = np.sort(extra_rng.choice(batch_sizes[b], size=nps[b], replace=False))
        mp[prefix[b]:prefix[b + 1]] = full_start[b] + sel
    ncp = (ncs + 3) // 4 * 4
    wave_offset = np.concatenate([[0], np.cumsum(ncp.astype(np.int64) * nps)]).astype(np.int64)
    return batch_sizes.astype(np.int32), prefix, mp, n_full, wave_offset


def make_rho1_problem(n_atoms, box_ang, n_batches, np_range, nc_range=None, config_k=1, fill_wave=True,
                      batch_spread_bohr=1.5, zero_nc=(), zero_np=(), host_atoms=None):
    """Batched rho1 inputs.

    nc_range None: natural compute sets (basis functions with max_batch |phi| > 1e-10, the rule of
    the small CPU-runnable configuration); otherwise nc ~ U{nc_range} nearest basis functions.
    np_range: int (fixed) or (lo, hi) inclusive.
    zero_nc / zero_np: batch indices forced to n_compute_c = 0 / n_points = 0 (edge cases).
    host_atoms: atoms the batches are centred on, in turn (default: all atoms); a few atoms give a
    few hundred batches the density of a full grid (tens of batches per atom).
    """
    sd = seeds(config_k)
    xyz = make_geometry(n_atoms, box_ang * ANG, 1.5 * ANG, sd["geometry"])
    basis = make_basis(n_atoms, sd["basis"])
    rng = np.random.default_rng(sd["grid"])
    if isinstance(np_range, int):
        nps = np.full(n_batches, np_range, dtype=np.int64)
    else:
        nps = rng.integers(np_range[0], np_range[1] + 1, size=n_batches)
    nps[list(zero_np)] = 0
    hosts = np.arange(n_atoms) if host_atoms is None else np.asarray(host_atoms)
    host = hosts[np.arange(n_batches) % len(hosts)]
    if nc_range is None:
        centres = xyz[host]
    else:
        centres = xyz[host] + rng.normal(0.0, batch_spread_bohr, size=(n_batches, 3))
        # integration batches come from a spatial partition of the grid: list them along a
        # space-filling curve, as neighbouring batches are in a real grid
        centres = centres[morton_order(centres)]
    pts_per_batch = [centres[b] + rng.normal(0.0, 1.0, size=(nps[b], 3)) for b in range(n_batches)]

    nb_basis = basis.n_basis
    basis_xyz = xyz[basis.atom]
    ibasis_list = []
    if nc_range is None:
        for b in range(n_batches):
            phi = eval_basis(xyz, basis, np.arange(nb_basis), pts_per_batch[b])
            keep = np.nonzero(np.max(np.abs(phi), axis=1) > 1e-10)[0]
            ibasis_list.append(keep.astype(np.int32) + 1)
    else:
        ncs = rng.integers(nc_range[0], nc_range[1] + 1, size=n_batches)
        ncs = np.minimum(ncs, nb_basis)
        ncs[list(zero_nc)] = 0
        # nearest basis functions to the batch centre: all functions of an atom share its distance,
        # so whole atoms are taken in (distance, atom index) order and the last one is truncated
        # (the basis of atom a is the contiguous range [start[a], start[a + 1])).
        start = np.searchsorted(basis.atom, np.arange(n_atoms + 1))
        for c0 in range(0, n_batches, 4096):
            c1 = min(c0 + 4096, n_batches)
            d2 = np.sum((xyz[None, :, :] - centres[c0:c1, None, :]) ** 2, axis=2)
            order = np.argsort(d2, axis=1, kind="stable")
            for b in range(c0, c1):
                need, got, parts = int(ncs[b]), 0, []
                for a in order[b - c0]:
                    if got >= need:
                        break
                    parts.append(np.arange(start[a], start[a + 1]))
                    got += start[a + 1] - start[a]
                idx = np.concatenate(parts)[:need] if parts else np.zeros(0, np.int64)
                ibasis_list.append(np.sort(idx).astype(np.int32) + 1)
    ncs = np.asarray([len(v) for v in ibasis_list], dtype=np.int64)
    batch_sizes, prefix, mp, n_full, wave_offset = _batch_layout(nps, ncs, rng)

    coord_points = np.zeros((3, n_full), order="F")
    for b in range(n_batches):
        coord_points[:, mp[prefix[b]:prefix[b + 1]]] = pts_per_batch[b].T

    ibasis_ptr = np.concatenate([[0], np.cumsum(ncs)]).astype(np.int64)
    ibasis = np.concatenate(ibasis_list).astype(np.int32) if ibasis_list else np.zeros(0, np.int32)
    prob = Rho1Problem(
        n_atoms=n_atoms, n_basis=nb_basis, n_full_points=n_full, n_max_batch_size=int(batch_sizes.max()),
        n_max_compute_ham=int(ncs.max()), atom_xyz=xyz, basis=basis, batch_sizes=batch_sizes,
        n_point_batches=nps.astype(np.int32), n_point_batches_prefix_sum=prefix,
        i_valid_point_2_i_full_points_map=mp, n_compute_c_batches=ncs.astype(np.int32),
        ibasis_ptr=ibasis_ptr, ibasis=ibasis, wave_offset=wave_offset, coord_points=coord_points)
    if fill_wave:
        prob.wave = build_wave(prob)
    return prob


def build_wave(prob):
    """wave_batches_compress: batch b's (nc_pad, np) block of phi, NaN padding rows."""
    w = np.full(int(prob.wave_offset[-1]), np.nan)
    for b in range(prob.n_batches):
        nc, npt = int(prob.n_compute_c_batches[b]), int(prob.n_point_batches[b])
        if nc == 0 or npt == 0:
            continue
        ncp = (nc + 3) // 4 * 4
        idx = prob.ibasis[prob.ibasis_ptr[b]:prob.ibasis_ptr[b + 1]] - 1
        pts = prob.coord_points[:, prob.i_valid_point_2_i_full_points_map[
            prob.n_point_batches_prefix_sum[b]:prob.n_point_batches_prefix_sum[b + 1]]].T
        blk = np.full((ncp, npt), np.nan)
        blk[:nc] = eval_basis(prob.atom_xyz, prob.basis, idx, pts)
        w[prob.wave_offset[b]:prob.wave_offset[b + 1]] = blk.ravel(order="F")
    return w


def random_like_wave(prob, k, seed):
    """(k * wave_offset[-1],) uniform [-1, 1] data with NaN padding rows, for gradient/hessian waves."""
    rng = np.random.default_rng(seed)
    out = rng.uniform(-1.0, 1.0, size=k * int(prob.wave_offset[-1]))
    for b in range(prob.n_batches):
        nc, npt = int(prob.n_compute_c_batches[b]), int(prob.n_point_batches[b])
        ncp = (nc + 3) // 4 * 4
        if ncp == nc or npt == 0:
            continue
        blk = out[k * prob.wave_offset[b]:k * prob.wave_offset[b + 1]].reshape((ncp, k, npt), order="F")
        blk[nc:] = np.nan
    return out


# ----------------------------------------------------------------------------------------------
# Far field
# ----------------------------------------------------------------------------------------------

@dataclass
class FarProblem:
    """Inputs of the far-field branch of sum_up_whole_potential (sum_up_whole_potential.h:3-48)."""
    n_atoms: int
    n_full_points: int
    n_valid_points: int
    l_pot_max: int
    l_max_analytic_multipole: int
    hartree_force_l_add: int
    n_tile: int
    coords_center: np.ndarray                 # (3, n_centers) F
    coord_points: np.ndarray                  # (3, n_full_points) F
    i_valid_point_2_i_full_points_map: np.ndarray  # int32 (n_valid_points,)
    species: np.ndarray                       # int32 (n_atoms,) 1-based
    l_hartree: np.ndarray                     # int32 (n_species,)
    centers_hartree_potential: np.ndarray     # int32 1-based
    center_to_atom: np.ndarray                # int32 1-based
    l_hartree_max_far_distance: np.ndarray    # int32 (n_atoms, T) F
    multipole_radius_sq: np.ndarray           # (n_atoms, T) F
    outer_potential_radius: np.ndarray        # (l_pot_max+1, n_atoms, T) F
    multipole_moments: np.ndarray             # ((l_pot_max+1)^2, n_atoms, T) F
    tables: dict = field(default_factory=dict)
    near: dict = None                         # near-field inputs (add_near_field)

    @property
    def n_centers(self):
        return self.coords_center.shape[1]


def moments(l_max, l_pot_max, n_atoms, n_tile, seed):
    rng = np.random.default_rng(seed)
    M = np.zeros(((l_pot_max + 1) ** 2, n_atoms, n_tile), order="F")
    fact = np.array([float(np.prod(np.arange(1, l + 2))) for l in range(l_pot_max + 1)])
    for l in range(l_max + 1):
        for m in range(-l, l + 1):
            M[l * l + l + m] = rng.standard_normal((n_atoms, n_tile)) / fact[l]
    return M


def make_far_problem(n_atoms, box_ang, n_points, l_max, config_k=2, min_dist_bohr=8.0, all_far=True,
                     n_tile=1, outer_radius=None, xyz=None, point_margin_bohr=8.0, grid=None):
    """Far-field inputs.

    all_far: points rejected unless >= min_dist_bohr from every atom (the far-field-only case);
    otherwise uniform in the padded box, pairs closer than min_dist_bohr are near-field and skipped.
    outer_radius: None -> 1e30 (no l reduction); 'decreasing' -> a per-atom table that lowers l
    with distance, exercising the reference's l-reduction loop.
    grid: (coord_points (3, n_full) F, i_valid_point_2_i_full_points_map) of an existing integration
    grid (e.g. a Rho1Problem's batches) used instead of generated points; n_points is then ignored.
    """
    sd = seeds(config_k)
    if xyz is None:
        xyz = make_geometry(n_atoms, box_ang * ANG, 1.5 * ANG, sd["geometry"])
    rng = np.random.default_rng(sd["grid"])
    if grid is not None:
        return _far_problem(xyz, np.asfortranarray(grid[0]), np.ascontiguousarray(grid[1], np.int32), l_max,
                            sd, rng, min_dist_bohr, n_tile, outer_radius)
    lo = xyz.min(axis=0) - point_margin_bohr
    hi = xyz.max(axis=0) + point_margin_bohr
    pts = []
    need = n_points
    while need > 0:
        cand = rng.uniform(lo, hi, size=(max(2 * need, 1024), 3))
        if all_far:
            ok = np.ones(len(cand), bool)
            for s in range(0, n_atoms, 64):
                d2 = np.sum((cand[:, None, :] - xyz[None, s:s + 64, :]) ** 2, axis=2)
                ok &= np.all(d2 >= min_dist_bohr ** 2, axis=1)
            cand = cand[ok]
        pts.append(cand[:need])
        need -= len(pts[-1])
    pts = np.concatenate(pts)
    # a few invalid full points between the valid ones: the map is a strided subset
    n_full = n_points + n_points // 16
    full_idx = np.sort(rng.choice(n_full, size=n_points, replace=False)).astype(np.int32)
    coord_points = np.zeros((3, n_full), order="F")
    coord_points[:, full_idx] = pts.T
    return _far_problem(xyz, coord_points, full_idx, l_max, sd, rng, min_dist_bohr, n_tile, outer_radius)


def _far_problem(xyz, coord_points, full_idx, l_max, sd, rng, min_dist_bohr, n_tile, outer_radius):
    n_atoms = len(xyz)
    n_full = coord_points.shape[1]
    n_points = len(full_idx)
    l_pot_max = l_max
    lama = l_pot_max + 1
    tabs = far_tables(lama)
    M = moments(l_max, l_pot_max, n_atoms, n_tile, sd["p1"])
    opr = np.full((l_pot_max + 1, n_atoms, n_tile), 1e30, order="F")
    if outer_radius == "decreasing":
        base = rng.uniform(15.0, 40.0, size=(n_atoms, n_tile))
        for l in range(l_pot_max + 1):
            opr[l] = (base * (1.0 + 0.35 * (l_pot_max - l))) ** 2
    return FarProblem(
        n_atoms=n_atoms, n_full_points=n_full, n_valid_points=n_points, l_pot_max=l_pot_max,
        l_max_analytic_multipole=lama, hartree_force_l_add=1, n_tile=n_tile,
        coords_center=np.asfortranarray(xyz.T), coord_points=coord_points,
        i_valid_point_2_i_full_points_map=full_idx, species=np.ones(n_atoms, np.int32),
        l_hartree=np.array([l_max], np.int32),
        centers_hartree_potential=np.arange(1, n_atoms + 1, dtype=np.int32),
        center_to_atom=np.arange(1, n_atoms + 1, dtype=np.int32),
        l_hartree_max_far_distance=np.full((n_atoms, n_tile), l_max, np.int32, order="F"),
        multipole_radius_sq=np.full((n_atoms, n_tile), min_dist_bohr ** 2, order="F"),
        outer_potential_radius=opr, multipole_moments=M, tables=tabs)


def add_near_field(fp, n_hartree_grid=80, kind="smooth", n_species=2, r_max_bohr=40.0, seed=0):
    """Near-field inputs of sum_up_whole_potential on a FarProblem (in place; returns fp).

    Atoms alternate over n_species species (1 or 2); species 2 has l_hartree lowered by 2, a larger
    r_grid_min and n_grid = n_hartree_grid - 7, so species indexing and the per-species clamp of the
    spline interval are exercised.  Log grid of species s: r_i = r_grid_min(s) exp((i-1) log_r_grid_inc(s)),
    i = 1..n_grid(s), r_{n_grid} = r_max_bohr.  current_delta_v_hart_part_spl_tile (lm, 2, grid, atom, t):
      kind 'smooth': f_lm(r) = A_lm r^l / (1 + r)^(2l+1) with the derivative d f / d i = f'(r) r log_inc;
      kind 'cubic' : f_lm(i) = a + b i + c i^2 + d i^3 in the grid index (the cubic Hermite spline is then
                     exact, also outside the grid, which gives the closed form of the near sum).
    Grid entries beyond a species' n_grid are NaN (the spline must never read them).  fp.near['coef']
    keeps the cubic coefficients (lm, 4, atom, t) for the closed form.
    """
    assert n_species in (1, 2) and n_hartree_grid >= 9
    rng = np.random.default_rng(seed)
    n_atoms, T, L = fp.n_atoms, fp.n_tile, fp.l_pot_max
    L2 = (L + 1) ** 2
    if n_species == 2:
        fp.species = (1 + np.arange(n_atoms) % 2).astype(np.int32)
        fp.l_hartree = np.array([L, max(L - 2, 0)], np.int32)
    species_center = fp.species[fp.center_to_atom - 1].astype(np.int32)
    n_grid = np.array([n_hartree_grid, n_hartree_grid - 7][:n_species], np.int32)
    r_grid_min = np.array([1e-4, 3e-4][:n_species])
    log_inc = np.log(r_max_bohr / r_grid_min) / (n_grid - 1)
    spl = np.full((L2, 2, n_hartree_grid, n_atoms, T), np.nan, order="F")
    coef = np.zeros((L2, 4, n_atoms, T), order="F")
    lvec = np.array([l for l in range(L + 1) for _ in range(2 * l + 1)])
    for a in range(n_atoms):
        s = fp.species[a] - 1
        ng = n_grid[s]
        i = np.arange(1, ng + 1, dtype=np.float64)
        for t in range(T):
            if kind == "smooth":
                A = rng.standard_normal(L2) / np.array([math.factorial(l + 1) for l in lvec])
                r = r_grid_min[s] * np.exp((i - 1) * log_inc[s])
                lr = lvec[:, None]
                f = r[None, :] ** lr / (1.0 + r[None, :]) ** (2 * lr + 1)
                df = f * (lr / r[None, :] - (2 * lr + 1) / (1.0 + r[None, :]))
                spl[:, 0, :ng, a, t] = A[:, None] * f
                spl[:, 1, :ng, a, t] = A[:, None] * df * r[None, :] * log_inc[s]
            elif kind == "cubic":
                c = rng.standard_normal((L2, 4)) / np.array([1.0, ng, ng ** 2, ng ** 3])
                coef[:, :, a, t] = c
                spl[:, 0, :ng, a, t] = c[:, 0:1] + c[:, 1:2] * i + c[:, 2:3] * i ** 2 + c[:, 3:4] * i ** 3
                spl[:, 1, :ng, a, t] = c[:, 1:2] + 2 * c[:, 2:3] * i + 3 * c[:, 3:4] * i ** 2
            else:
                raise ValueError(kind)
    fp.near = dict(n_hartree_grid=int(n_hartree_grid), species_center=species_center, n_grid=n_grid,
                   r_grid_min=r_grid_min, log_r_grid_inc=log_inc, spl=spl, kind=kind, coef=coef)
    return fp
