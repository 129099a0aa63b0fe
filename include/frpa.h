/*
 * fast_rpa_eval_amd: MI355X (gfx950) implementation of the two DFPT grid hot paths of
 * fast_rpa_eval -- the first-order density rho1 (evaluate_first_order_rho) and the Hartree
 * potential sum (sum_up_whole_potential: far-field multipoles and near-field splines) -- plus the
 * adjoints of rho1 and of the far field.
 *
 * Two layers, both C linkage, plain pointers and sizes:
 *
 *  1. Reference entry points (Fortran calling convention: trailing underscore, scalars by pointer,
 *     column-major arrays, the reference's index bases).  Same names and argument lists as the
 *     reference; arrays are HOST pointers.  Errors print and exit(1), as the reference's DEV_CHECK.
 *
 *  2. The frpa_* C-ABI: DEVICE pointers, an explicit HIP stream (hipStream_t passed as void*,
 *     NULL = default stream), int status return (FRPA_OK = 0), frpa_last_error() for the message.
 *     Calls are asynchronous on the stream unless stated.  Plans hold the batch/centre metadata
 *     resident in HBM; a plan is not thread-safe, distinct plans are independent.
 *
 * Array conventions (both layers): Fortran column-major; i_basis lists and centre/atom/species
 * indices 1-based; the valid -> full point map 0-based; wave_batches_compress holds batch b at
 * offset wave_offset[b] as (nc_pad, np) with nc_pad = ceil(nc / 4) * 4 (rows nc..nc_pad-1 are never
 * read); ATOM_TILE_SIZE of the reference is the run-time argument atom_tile_size (T).
 */
#ifndef FRPA_H
#define FRPA_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------ */
/* Status                                                                                     */
/* ------------------------------------------------------------------------------------------ */
#define FRPA_OK 0
#define FRPA_ERR_INVALID 1     /* bad argument or inconsistent sizes */
#define FRPA_ERR_HIP 2         /* HIP runtime error */
#define FRPA_ERR_UNSUPPORTED 3 /* valid input outside what this build implements */
#define FRPA_ERR_STATE 4       /* call order (e.g. reference entry before init_dfpt_device_data) */

/* Message of the last failing call on this thread ("" if none). */
const char *frpa_last_error(void);
/* Library version string. */
const char *frpa_version(void);

/* ------------------------------------------------------------------------------------------ */
/* rho1: first_order_rho(map(p), t) = sum_ij phi_i(p) P1(ib_i, ib_j, t) phi_j(p) + bias          */
/* ------------------------------------------------------------------------------------------ */
typedef struct frpa_rho1_plan frpa_rho1_plan;

/* Precision of the rho1 / adjoint GEMM. */
#define FRPA_PREC_FP64 0   /* FP64 MFMA, FP64 accumulation                                     */
#define FRPA_PREC_OZAKI 1  /* INT8-MFMA split, 5 radix-254 digits per operand, 15 exact INT8     */
                           /* GEMMs, FP64 recombination (max|d rho1|/max|rho1| ~ 1e-12)          */
#define FRPA_PREC_OZAKI4 2 /* INT8-MFMA split with 4 digits, 10 INT8 GEMMs (~ 5e-10)            */
#define FRPA_PREC_FP32 3   /* rho1 forward only, for comparison: FP32 MFMA, FP32 sums per 64-row */
                           /* K step added in FP64 (~1e-7)                                       */
#define FRPA_PREC_BF16X3 4 /* rho1 forward only, for comparison: bf16 hi/lo split, 3 products,   */
                           /* FP32 sums per K step added in FP64 (~5e-6)                         */
#define FRPA_PREC_FP64_SYM 5 /* rho1 forward only, no gradient: FP64 MFMA on the symmetric half */
                             /* of P1 (tile pairs J >= I, the INT8 split's algorithm and launch), */
                             /* wave 16-byte aligned; FRPA_PREC_FP64 multiplies the full P1 as    */
                             /* the reference does                                                */

/*
 * Batch metadata (host arrays, copied into the plan):
 *   n_point_batches[n_batches], n_point_batches_prefix_sum[n_batches + 1],
 *   i_valid_point_2_i_full_points_map[prefix[n_batches]] (0-based full-grid index),
 *   n_compute_c_batches[n_batches],
 *   ibasis_ptr[n_batches + 1] + ibasis[ibasis_ptr[n_batches]]: batch b's 1-based basis indices
 *     (the non-padded prefix of the reference's i_basis_batches(:, b)),
 *   wave_offset[n_batches + 1]: 64-bit offsets of each batch block in wave_batches_compress.
 * Replaces the rho1 part of init_dfpt_device_data (fast_rap_src_dfpt_gpu_eval/device_data.cu:10-361).
 */
int frpa_rho1_plan_create(frpa_rho1_plan **plan, int n_batches, int n_basis, int n_full_points, int n_max_batch_size,
                          const int *n_point_batches, const int *n_point_batches_prefix_sum,
                          const int *i_valid_point_2_i_full_points_map, const int *n_compute_c_batches,
                          const long long *ibasis_ptr, const int *ibasis, const long long *wave_offset);
int frpa_rho1_plan_destroy(frpa_rho1_plan *plan);

/*
 * Forward (fast_rap_src_dfpt_gpu_eval/evaluate_first_order_rho.cu:188-371, same maths as
 * evaluate_first_order_rho.cpp:113-209).  Every precision computes the GGA gradient when
 * gradient_wave is given; the INT8 split then multiplies every tile pair of P1 (twice the products of
 * rho1 alone, which only needs the symmetric half).  Device arrays:
 *   wave                       wave_batches_compress
 *   density_matrix             first_order_density_matrix (n_basis, n_basis, T)
 *   rho_bias (nullable)        first_order_rho_bias_part2 (n_max_batch_size, n_batches, T)
 *   first_order_rho            (n_full_points, T): zeroed, then batch points of t < n_tile written
 *   gradient_wave (nullable)   gradient_wave_batches_compress (3 * wave_offset[n_batches])
 *   grad_bias (nullable)       first_order_gradient_rho_bias_batches_atoms (3, nmb, n_batches, T)
 *   first_order_gradient_rho (nullable, required with gradient_wave) (3, nmb, n_batches, T)
 * n_tile = j_atom_end - j_atom_begin + 1 (<= T).
 */
int frpa_rho1_forward(frpa_rho1_plan *plan, int atom_tile_size, int n_tile, int precision, const double *wave,
                      const double *density_matrix, const double *rho_bias, double *first_order_rho,
                      const double *gradient_wave, const double *grad_bias, double *first_order_gradient_rho,
                      void *stream);

/*
 * Adjoint of the forward with respect to P1 (LDA first-order H integration,
 * fast_rap_src_dfpt_gpu_eval/evaluate_first_order_H.cpp:144-215 with the GGA terms zero):
 *   first_order_H(ib_i, ib_j, t) += sum_{batches, p} g(map(p), t) phi_i(p) phi_j(p)
 * g (n_full_points, T) device; first_order_H (n_basis, n_basis, T) device, ACCUMULATED into.
 * With g = partition_tab * (v_hartree_gradient + v2rho2 * rho1) + H_prefactor1 this is the
 * reference's LDA H1; with g = d L / d rho1 it is d L / d P1.  The INT8-split precisions accumulate in
 * 64-bit fixed point (integer atomics, scale from a bound on |H|) and are bit-reproducible; FP64 uses
 * FP64 atomics (run-to-run differences in the last bits).  Non-finite g or wave leave H undefined.
 * The INT8-split precisions merge consecutive batches with contiguous points into one work item over
 * the union of their compute sets (one atomic per H element per group); env FRPA_ADJOINT_MERGE sets
 * the threshold (|union|^2 <= merge * sum n_compute^2; default 0.8, 0 = every batch alone), read per
 * call; the groups and digits are rebuilt when it changes.
 */
int frpa_rho1_adjoint(frpa_rho1_plan *plan, int atom_tile_size, int precision, const double *wave, const double *g,
                      double *first_order_H, void *stream);

/* Autotuner of the rho1 forward launch: times every launch configuration valid for the plan and
 * precision (batch order: descending work, XCD-aware spatial, spatial with point passes as
 * separate items, or spatial with ranges of row tiles of point passes as items; one workgroup per
 * batch or a
 * persistent grid with per-XCD work counters; points per pass) on the given inputs, `reps` calls
 * each after one warm-up, with HIP events on `stream`; keeps the fastest in the plan for every later
 * frpa_rho1_forward of that precision (the GGA gradient launch included) and writes a description
 * ("order=1 persistent=0 npg=4 ms=..." per candidate, the chosen one first) to `choice`.  The
 * configurations give bit-identical results; first_order_rho is overwritten as by
 * frpa_rho1_forward (no bias).  No counterpart in the reference, whose batch tiling is fixed at
 * compile time (N_BATCHES_TILE, fast_rap_src_dfpt_gpu_eval/evaluate_first_order_rho.cu:188-371). */
int frpa_rho1_autotune(frpa_rho1_plan *plan, int atom_tile_size, int precision, const double *wave,
                       const double *density_matrix, double *first_order_rho, int reps, char *choice,
                       int choice_len, void *stream);

/* Sets the launch configuration of frpa_rho1_forward for one precision explicitly (what the
 * autotuner chooses): order 0 = descending work, 1 = XCD-aware spatial, 2 = (not FP64) the
 * spatial order with every point pass of a batch as its own work item, the passes of a batch
 * adjacent on one XCD so they share its P' gathers in L2; 3 = (not FP64) every range of
 * 64-row tiles (about 16 K steps of work) of a point pass its own item, a batch's items adjacent on
 * one XCD so they also share its phi digits in L2 (per-tile sums go through a device buffer of
 * n_row_tiles x points doubles per batch, 4x with the gradient, times atom_tile_size, and a
 * finishing kernel; environment variable FRPA_TILE_WORK sets the K steps per item); persistent 0 = one
 * workgroup per batch, 1 = a persistent grid sized from the occupancy, k >= 8 = a persistent grid of
 * k workgroups (rounded down to a multiple of 8); points_per_wave = groups of 16 points per wave and
 * pass (FP64: 4, 8, 13; INT8 split S = 5: 4; S = 4: 4, 5; 0 = default), and for the INT8 split 2 =
 * four-wave workgroups over 64-point passes (two per CU, where the 8-wave workgroups over 128-point
 * passes fill a CU's LDS alone), 1 = those workgroups with the phi digits staged by direct
 * global -> LDS loads and the P1 digit gathers issued one K step ahead through a buffer resource
 * (n_basis^2 * 8 < 2^32; larger bases use 2); FP32, BF16x3 and the FP64 symmetric half have that one
 * block shape (1, direct global -> LDS staging, 64-point passes; n_basis^2 * 4 < 2^32, FP64
 * n_basis^2 * 8 < 2^32).  Results do not depend on it.  A new plan starts with FP64 (1, 1, 13) and
 * the INT8 split, FP32, BF16x3 and FP64 symmetric half (3, 1, 1), the autotuner's choices on a
 * 60k-batch workload.  No counterpart in the reference
 * (compile-time tiling). */
int frpa_rho1_set_launch(frpa_rho1_plan *plan, int precision, int order, int persistent, int points_per_wave);

/* One-off conversion of wave_batches_compress into the INT8 digits used by the INT8-split
 * precisions (FRPA_PREC_FP32 / FRPA_PREC_BF16X3: into the FP32 values / bf16 hi and lo planes of
 * those forwards, 4 bytes per element).  The forward and the adjoint each keep digits of their own
 * in the plan, made on their first INT8-split call and reused while the same wave pointer is passed
 * (the FP32 and BF16x3 forwards likewise keep a converted wave each, so alternating them converts
 * nothing and can be captured in a graph): after changing the wave in place (or passing new
 * contents at the same address), call this
 * function, which rebuilds the forward's digits and, once the adjoint has made its own, the
 * adjoint's.  Optional otherwise.  Returns in *bytes (nullable) the device bytes the digits it made
 * hold (the forward's, plus the adjoint's when it rebuilt them). */
int frpa_rho1_prepare_wave(frpa_rho1_plan *plan, int precision, const double *wave, long long *bytes,
                           void *stream);

/* The INT8-split adjoint hands blocks between its waves through LDS slots; every wait for a slot is
 * bounded (default 65536 polls, ~2 ms).  A wait that gives up leaves first_order_H incomplete and
 * is counted on the device: frpa_rho1_adjoint reads the count on every call not under stream
 * capture and returns FRPA_ERR_HIP (first_order_H then not updated by that call);
 * frpa_rho1_adjoint_status reads and clears it (synchronising the stream) and returns FRPA_ERR_HIP
 * if any wait gave up since the last check -- the check for calls replayed from a HIP graph.
 * frpa_rho1_set_adjoint_wait_limit sets the bound in polls (0 = default; small values make the waits
 * give up, for tests of this path).  No counterpart in the reference (its H integration is a CPU
 * function, evaluate_first_order_H.cpp:13-222). */
int frpa_rho1_set_adjoint_wait_limit(frpa_rho1_plan *plan, int polls);
int frpa_rho1_adjoint_status(frpa_rho1_plan *plan, void *stream);

/*
 * Full first-order H (evaluate_first_order_H.cpp:13-222, LDA and GGA terms; a CPU function in the
 * reference): for the perturbations t < n_tile (atom j = j_atom_begin + t, coordinate j_coord),
 *   first_order_H(:, :, t) += sum_batches scatter(C phi^T + phi C^T),
 *   C = (H_prefactor1 + pref3) / 2 phi + sum_c pref2_c dphi_c - [rows of atom j] sum_c pref5_c d2phi_{c, j_coord}
 * with the reference's per-point prefactors (sigma1, pref2, pref3, pref5 of evaluate_first_order_H.cpp:
 * 144-168).  Device arrays in the reference's layouts: atom_valid_n_compute_c_batches (n_atoms + 1,
 * n_batches) int; wave, gradient_wave (3x), hessian_wave (6x: xx, xy, xz, yy, yz, zz) as
 * wave_batches_compress; vsigma_batches and v2rho2_batches (3, nmb) per batch, v2rhosigma_batches and
 * v2sigma2_batches (6, nmb) per batch (component 0 used, as in the reference); gradient_rho (3,
 * n_full_points); H_prefactor1_batches_tile (nmb, n_batches, T); first_order_gradient_rho_batches_tile
 * (3, nmb, n_batches, T); v_hartree_gradient, first_order_rho (n_full_points, T); partition_tab
 * (n_full_points); first_order_H_tile (n_basis, n_basis, T) ACCUMULATED.  FP64 (MFMA).  With every XC
 * derivative zero it equals frpa_rho1_adjoint with g = partition_tab * v_hartree_gradient + H_prefactor1.
 * The plan keeps a device scratch of twice the wave's length (C and phi per batch, in its sorted row
 * order; allocated on the first call) and the relabelled n_basis^2 x n_tile accumulator.
 */
int frpa_rho1_h1(frpa_rho1_plan *plan, int atom_tile_size, int j_atom_begin, int n_tile, int j_coord, int n_atoms,
                 const int *atom_valid_n_compute_c_batches, const double *wave, const double *gradient_wave,
                 const double *hessian_wave, const double *vsigma_batches, const double *gradient_rho,
                 const double *H_prefactor1_batches_tile, const double *first_order_gradient_rho_batches_tile,
                 const double *v_hartree_gradient, const double *first_order_rho, const double *v2rho2_batches,
                 const double *v2rhosigma_batches, const double *v2sigma2_batches, const double *partition_tab,
                 double *first_order_H_tile, void *stream);

/*
 * Bias input of rho1 (generate_first_order_rho_bias_part2.cpp:9-167, a CPU function in the
 * reference): for the atoms j = j_atom_begin .. j_atom_begin + n_tile - 1 (t = j - j_atom_begin) and
 * every batch with n_compute_c > 0,
 *   first_order_rho_bias_part2(p, b, t) = sum_{k in V} phi'(k, p) sum_i (DM(i, k) + DM(k, i)) phi(i, p),
 * V = the batch's compute rows of atom j ([avn(j-1, b), avn(j, b))), phi' = -d phi / d x_{j_coord},
 * DM(i, k) = density_matrix(i_basis(i), i_basis(k)); 0 for empty V; other entries untouched.
 * Device arrays: atom_valid_n_compute_c_batches (n_atoms + 1, n_batches) int, wave and gradient_wave
 * as for frpa_rho1_forward, density_matrix (n_basis, n_basis), first_order_rho_bias_part2
 * (n_max_batch_size, n_batches, atom_tile_size).  n_tile <= min(atom_tile_size, 8), j_coord 1..3.
 * Each call first writes DM(row, c) + DM(c, row) for the tile's row functions into a plan-owned
 * table (allocated on the first call: 256 x n_basis doubles, capped at 1 GiB; environment variable
 * FRPA_BIAS2_TABLE_ROWS sets the row capacity, 0 = none); a tile with more rows than that gathers
 * from density_matrix directly, with the same result bits.
 */
int frpa_rho1_bias_part2(frpa_rho1_plan *plan, int atom_tile_size, int j_atom_begin, int n_tile, int j_coord,
                         int n_atoms, const int *atom_valid_n_compute_c_batches, const double *wave,
                         const double *gradient_wave, const double *density_matrix,
                         double *first_order_rho_bias_part2, void *stream);

/*
 * The inline bias of the v6 variant of the rho1 function
 * (task4_common_calculation_eliminated/version6_rho_with_bias_calculate/evaluate_first_order_rho.cpp:
 * 179-231, its second dgemm DM(V, :) phi and the acc_bias loop), ADDED to first_order_rho_bias:
 *   first_order_rho_bias(p, b, t) += -sum_{i in V} grad_{j_coord} phi(i, p) sum_k DM(i, k) phi(k, p)
 * with V, DM(i, k) and the arrays as for frpa_rho1_bias_part2 (one side of the density matrix where
 * part 2 has both: inline(DM) + inline(DM^T) = part 2 of DM).  v6 adds this term to rho1 on top of
 * its bias input; passing a copy of the bias input through this call and then to frpa_rho1_forward
 * gives v6's rho1.  Entries of batches with n_compute_c = 0, of points beyond a batch's n_points and
 * of perturbations t >= n_tile are unchanged.  Same limits and table as frpa_rho1_bias_part2.
 */
int frpa_rho1_bias_inline(frpa_rho1_plan *plan, int atom_tile_size, int j_atom_begin, int n_tile, int j_coord,
                          int n_atoms, const int *atom_valid_n_compute_c_batches, const double *wave,
                          const double *gradient_wave, const double *density_matrix, double *first_order_rho_bias,
                          void *stream);

/*
 * Bias part 2 of the first-order H (fast_rap_src_dfpt_gpu_eval/evaluate_first_order_H_bias_part2.cpp:
 * 13-218), ADDED to first_order_H_tile for the atoms j = j_atom_begin .. + n_tile - 1 (t = j -
 * j_atom_begin): per batch, with V the compute rows of atom j, w = partition_tab(q),
 *   C(i, p) = w H_times_psi(i, p) + w vrho(p) phi(i, p) + sum_c 2 w vsigma(0, p) gradient_rho(c, q) d_c phi(i, p)
 *   L(i, k) = sum_p C(i, p) (-d_{j_coord} phi(k, p)),  k in V
 *   first_order_H(ib_i, ib_k, t) += L(i, k),  first_order_H(ib_k, ib_i, t) += L(i, k).
 * Device arrays: atom_valid_n_compute_c_batches (n_atoms + 1, n_batches) int; wave and gradient_wave
 * as for frpa_rho1_forward; H_times_psi in the wave's layout; vrho_batches (n_max_batch_size,
 * n_batches); vsigma_batches (3, n_max_batch_size, n_batches) (component 0 is read); gradient_rho
 * (3, n_full_points); partition_tab (n_full_points); first_order_H_tile (n_basis, n_basis,
 * atom_tile_size).  n_tile <= min(atom_tile_size, 8), j_coord 1..3.  FP64; the scatter is FP64
 * atomics.
 */
int frpa_rho1_h1_bias_part2(frpa_rho1_plan *plan, int atom_tile_size, int j_atom_begin, int n_tile, int j_coord,
                            int n_atoms, const int *atom_valid_n_compute_c_batches, const double *wave,
                            const double *gradient_wave, const double *H_times_psi, const double *vrho_batches,
                            const double *vsigma_batches, const double *gradient_rho, const double *partition_tab,
                            double *first_order_H_tile, void *stream);

/* ------------------------------------------------------------------------------------------ */
/* Far-field Hartree multipole sum                                                            */
/* ------------------------------------------------------------------------------------------ */
typedef struct frpa_far_plan frpa_far_plan;

#define FRPA_FAR_TERMS 0   /* term-by-term Cartesian sum, as the reference (plain FP64 baseline) */
#define FRPA_FAR_HORNER 1  /* merged (i,j,k,nn) coefficients, nested Horner; FP64 throughout      */

/*
 * Geometry and tables (host arrays, copied).  Arguments keep the reference's names and layouts
 * (fast_rap_src_dfpt_gpu_eval/sum_up_whole_potential.h:3-48):
 *   species[n_atoms], l_hartree[n_species], centers_hartree_potential[n_centers_hartree_potential],
 *   center_to_atom[n_centers], coords_center (3, n_centers), index_cc (index_cc_dim_0, 6),
 *   cc[index_cc_dim_0], n_cc_lm_ijk[l_max_analytic_multipole + 1],
 *   index_ijk_max_cc (3, l_max_analytic_multipole + 1).
 * l_pot_max <= 10.
 */
int frpa_far_plan_create(frpa_far_plan **plan, int l_pot_max, int l_max_analytic_multipole, int index_cc_dim_0,
                         int n_centers, int n_centers_hartree_potential, int n_atoms, int n_species,
                         int hartree_force_l_add, const int *species, const int *l_hartree,
                         const int *centers_hartree_potential, const int *center_to_atom,
                         const double *coords_center, const int *index_cc, const double *cc,
                         const int *n_cc_lm_ijk, const int *index_ijk_max_cc);
int frpa_far_plan_destroy(frpa_far_plan *plan);

/*
 * delta_v_hartree(p, t) += sum_centres far-field term, for the valid points, t < n_tile
 * (far branch of sum_up_whole_potential.cpp:382-503; pairs with dist^2 < multipole_radius_sq are
 * the near field and are not computed here).  Device arrays:
 *   coord_points (3, n_full_points), i_valid_point_2_i_full_points_map[n_valid_points],
 *   l_hartree_max_far_distance (n_atoms, T) int, multipole_radius_sq (n_atoms, T),
 *   outer_potential_radius (l_pot_max + 1, n_atoms, T), multipole_moments ((l_pot_max+1)^2, n_atoms, T),
 *   delta_v_hartree (n_full_points, T) ACCUMULATED into.
 */
int frpa_far_forward(frpa_far_plan *plan, int atom_tile_size, int n_tile, int variant, int n_full_points,
                     int n_valid_points, const double *coord_points, const int *i_valid_point_2_i_full_points_map,
                     const int *l_hartree_max_far_distance, const double *multipole_radius_sq,
                     const double *outer_potential_radius, const double *multipole_moments,
                     double *delta_v_hartree, void *stream);

/*
 * frpa_far_forward without any allocation: same arguments, kernels and result bits, for callers that
 * cannot allocate device memory during the call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The per-call coefficient
 * table ((n_cc(l_pot_max) for FRPA_FAR_TERMS, or the merged Horner count for FRPA_FAR_HORNER) x
 * n_atoms x atom_tile_size doubles; the Horner count is 495 at l_pot_max = 8, 1001 at 10) goes into a
 * fixed array of FRPA_FAR_NOALLOC_SCRATCH_DOUBLES doubles (32 MiB) that each device the library runs
 * on holds until the process ends; a call needing more returns FRPA_ERR_UNSUPPORTED before launching
 * anything.  The plan must exist (frpa_far_plan_create allocates; this call does not).  The first
 * call on a device looks up the array's address there (the current device must be the stream's).
 * Every call on a device shares that device's array: calls that overlap on one GPU (other streams,
 * other threads) race; calls in one stream are ordered.
 */
#define FRPA_FAR_NOALLOC_SCRATCH_DOUBLES 4194304
int frpa_far_forward_noalloc(frpa_far_plan *plan, int atom_tile_size, int n_tile, int variant, int n_full_points,
                             int n_valid_points, const double *coord_points,
                             const int *i_valid_point_2_i_full_points_map, const int *l_hartree_max_far_distance,
                             const double *multipole_radius_sq, const double *outer_potential_radius,
                             const double *multipole_moments, double *delta_v_hartree, void *stream);

/*
 * Adjoint with respect to the multipole moments:
 *   grad_moments(lm, a, t) += sum_points gV(p, t) d delta_v(p, t) / d M(lm, a, t)
 * gV (n_full_points, T) device; grad_moments ((l_pot_max+1)^2, n_atoms, T) device, ACCUMULATED.
 * Moments with |M| <= 1e-10 are treated as the reference treats them (their terms dropped), so
 * their gradient entries receive 0.
 */
int frpa_far_adjoint(frpa_far_plan *plan, int atom_tile_size, int n_tile, int n_full_points, int n_valid_points,
                     const double *coord_points, const int *i_valid_point_2_i_full_points_map,
                     const int *l_hartree_max_far_distance, const double *multipole_radius_sq,
                     const double *outer_potential_radius, const double *multipole_moments, const double *gV,
                     double *grad_moments, void *stream);

/*
 * Near-field (spline) tables of the hartree centres, host arrays, copied: species_center[n_centers]
 * (1-based), and per species n_grid[n_species], r_grid_min[n_species], log_r_grid_inc[n_species]
 * (the arrays of the same names in sum_up_whole_potential.h:3-48).  Required before
 * frpa_near_forward; n_grid >= 2, r_grid_min > 0, log_r_grid_inc > 0.
 */
int frpa_far_plan_set_near(frpa_far_plan *plan, int n_species, const int *species_center, const int *n_grid,
                           const double *r_grid_min, const double *log_r_grid_inc);

/*
 * delta_v_hartree(p, t) += sum over hartree centres c with dist^2 < multipole_radius_sq(atom, t) of
 *   sum_{lm < (l+1)^2} Y_lm(dir) S_lm(i_r_log), i_r_log = 1 + log(dist / r_grid_min) / log_r_grid_inc,
 * S_lm the cubic Hermite spline of current_delta_v_hart_part_spl_tile on the log-grid index, l the
 * atom's l_hartree lowered by outer_potential_radius (near branch of sum_up_whole_potential.cpp:382-477;
 * sum_up_close_part_cu, sum_up_whole_potential.cu:168-371).  Device arrays:
 *   coord_points (3, n_full_points), i_valid_point_2_i_full_points_map[n_valid_points],
 *   multipole_radius_sq (n_atoms, T), outer_potential_radius (l_pot_max + 1, n_atoms, T),
 *   current_delta_v_hart_part_spl_tile ((l_pot_max+1)^2, 2, n_hartree_grid, n_atoms, T),
 *   delta_v_hartree (n_full_points, T) ACCUMULATED into.
 * n_hartree_grid >= n_grid of every centre's species.  l_pot_max <= 10 (the reference's device code
 * stops at 6).  frpa_near_forward followed by frpa_far_forward is the whole sum_up_whole_potential.
 */
int frpa_near_forward(frpa_far_plan *plan, int atom_tile_size, int n_tile, int n_full_points, int n_valid_points,
                      int n_hartree_grid, const double *coord_points, const int *i_valid_point_2_i_full_points_map,
                      const double *multipole_radius_sq, const double *outer_potential_radius,
                      const double *current_delta_v_hart_part_spl_tile, double *delta_v_hartree, void *stream);

/* ------------------------------------------------------------------------------------------ */
/* Synthetic wave generation on the device (benchmark inputs too large to build on the host)   */
/* ------------------------------------------------------------------------------------------ */
/*
 * wave(off_b + i + nc_pad * p) = exp(-alpha_mu |r_p - R_a|^2) Y_{l_mu m_mu}(r^) for mu = ibasis(i) - 1,
 * padding rows set to NaN.  basis_atom/basis_l/basis_m/basis_alpha (n_basis) and atom_xyz (3, n_atoms)
 * are host arrays; coord_points (3, n_full_points) is a device array.  l <= 2.
 */
int frpa_synth_wave(frpa_rho1_plan *plan, int n_atoms, const double *atom_xyz, const int *basis_atom,
                    const int *basis_l, const int *basis_m, const double *basis_alpha, const double *coord_points,
                    double *wave, void *stream);

/* ------------------------------------------------------------------------------------------ */
/* Reference entry points (drop-in names; host arrays; exit(1) on error)                       */
/* ------------------------------------------------------------------------------------------ */

/* The reference declares init_dfpt_device_info / free_dfpt_device_info / init_dfpt_device_data /
 * free_dfpt_device_data and sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_ with
 * C++ linkage.  The library exports each of them twice: with C linkage (declared here) and with C++
 * linkage and the reference's parameter lists (the mangled names a host compiled against the
 * reference's device.hpp, device_data.hpp and sum_up_whole_potential.h links to); both forms run
 * the same code. */

/* fast_rap_src_dfpt_gpu_eval/device.hpp:191-192. */
void init_dfpt_device_info(void);
void free_dfpt_device_info(void);

/* fast_rap_src_dfpt_gpu_eval/device_data.hpp:159-227.
 * Uploads the constant arrays once and builds the rho1 and far-field plans; allocates (does not
 * copy) the per-call rho1 arrays, as the reference.  The reference's compile-time ATOM_TILE_SIZE is
 * taken from the environment variable FRPA_ATOM_TILE_SIZE (default 1), since this signature has no
 * argument for it; FRPA_ALL_KERNELS_ON_GPU=1 selects the device-resident mode below. */
void init_dfpt_device_data(
    int *n_my_batches_work_, int *n_full_points_, int *n_valid_points_, int *n_basis_, int *n_atoms_,
    int *n_max_compute_ham_, int *n_centers_basis_I_, int *n_max_batch_size_, int *l_pot_max_, int *n_max_radial_,
    int *n_hartree_grid_, int *n_centers_, int *n_centers_hartree_potential_, int *hartree_force_l_add_,
    int *n_species_, int *l_max_analytic_multipole_, int *index_cc_dim_0_,
    int *basis_atom_ptr, int *batch_sizes_ptr, int *n_point_batches_ptr, int *n_point_batches_prefix_sum_ptr,
    int *i_valid_point_2_i_full_points_map_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *i_batch_2_wave_offset_ptr, double *wave_batches_compress_ptr,
    double *gradient_wave_batches_compress_ptr, double *density_matrix_ptr,
    int *n_radial_ptr, int *species_ptr, int *species_center_ptr, double *coord_points_ptr, int *l_hartree_ptr,
    int *n_grid_ptr, int *n_cc_lm_ijk_ptr, int *centers_hartree_potential_ptr, int *center_to_atom_ptr,
    int *index_cc_ptr, int *index_ijk_max_cc_ptr, double *cc_ptr, double *coords_center_ptr,
    double *r_grid_min_ptr, double *log_r_grid_inc_ptr, double *scale_radial_ptr, double *partition_tab_ptr,
    double *first_order_density_matrix_ptr, double *first_order_rho_ptr, double *first_order_rho_bias_part2_ptr,
    double *first_order_gradient_rho_ptr, double *first_order_gradient_rho_bias_batches_atoms_ptr,
    int *l_hartree_max_far_distance_ptr, double *multipole_radius_sq_ptr, double *outer_potential_radius_ptr,
    double *multipole_moments_ptr, double *current_delta_v_hart_part_spl_tile_ptr, double *delta_v_hartree_ptr);
void free_dfpt_device_data(void);

/* Device-resident mode, the run-time counterpart of the reference's compile-time ALL_KERNELS_ON_GPU
 * (evaluate_first_order_rho.cu:17,259-263,363-367; its runner fills and reads the global devPs
 * arrays, direct_test/evaluate_first_order_rho_run.cpp:163-166,200-202): on, the rho1 _cu_host_
 * function reads first_order_density_matrix and the two bias arrays from device arrays and leaves
 * first_order_rho and first_order_gradient_rho in device arrays, without copies and without
 * synchronising; its host pointer arguments are not used.  frpa_ref_device_array returns the device
 * address of the array of that name (first_order_density_matrix, first_order_rho_bias_part2,
 * first_order_gradient_rho_bias_batches_atoms, first_order_rho, first_order_gradient_rho,
 * wave_batches_compress, gradient_wave_batches_compress, coord_points, density_matrix (the copy made
 * by init_dfpt_device_data, device_data.cu:328, used by the inline bias); the layouts and sizes of the
 * reference's arguments with ATOM_TILE_SIZE = FRPA_ATOM_TILE_SIZE), NULL before
 * init_dfpt_device_data or for another name; frpa_ref_stream the stream the entry points run on
 * (the reference's devInfo.stream).  Off by default (the environment variable FRPA_ALL_KERNELS_ON_GPU=1
 * at init_dfpt_device_data, or this setter, turns it on). */
void frpa_ref_set_all_kernels_on_gpu(int on);
double *frpa_ref_device_array(const char *name);
void *frpa_ref_stream(void);

/* fast_rap_src_dfpt_gpu_eval/evaluate_first_order_rho.h:37-69.  Runs on the device-resident
 * batch data of init_dfpt_device_data; uploads first_order_density_matrix and the bias arrays from
 * the given host pointers and downloads first_order_rho and first_order_gradient_rho into them
 * (the reference's non-ALL_KERNELS_ON_GPU mode, evaluate_first_order_rho.cu:240-244,362-366), or
 * works on the device arrays only in the device-resident mode above.
 * Precision from the environment variable FRPA_RHO1_PRECISION: ozaki (the default: INT8 split with 5
 * digits, ~1e-12 of max|rho1| and of max|grad rho1|), ozaki4 (4 digits) or fp64 (FP64 MFMA);
 * first_order_gradient_rho is computed (GGA) unless FRPA_RHO1_GRADIENT=0 (then left zero, and the INT8
 * split multiplies only the symmetric half of P1).  FRPA_RHO1_INLINE_BIAS=1 selects the v6 variant
 * (task4_common_calculation_eliminated/version6_rho_with_bias_calculate/evaluate_first_order_rho.cpp:
 * 179-231): rho1 also gets the inline bias of frpa_rho1_bias_inline for atoms j_atom_begin..j_atom_end
 * and coordinate j_coord, from the density_matrix of init_dfpt_device_data; the bias input arrays are
 * not modified. */
void evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_(
    int *n_my_batches_work_, int *j_atom_begin_, int *j_atom_end_, int *j_coord_, int *n_full_points_,
    int *n_basis_, int *n_atoms_, int *n_max_compute_ham_, int *n_centers_basis_I_, int *n_max_batch_size_,
    int *batch_sizes_ptr, int *n_point_batches_ptr, int *n_point_batches_prefix_sum_ptr,
    int *i_valid_point_2_i_full_points_map_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *i_batch_2_wave_offset_ptr, double *wave_batches_compress_ptr,
    double *gradient_wave_batches_compress_ptr, double *first_order_density_matrix_ptr, double *first_order_rho_ptr,
    double *first_order_rho_bias_part2_ptr, double *first_order_gradient_rho_ptr,
    double *first_order_gradient_rho_bias_batches_atoms_ptr, int *basis_atom_ptr, double *partition_tab_ptr,
    double *density_matrix_ptr);

/* fast_rap_src_dfpt_gpu_eval/evaluate_first_order_rho.h:3-35: the reference's CPU function (same
 * name, C linkage, same 28 arguments, host arrays), run on the GPU.  Needs no init_dfpt_* call: it
 * keeps a plan of its own, rebuilt when the batch metadata differs from the previous call's, and
 * uploads the wave arrays, first_order_density_matrix and the bias arrays on every call (the CPU
 * function reads them afresh); first_order_rho is overwritten (zeroed, then the tile's
 * perturbations written) and first_order_gradient_rho computed, as evaluate_first_order_rho.cpp:
 * 113-209, with the precision switches of the _cu_host_ function.  With FRPA_RHO1_INLINE_BIAS=1 it
 * computes the v6 variant (see the _cu_host_ function) from this call's density_matrix and
 * atom_valid_n_compute_c_batches, which are then uploaded too. */
void evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_(
    int *n_my_batches_work_, int *j_atom_begin_, int *j_atom_end_, int *j_coord_, int *n_full_points_,
    int *n_basis_, int *n_atoms_, int *n_max_compute_ham_, int *n_centers_basis_I_, int *n_max_batch_size_,
    int *batch_sizes_ptr, int *n_point_batches_ptr, int *n_point_batches_prefix_sum_ptr,
    int *i_valid_point_2_i_full_points_map_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *i_batch_2_wave_offset_ptr, double *wave_batches_compress_ptr,
    double *gradient_wave_batches_compress_ptr, double *first_order_density_matrix_ptr, double *first_order_rho_ptr,
    double *first_order_rho_bias_part2_ptr, double *first_order_gradient_rho_ptr,
    double *first_order_gradient_rho_bias_batches_atoms_ptr, int *basis_atom_ptr, double *partition_tab_ptr,
    double *density_matrix_ptr);

/* fast_rap_src_dfpt_gpu_eval/sum_up_whole_potential.h:3-48: the reference's CPU function (same name,
 * C linkage, scalars by pointer), run on the GPU: both branches (near-field splines, far-field
 * multipoles) accumulated into delta_v_hartree (n_full_points, ATOM_TILE_SIZE), as
 * sum_up_whole_potential.cpp:195-529.  Needs no init_dfpt_* call: it keeps a far-field plan of its
 * own, rebuilt when the geometry or term tables differ from the previous call's. */
void sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_(
    const int *j_atom_begin_, const int *j_atom_end_, const int *n_full_points_, const int *l_max_analytic_multipole_,
    const int *index_cc_dim_0_, const int *n_valid_points_, const int *l_pot_max_, const int *n_max_radial_,
    const int *n_hartree_grid_, const int *n_centers_, const int *n_centers_hartree_potential_, const int *n_atoms_,
    const int *hartree_force_l_add_, const int *n_species_, const int *n_radial_ptr, const int *species_ptr,
    const int *species_center_ptr, const int *l_hartree_max_far_distance_ptr,
    const int *i_valid_point_2_i_full_points_map_ptr, const double *coord_points_ptr,
    const double *multipole_radius_sq_ptr, const double *outer_potential_radius_ptr,
    const double *multipole_moments_ptr, const double *current_delta_v_hart_part_spl_tile_ptr,
    const int *l_hartree_ptr, const int *n_grid_ptr, const int *n_cc_lm_ijk_ptr,
    const int *centers_hartree_potential_ptr, const int *center_to_atom_ptr, const int *index_cc_ptr,
    const int *index_ijk_max_cc_ptr, const double *cc_ptr, const double *coords_center_ptr,
    const double *r_grid_min_ptr, const double *log_r_grid_inc_ptr, const double *scale_radial_ptr,
    const double *partition_tab_ptr, double *delta_v_hartree_ptr);

/* fast_rap_src_dfpt_gpu_eval/generate_first_order_rho_bias_part2.cpp:9-30 (a CPU function there;
 * same name, C linkage, 21 pointer arguments).  Runs on the GPU with the batch data that
 * init_dfpt_device_data put on the device (which must come first); uploads density_matrix and
 * first_order_rho_bias_part2, writes the tile's bias and copies it back.  The reference file fixes
 * the array's tile extent at 4 (its line 7); here it is ATOM_TILE_SIZE (FRPA_ATOM_TILE_SIZE, set it
 * to 4 for the reference's layout). */
void generate_first_order_rho_bias_part2_(
    int *n_my_batches_work_, int *j_atom_begin_, int *j_atom_end_, int *j_coord_, int *n_full_points_,
    int *n_basis_, int *n_atoms_, int *n_max_compute_ham_, int *n_centers_basis_I_, int *n_max_batch_size_,
    int *batch_sizes_ptr, int *n_point_batches_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *i_batch_2_wave_offset_ptr, double *wave_batches_compress_ptr,
    double *gradient_wave_batches_compress_ptr, double *first_order_rho_bias_part2_ptr, double *partition_tab_ptr,
    double *density_matrix_ptr);

/* fast_rap_src_dfpt_gpu_eval/evaluate_first_order_H.h:3-43 (a CPU function there; same name, C
 * linkage, 34 pointer arguments).  Runs frpa_rho1_h1 on the GPU with the wave and gradient wave that
 * init_dfpt_device_data put on the device (which must come first); uploads the hessian wave, the
 * XC and per-point arrays and first_order_H_tile, accumulates and copies first_order_H_tile back.
 * vrho_batches and H_times_psi_batches_compress are not read by the reference either. */
void evaluate_first_order_h_reduce_memory_c_v2_(
    int *j_atom_begin_, int *j_atom_end_, int *n_full_points_, int *j_coord_, int *n_basis_, int *n_atoms_,
    int *n_max_compute_ham_, int *n_max_batch_size_, int *n_my_batches_work_, int *n_centers_basis_I_,
    int *n_point_batches_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *n_point_batches_prefix_sum_ptr,
    int *i_valid_point_2_i_full_points_map_ptr, double *vrho_batches_ptr, double *vsigma_batches_ptr,
    double *gradient_rho_ptr, double *H_times_psi_batches_compress_ptr, int *i_batch_2_wave_offset_ptr,
    double *wave_batches_compress_ptr, double *gradient_wave_batches_compress_ptr,
    double *hessian_wave_batches_compress_ptr, double *first_order_H_tile_ptr, double *H_prefactor1_batches_tile_ptr,
    double *first_order_gradient_rho_batches_tile_ptr, double *v_hartree_gradient_ptr, double *first_order_rho_ptr,
    double *v2rho2_batches_ptr, double *v2rhosigma_batches_ptr, double *v2sigma2_batches_ptr,
    double *partition_tab_ptr, int *not_use);

/* fast_rap_src_dfpt_gpu_eval/evaluate_first_order_H_bias_part2.cpp:13-52 (a CPU function there; same
 * name, C linkage, 33 pointer arguments).  Runs frpa_rho1_h1_bias_part2 on the GPU with the wave and
 * gradient wave that init_dfpt_device_data put on the device (which must come first); uploads
 * H_times_psi, vrho, vsigma, gradient_rho, partition_tab and first_order_H_tile, accumulates and
 * copies first_order_H_tile back.  The arguments the reference does not read are not read here. */
void evaluate_first_order_h_bias_part2_(
    int *j_atom_begin_, int *j_atom_end_, int *n_full_points_, int *j_coord_, int *n_basis_, int *n_atoms_,
    int *n_max_compute_ham_, int *n_max_batch_size_, int *n_my_batches_work_, int *n_centers_basis_I_,
    int *n_point_batches_ptr, int *n_compute_c_batches_ptr, int *i_basis_batches_ptr,
    int *atom_valid_n_compute_c_batches_ptr, int *n_point_batches_prefix_sum_ptr,
    int *i_valid_point_2_i_full_points_map_ptr, double *vrho_batches_ptr, double *vsigma_batches_ptr,
    double *gradient_rho_ptr, double *H_times_psi_batches_compress_ptr, int *i_batch_2_wave_offset_ptr,
    double *wave_batches_compress_ptr, double *gradient_wave_batches_compress_ptr, double *first_order_H_tile_ptr,
    double *H_prefactor1_batches_tile_ptr, double *first_order_gradient_rho_batches_tile_ptr,
    double *v_hartree_gradient_ptr, double *first_order_rho_ptr, double *v2rho2_batches_ptr,
    double *v2rhosigma_batches_ptr, double *v2sigma2_batches_ptr, double *partition_tab_ptr, int *not_use);

/* fast_rap_src_dfpt_gpu_eval/sum_up_whole_potential.h:50-95.
 * Both branches, as the reference (sum_up_whole_potential.cu:641-769): uploads the per-call arrays
 * (including current_delta_v_hart_part_spl_tile), runs the near-field spline sum and the far-field
 * multipole sum, accumulates into delta_v_hartree and synchronises.  species_center, n_grid,
 * r_grid_min and log_r_grid_inc must describe valid log grids (n_grid >= 2, n_grid <= n_hartree_grid). */
void sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_(
    const int j_atom_begin, const int j_atom_end, const int n_full_points, const int l_max_analytic_multipole,
    const int index_cc_dim_0, const int n_valid_points, const int l_pot_max, const int n_max_radial,
    const int n_hartree_grid, const int n_centers, const int n_centers_hartree_potential, const int n_atoms,
    const int hartree_force_l_add, const int n_species, const int *n_radial_ptr, const int *species_ptr,
    const int *species_center_ptr, const int *l_hartree_max_far_distance_ptr,
    const int *i_valid_point_2_i_full_points_map_ptr, const double *coord_points_ptr,
    const double *multipole_radius_sq_ptr, const double *outer_potential_radius_ptr,
    const double *multipole_moments_ptr, const double *current_delta_v_hart_part_spl_tile_ptr,
    const int *l_hartree_ptr, const int *n_grid_ptr, const int *n_cc_lm_ijk_ptr,
    const int *centers_hartree_potential_ptr, const int *center_to_atom_ptr, const int *index_cc_ptr,
    const int *index_ijk_max_cc_ptr, const double *cc_ptr, const double *coords_center_ptr,
    const double *r_grid_min_ptr, const double *log_r_grid_inc_ptr, const double *scale_radial_ptr,
    const double *partition_tab_ptr, double *delta_v_hartree_ptr);

#ifdef __cplusplus
}
#endif
#endif /* FRPA_H */
