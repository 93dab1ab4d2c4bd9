/*
 * pxh.h -- C ABI of libpxh, the MI355X (gfx950) frequency-response solver.
 *
 * A drop-in for the reference's frequency-dependent response path
 * (src/solvers, src/matrix, src/wfq, src/rhs, src/bem).  The Fortran host
 * calls it through iso_c_binding (plasmonx_amd/fortran/pxh.f90); Python
 * through ctypes (plasmonx_amd/_lib.py).
 *
 * Conventions (those of the reference's Fortran data):
 *   - integers are int32 (the reference's default INTEGER), reals double;
 *   - complex arrays are interleaved (re, im) doubles, the memory layout of
 *     Fortran complex(dp) / complex(c_double_complex);
 *   - arrays are column-major; atom types are 1-based as map_atomtypes;
 *   - per-type-pair tables are n_types x n_types, element (ti, tj) at
 *     [(ti-1) + n_types*(tj-1)];
 *   - the unknowns of omegaFQFmu are [q_1..q_N, mu_1x, mu_1y, mu_1z, ...]
 *     (n_var = 4N), those of omegaFQ [q_1..q_N], those of BEM one value per
 *     tessera (DPCM: surface charge density, IEF: charge);
 *   - all lengths in bohr, frequencies in atomic units.
 *
 * Ownership: the host owns every array it passes.  Inputs are copied to HBM
 * by the pxh_set_* calls and stay resident in the context; outputs are
 * written before the call returns; no pointer is kept after a call returns.
 * A context is used by one host thread at a time.
 *
 * Status codes: 0 ok; 1 bad argument; 2 GMRES not converged within the
 * Krylov dimension (reference: fatal, gmres_module.f90:216); 3 singular
 * pivot (reference: info /= 0, algorithm_module.f90:529,539); 4 HBM budget
 * exceeded (reference: memory_manager_module.f90:394-397); 5 HIP runtime
 * error; 6 the HIP device or library is unavailable.  pxh_last_error()
 * returns the message of the last failure.
 */
#ifndef PXH_H
#define PXH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pxh_ctx pxh_ctx;

enum {
    PXH_OK = 0,
    PXH_EINVAL = 1,
    PXH_ENOCONV = 2,
    PXH_ESINGULAR = 3,
    PXH_ENOMEM = 4,
    PXH_EHIP = 5,
    PXH_ENODEV = 6
};

/* target models (reference target_%name_: wfq, wfqfmu, bem; static fq, fqfmu) */
enum { PXH_WFQ = 1, PXH_WFQFMU = 2, PXH_BEM = 3, PXH_FQ = 4, PXH_FQFMU = 5 };
/* keyword 'what' of the static models: energy | static response */
enum { PXH_ENERGY = 0, PXH_STATIC_FIELD = 1 };
/* charge-charge kernels (keyword 'forcefield kernel') */
enum { PXH_GAUSSIAN = 1, PXH_COULOMB = 2, PXH_OHNO = 3 };
/* keyword 'algorithm method': inversion | iterative | iterative on the fly */
enum { PXH_INVERSION = 0, PXH_ITERATIVE = 1, PXH_ITERATIVE_OTF = 2 };
/* arithmetic of the solve */
enum { PXH_FP64 = 0, PXH_MIXED = 1 };
/* keyword 'field rhs type': field | potential */
enum { PXH_RHS_FIELD = 1, PXH_RHS_POTENTIAL = 2 };
/* BEM keywords 'bem variant' and 'bem green function' */
enum { PXH_DPCM = 1, PXH_IEFPCM = 2 };
enum { PXH_APPROXIMATE = 1, PXH_ACCURATE = 2 };

/* Create / destroy a context on HIP device `device`.  hbm_budget_bytes = 0
 * means "all free HBM".  Replaces the reference's global state set up by
 * inp_%read_ (src/input_output/input_module.f90:156-258). */
int pxh_ctx_create(int device, size_t hbm_budget_bytes, pxh_ctx **out);
int pxh_ctx_destroy(pxh_ctx *ctx);

/* Atomistic target (omegaFQ / omegaFQFmu).  Replaces the data the reference
 * keeps in target_ and parameters (target_module.f90:20-131,
 * parameters_module.f90:64-98) and get_neighbours (target_module.f90:417-464):
 *   coord_bohr 3 x n_atoms; type_of_atom n_atoms (1-based);
 *   eta, r_q, r_mu, a_ij, tau, density: n_types (tau and density after the
 *   'scaling sigma0-tau' of wfq_module.f90:96-127);
 *   fermi_d, fermi_s, r0: n_types x n_types (r0 in bohr,
 *   target_module.f90:472-518);
 *   nb_counts n_types x n_atoms and nb_total n_atoms: neighbours within r0
 *   (self included), used by the omegaFQFmu diagonal; may be NULL for wfq. */
int pxh_set_atoms(pxh_ctx *ctx, int model, int kernel, int n_atoms, const double *coord_bohr,
                  const int32_t *type_of_atom, int n_types, const double *eta, const double *r_q,
                  const double *r_mu, const double *a_ij, const double *fermi_d, const double *fermi_s,
                  const double *r0, const double *tau, const double *density, const int32_t *nb_counts,
                  const int32_t *nb_total);

/* Static FQ / FQFMu targets (model PXH_FQ / PXH_FQFMU in pxh_set_atoms, where
 * a_ij, the Fermi tables, tau, density and the neighbour counts are unused
 * and may be NULL): the electronegativities chi and, for FQFMu, the atomic
 * polarizabilities alpha per type (a.u.), and the molecules as runs of
 * consecutive atoms, atoms_per_molecule[n_mol] (each molecule's charges sum
 * to zero: one Lagrange multiplier each).  Replaces the chi / alpha of
 * target_ and create_n_atoms_per_molecule (input_module.f90:1007-1031); the
 * reference's add_lagrangian_blocks (matrix_module.f90:351-382) places
 * molecules correctly only when they all have the same size, this one for
 * any sizes. */
int pxh_set_static(pxh_ctx *ctx, int n_types, const double *chi, const double *alpha, int n_mol,
                   const int32_t *atoms_per_molecule);

/* Static solve of an FQ / FQFMu target, unknowns [q_1..q_N, lambda_1..
 * lambda_M, mu_1x, mu_1y, mu_1z, ...] (n_var = N + n_mol, + 3N for FQFMu).
 * Replaces algorithm%solve_ground_state and solve_static_field
 * (algorithm_module.f90:95-148) with real_inversion (:469-504) and the
 * models' right-hand sides, energies and polarizability
 * (fq_module.f90:179-597, fqfmu_module.f90:81-349).
 *   what:      PXH_ENERGY (right-hand side -chi; energy = -1/2 x^T b) or
 *              PXH_STATIC_FIELD (three right-hand sides: rhs_form potential
 *              of the field e0 along x, y, z on the charge rows, and for
 *              FQFMu 1 on dipole component k, 0 when e0 is 0, as
 *              rhs_module.f90:96-127 writes it);
 *   variables: n_var x 1 (energy) or n_var x 3 (static field) real, may be NULL;
 *   polar:     3 x 3 static polarizability of FQ (static field), may be NULL;
 *              FQFMu returns PXH_EINVAL for it, as the reference stops there
 *              (fqfmu_module.f90:340-347);
 *   energy:    may be NULL. */
int pxh_solve_static(pxh_ctx *ctx, int what, double e0, int rhs_form, double *variables, double *polar,
                     double *energy);

/* Interband polarizability alpha_t(omega) per type and frequency (complex,
 * n_types x n_freq).  Replaces read_alpha (wfqfmu_module.f90:97-166). */
int pxh_set_alpha(pxh_ctx *ctx, int n_types, int n_freq, const double *alpha_w);

/* BEM target.  verts9: 9 x n_tess, the three vertices (in the order of the
 * GMSH face) of each triangle in bohr.  Replaces read_gmsh_file +
 * create_tesserae (bem_module.f90:224-491) and the bem keywords. */
int pxh_set_mesh(pxh_ctx *ctx, int variant, int green, int n_tess, const double *verts9, double normal_factor,
                 double sphere_r_bohr, double eps_solvent, int charge_constraint);

/* Permittivity eps(omega) of the BEM metal (complex, n_freq).  Replaces
 * read_permittivity_bem (bem_module.f90:154-219). */
int pxh_set_permittivity(pxh_ctx *ctx, int n_freq, const double *eps_w);

/* Solve the frequency-dependent response for n_freq frequencies and the
 * field directions of pol_mask (bit 0 = x, 1 = y, 2 = z).  Replaces the body
 * of algorithm%solve_dynamic_field (algorithm_module.f90:150-179) with its
 * three loops solve_w_inversion / solve_w_iter_mem / solve_w_iter_on_the_fly
 * (:189-465).
 *   method:      PXH_INVERSION | PXH_ITERATIVE | PXH_ITERATIVE_OTF
 *                (PXH_INVERSION factorises up to 128 frequencies at once,
 *                within 32 GB of matrices, half the free HBM and the
 *                context's budget; the factorisation is of A^T, so its
 *                partial pivoting picks along rows of A, a rounding-level
 *                difference from ZGETRF);
 *   tol_abs:     absolute GMRES threshold, already scaled by sqrt(n_var) as
 *                input_module.f90:251 does; the true residual
 *                ||b - A x|| < tol_abs is required on return;
 *   krylov_dim:  'algorithm gmres dimension' (no restart, as the reference);
 *   precision:   PXH_FP64 (every operation in FP64) or PXH_MIXED (FP32
 *                interaction tiles inside FP64 iterative refinement; the
 *                stopping test is the same FP64 true residual);
 *   variables:   n_var x 3 x n_freq complex, may be NULL;
 *   polar:       3 x 3 x n_freq complex (polar_w per frequency);
 *   results:     n_freq x 20, column-major (target_%results);
 *   iters:       3 x n_freq Krylov iterations (0 for inversion), may be NULL;
 *   true_residual: 3 x n_freq, may be NULL. */
int pxh_solve_spectrum(pxh_ctx *ctx, int n_freq, const double *freq_au, double e0, int rhs_form, int pol_mask,
                       int method, double tol_abs, int krylov_dim, int precision, double *variables, double *polar,
                       double *results, int32_t *iters, double *true_residual);

/* Dense inversion ('inversion' method) of many atomistic targets at once:
 * ctxs[0..n_ctx) each hold a homogeneous omegaFQ / omegaFQFmu target (and
 * its alpha table), all on one device; every target is solved at the same
 * n_freq frequencies.  The targets are grouped by size and the (target,
 * frequency) matrices of a group are factorised together, padded to the
 * group's largest n, so one launch of each LU kernel covers many targets
 * (no counterpart in the reference, which solves one structure per run: the
 * loop of solve_w_inversion, algorithm_module.f90:189-265, over targets).
 * Homogeneous omegaFQ targets for which one eigendecomposition is cheaper
 * than their LUs (option "spectral") are solved by it, one after another on
 * their own contexts' streams.
 * Work runs on ctxs[0]'s stream; errors are reported by ctxs[0].
 *   polar:   n_ctx x n_freq x (3 x 3) complex, may be NULL;
 *   results: n_ctx blocks of n_freq x 20 (column-major each), may be NULL. */
int pxh_solve_group(int n_ctx, pxh_ctx *const *ctxs, int n_freq, const double *freq_au, double e0, int rhs_form,
                    int pol_mask, double *polar, double *results);

/* y = A(omega_i) x for n_vec vectors of n_var complex values: the on-the-fly
 * product plus the diagonal shift.  Replaces target_%product_matrix_vector
 * followed by target_%new_gmres_diagonal_shift (gmres_module.f90:111-121;
 * wfq_module.f90:321-348,398-424; wfqfmu_module.f90:584-626,698-746;
 * bem_module.f90:977-997).  i_freq indexes the table of pxh_set_alpha /
 * pxh_set_permittivity (0-based). */
int pxh_matvec(pxh_ctx *ctx, double freq_au, int i_freq, int n_vec, const double *x, double *y);

/* y = A(omega_i) x as pxh_matvec, in FP64, for atomistic targets, with x and
 * y in device memory the caller owns (n_var x n_vec complex each), and
 * without any device allocation: the product's working memory is a 64 MiB
 * array the library holds on each GPU it runs on until the process ends.
 * No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Returns PXH_ENOMEM when n_vec vectors need more than that
 * array (about n_var x n_vec x (splits + 1) x 16 bytes); every call on one
 * GPU shares the array, so calls that overlap on one GPU (other threads,
 * other contexts' streams) race.  The results are complete on return. */
int pxh_matvec_noalloc(pxh_ctx *ctx, double freq_au, int i_freq, int n_vec, const double *x_dev, double *y_dev);

/* The frequency-independent real matrix, n_var x n_var column-major.
 * Replaces target_%construct_constant_matrix (wfq_module.f90:192-206,
 * wfqfmu_module.f90:275-323, bem_module.f90:678-782). */
int pxh_constant_matrix(pxh_ctx *ctx, double *m);

/* The dynamic right-hand sides, n_var x 3 complex.  Replaces
 * target_%construct_dynamic_field_rhs(_on_the_fly) and, for alloys,
 * assign_variables_w (wfq_module.f90:260-301; wfqfmu_module.f90:177-266,
 * 329-428; rhs_module.f90:43-430; bem_module.f90:788-873). */
int pxh_dynamic_rhs(pxh_ctx *ctx, double freq_au, double e0, int rhs_form, double *rhs);

/* Field at n_points points for the variables of one frequency and one field
 * direction i_pol (0, 1, 2) of an atomistic target with the Gaussian kernel:
 * the incident e0 along i_pol plus the fields of the Gaussian charges and,
 * for omegaFQFmu, dipoles (widths R_q, R_mu of each atom).  points_bohr
 * 3 x n_points; variables n_var complex; field 3 x n_points complex
 * (point-major).  |field| / e0 on a grid is the plasmonX_analysis field
 * cube.  Replaces calculate_induced_field_at_point
 * (wfq_module.f90:592-617, wfqfmu_module.f90:973-1003) with
 * calculate_induced_field_q / _mu (matrix_module.f90:1140-1246) as the
 * per-point loop of calculate_field_3d / _2d calls it
 * (src/plasmonX_analysis/grid_module.f90:200-388). */
int pxh_induced_field(pxh_ctx *ctx, int n_points, const double *points_bohr, const double *variables, int i_pol,
                      double e0, double *field);

/* Field at n_points points of BEM surface charges (n_tess complex, point
 * charges at the tessera centroids, bohr, 3 x n_tess) for field direction
 * i_pol (0, 1, 2): zero at points inside the particle (not farther from the
 * centroids' mean than the nearest centroid is, plus 0.5 Angstrom), else the
 * incident e0 along i_pol plus the Coulomb fields of the charges.  Needs no
 * target in the context.  field 3 x n_points complex.  Replaces
 * calculate_induced_field_at_point_bem (bem_module.f90:1151-1213) with
 * calculate_induced_field_q_bem (matrix_module.f90:1444-1464) in the
 * per-point loops of grid_module.f90:200-388. */
int pxh_bem_induced_field(pxh_ctx *ctx, int n_tess, const double *centroids_bohr, const double *charges, int n_points,
                          const double *points_bohr, int i_pol, double e0, double *field);

/* Plasmon density at n_points points for the variables of one frequency and
 * one field direction of an atomistic target with the Gaussian kernel: the
 * normalised Gaussians (width R_q) of the charges and, for omegaFQFmu, minus
 * the derivatives along the dipoles of the Gaussians of width R_mu.
 * points_bohr 3 x n_points; variables n_var complex; rho n_points complex;
 * rho_q and rho_mu (the charge and dipole parts, rho = rho_q + rho_mu) are
 * written when not NULL.  Replaces calculate_density_at_point and
 * calculate_density_at_point_q_mu_separated (wfq_module.f90:625-648,
 * wfqfmu_module.f90:1012-1081) with calculate_density_q / _mu
 * (matrix_module.f90:1258-1324) as the per-point loop of
 * calculate_density_3d / _2d calls it
 * (src/plasmonX_analysis/grid_module.f90:248-301, 395-453). */
int pxh_plasmon_density(pxh_ctx *ctx, int n_points, const double *points_bohr, const double *variables, double *rho,
                        double *rho_q, double *rho_mu);

/* Copy the message of the last failure into buf (NUL-terminated). */
int pxh_last_error(pxh_ctx *ctx, char *buf, int len);

/* Tuning knobs (no counterpart in the reference): "kbar_cutoff" (1 = drop
 * K-bar pairs whose Fermi factor 1-f is below 1e-18, the default; 0 = keep
 * every pair), "batch" (systems sharing each product), "inner_tol" (relative
 * tolerance of the low-precision inner solves), "precond" (diagonal right
 * preconditioning of the batched path), "mfma" (omegaFQFmu low-precision
 * products on MFMA), "matvec_mixed" (pxh_matvec uses the low-precision
 * product), "splits", "max_outer", "reorth_kappa" (batched classical
 * Gram-Schmidt: second pass only when the first leaves less than kappa of
 * the norm; default 1/sqrt(2), >= 1 = always), "multishift" (1, the
 * default: spectra of BEM DPCM, BEM IEF without `variables`, and
 * homogeneous omegaFQ share one Krylov basis for all frequencies; 0 = the
 * batched solver), "x0_projection" (0..16, default 8: the batched solver
 * starts each system from the best combination of that many converged
 * neighbouring frequencies; 0 = x0 = 0), "initial_guess" (1: the iterative
 * methods start from the caller's `variables`; default 0, x0 = 0 as the
 * reference), "dcgs2" (1: the batched solver's Gram-Schmidt delays the
 * second pass of each vector to the next step), "res_batch" / "res_wait"
 * (FP64 residual products wait for 16 systems, at most 5 steps), "autotune"
 * (1: the first full low-precision pass of a new atom count times candidate
 * source splits and keeps the fastest for the process), "graphs" (1, the
 * default: the multi-shift Arnoldi step on a BEM target is captured once as
 * a HIP graph and replayed, the graph kept with the context while its sizes
 * and buffers stay the same; 0 = every kernel launched directly),
 * "mfma_levels" (1 or 2, default 2: FP32 accumulation levels of the MFMA
 * pair kernel), "fp64_mfma" (1, the default: omegaFQFmu FP64 products of
 * five or more systems contract on the FP64 matrix cores; 0 = the FP64
 * vector kernel), "bem_gemv" (1, the default: BEM FP64 products of up to 4
 * systems, the multi-shift spectra's, by a streaming matrix-vector pass; 0 =
 * the matrix-core tiles), "spectral" (homogeneous omegaFQ inversion, one atom type:
 * 1, the default, solves every frequency from one eigendecomposition of
 * L^T K L, T = L L^T, when that is estimated faster than one LU per
 * frequency; 2 whenever T is positive definite; 0 never; an indefinite T
 * falls back to the LU). */
int pxh_set_option(pxh_ctx *ctx, const char *key, double value);

/* Counters of the last pxh_solve_spectrum call (n <= 32): [0] vector
 * products, [1] FP32 block products, [2] FP64 block products, [3] / [4]
 * pair-kernel ms (FP32 / FP64, HIP events on the context's stream),
 * [5] wall ms, [6] pair evaluations, [7] outer refinements, [8] / [9]
 * pair-kernel launches (FP32 / FP64), [10] / [11] their algorithmic flops,
 * [12] / [13] ms and launches of the MFMA pair kernel alone, [14] / [15]
 * p50 / p99 service time of a frequency point in ms (admission of its first
 * field direction to convergence of its last; iterative methods),
 * [16] / [17] / [18] ms, launches and algorithmic flops of the BEM matrix
 * products (4 flop per matrix element and system), [19] / [20] / [21] ms,
 * launches and flops of the dense LU's trailing updates (8 flop per complex
 * multiply-add), [22] algorithmic bytes of the BEM matrix products (the
 * matrix once plus each vector read and written).  [16]-[18] and [22] count
 * the launches made directly only: the replays of a captured multi-shift
 * step (option "graphs") are neither timed nor counted there.  [23] source
 * splits of the last MFMA product; [24] targets solved by the spectral
 * route, [25] ms of their decompositions, [26] their nominal flops, [27]
 * the algorithmic bytes of their tridiagonal reductions (4 n^3 / 3 each).
 * pxh_solve_group reports on ctxs[0]. */
int pxh_get_stats(pxh_ctx *ctx, double *stats, int n);

/* ------------------------------------------------------------------ serving
 * A request loop on one device (no counterpart in the reference, which
 * solves one structure per process; admission by HBM footprint is the check
 * of memory_manager_module.f90:394-397 applied per request).  A request is
 * one pxh_solve_spectrum on a context the caller has prepared with
 * pxh_set_atoms / pxh_set_mesh (+ alpha / permittivity tables).  Requests are
 * admitted in arrival order as their footprint fits the budget and their
 * context is idle (a smaller later request may pass a larger waiting one);
 * up to max_inflight run at once, each on its own context's stream, so their
 * work overlaps on the device.  A context must not be used by the caller
 * while a request on it is outstanding.  A result is identical alone and in
 * the mix. */
typedef struct pxh_server pxh_server;

/* hbm_budget_bytes = 0: 90 % of the HBM free at creation. */
int pxh_server_create(int device, size_t hbm_budget_bytes, int max_inflight, pxh_server **out);

/* Queue a solve (arguments as pxh_solve_spectrum; outputs come from
 * pxh_server_wait).  The frequency list is copied. */
int pxh_server_submit(pxh_server *srv, pxh_ctx *ctx, int n_freq, const double *freq_au, double e0, int rhs_form,
                      int pol_mask, int method, double tol_abs, int krylov_dim, int precision, int64_t *request_id);

/* Block until the request is done, copy its outputs (each may be NULL;
 * layouts as pxh_solve_spectrum) and forget it.  times_ms[0] = queueing
 * time, times_ms[1] = submission to completion.  Returns the solve's status. */
int pxh_server_wait(pxh_server *srv, int64_t request_id, double *polar, double *results, int32_t *iters,
                    double *true_residual, double *times_ms);

int pxh_server_last_error(pxh_server *srv, char *buf, int len);

/* Waits for every submitted request, then stops the loop. */
int pxh_server_destroy(pxh_server *srv);

#ifdef __cplusplus
}
#endif
#endif
