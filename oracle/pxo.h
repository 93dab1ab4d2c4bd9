/*
 * pxo.h -- FP64 CPU oracle for the plasmonX frequency-response path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product path (libpxh, plasmonx_amd) never calls it.
 *
 * It restates, in plain C99 + OpenMP, the arithmetic of the reference's
 * frequency-dependent response solve.  Every function cites the reference
 * file:line it follows.  Conventions:
 *   - atomic units, coordinates in bohr;
 *   - all matrices column-major (Fortran order), complex = double _Complex;
 *   - atom types are 0-based here (the reference's map_atomtypes is 1-based);
 *   - per-type-pair tables (fermi_d, fermi_s, r0) are nt*nt, element
 *     [ti + nt*tj] (column-major like the reference's fermi_d(ti,tj)).
 *
 * What pins each function is listed in the header block of pxo_atoms.c,
 * pxo_linalg.c, pxo_bem.c and pxo_static.c, and in the tests under tests/.
 */
#ifndef PXO_H
#define PXO_H

#include <complex.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex pxo_cplx;

enum { PXO_WFQ = 1, PXO_WFQFMU = 2 };
enum { PXO_GAUSSIAN = 1, PXO_COULOMB = 2, PXO_OHNO = 3 };
enum { PXO_RHS_FIELD = 1, PXO_RHS_POTENTIAL = 2 };

typedef struct {
    int model;           /* PXO_WFQ / PXO_WFQFMU */
    int kernel;          /* PXO_GAUSSIAN / PXO_COULOMB / PXO_OHNO */
    int n_atoms;
    int n_types;
    int heterogeneous;   /* 1 when more than one metal type (alloys) */
    const double *coord; /* 3*n, bohr */
    const int *type;     /* n, 0-based */
    const double *eta, *r_q, *r_mu, *a_ij, *tau, *density; /* per type */
    const double *fermi_d, *fermi_s, *r0;                  /* nt*nt */
    const int *nb_counts; /* nt*n: neighbours of each type (self included) */
    const int *nb_total;  /* n: all neighbours (self included) */
} pxo_atoms;

/* ---- atomistic kernels (pxo_atoms.c) ---- */
double pxo_fermi(const pxo_atoms *a, int i, int j, double r);
void pxo_kbar(const pxo_atoms *a, double *K);
void pxo_hbar(const pxo_atoms *a, double freq, pxo_cplx *H);
void pxo_t_qq(const pxo_atoms *a, double *T);
void pxo_t_qmu(const pxo_atoms *a, double *T);
void pxo_t_muq(const pxo_atoms *a, double *T);
void pxo_t_mumu(const pxo_atoms *a, double *T);
int pxo_n_var(const pxo_atoms *a);
void pxo_constant_matrix(const pxo_atoms *a, double *M);
void pxo_q_shift(const pxo_atoms *a, double freq, pxo_cplx *zq);
void pxo_mu_shift(const pxo_atoms *a, const pxo_cplx *alpha_t, pxo_cplx *lam);
void pxo_dynamic_matrix(const pxo_atoms *a, const double *M, double freq,
                        const pxo_cplx *alpha_t, pxo_cplx *A);
void pxo_static_rhs_q(const pxo_atoms *a, double e0, int rhs_form, double *V);
void pxo_rhs(const pxo_atoms *a, double e0, int rhs_form, double freq,
             int on_the_fly, pxo_cplx *rhs);
void pxo_matvec_otf(const pxo_atoms *a, double freq, const pxo_cplx *alpha_t,
                    const pxo_cplx *x, pxo_cplx *y);
void pxo_matvec_otf_rows(const pxo_atoms *a, double freq,
                         const pxo_cplx *alpha_t, const pxo_cplx *x,
                         int n_rows, const int *rows, pxo_cplx *y_rows);
void pxo_polar(const pxo_atoms *a, double e0, const pxo_cplx *var,
               pxo_cplx *polar);
void pxo_results_row(double freq, const pxo_cplx *polar, double *res20);
void pxo_induced_field(const pxo_atoms *a, int n_points, const double *points,
                       const pxo_cplx *var, int i_pol, double e0, pxo_cplx *E);
void pxo_plasmon_density(const pxo_atoms *a, int n_points, const double *points,
                         const pxo_cplx *var, pxo_cplx *rho, pxo_cplx *rho_q,
                         pxo_cplx *rho_mu);
void pxo_neighbours(int n, int nt, const double *coord, const int *type,
                    const double *r0, int *nb_counts, int *nb_total);

/* ---- dense linear algebra and GMRES (pxo_linalg.c) ---- */
int pxo_zgesv(int n, int nrhs, pxo_cplx *A, pxo_cplx *B);
typedef void (*pxo_apply_fn)(void *ctx, const pxo_cplx *x, pxo_cplx *y);
int pxo_gmres(int n, pxo_apply_fn apply, void *ctx, const pxo_cplx *b,
              pxo_cplx *x, double tol, int m, int *iters, double *hist,
              double *true_res);
int pxo_gmres_dense(int n, const pxo_cplx *A, const pxo_cplx *b, pxo_cplx *x,
                    double tol, int m, int *iters, double *hist,
                    double *true_res);
int pxo_gmres_otf(const pxo_atoms *a, double freq, const pxo_cplx *alpha_t,
                  const pxo_cplx *b, pxo_cplx *x, double tol, int m,
                  int *iters, double *hist, double *true_res);

/* ---- spectra driver (pxo_atoms.c) ---- */
int pxo_solve_spectrum(const pxo_atoms *a, int n_freq, const double *freq,
                       const pxo_cplx *alpha_w /* nt*n_freq */, double e0,
                       int rhs_form, int pol_mask, int method, double tol,
                       int m, pxo_cplx *variables /* n_var*3*n_freq or NULL */,
                       pxo_cplx *polar /* 9*n_freq */,
                       double *results /* n_freq*20, column-major */,
                       int *iters /* 3*n_freq */);

/* ---- static FQ / FQFmu (pxo_static.c); model PXO_WFQ = charges only (FQ),
 * PXO_WFQFMU = charges and dipoles (FQFmu); what 0 = ground state (energy),
 * 1 = static field; mol_count[n_mol]: atoms per molecule, contiguous runs ---- */
int pxo_static_n_var(const pxo_atoms *a, int n_mol);
void pxo_static_matrix(const pxo_atoms *a, const double *alpha, int n_mol, const int *mol_count, double *A);
void pxo_static_rhs(const pxo_atoms *a, const double *chi, int n_mol, int what, double e0, int rhs_form,
                    double *rhs);
int pxo_static_solve(const pxo_atoms *a, const double *chi, const double *alpha, int n_mol, const int *mol_count,
                     int what, double e0, int rhs_form, double *variables, double *polar, double *energy);

/* ---- permittivity models (pxo_atoms.c) ---- */
void pxo_silver_etchegoin(int polar, int n, const double *freq, pxo_cplx *out);
void pxo_gold_etchegoin(int polar, int n, const double *freq, pxo_cplx *out);

/* ---- BEM (pxo_bem.c) ---- */
typedef struct {
    int n;                 /* tesserae */
    const double *coord;   /* 3*n centroids, bohr */
    const double *normal;  /* 3*n */
    const double *tang1;   /* 3*n */
    const double *tang2;   /* 3*n */
    const double *area;    /* n, bohr^2 */
    const double *verts;   /* 9*n: the 3 vertices of each tessera, bohr */
} pxo_mesh;

void pxo_tesserae(int n_tess, const double *verts9, double normal_factor,
                  double *coord, double *normal, double *tang1, double *tang2,
                  double *area);
void pxo_lgl(int npol, double *points, double *weights);
void pxo_bem_induced_field(int n_tess, const double *c, const pxo_cplx *q,
                           int n_points, const double *points, int i_pol,
                           double e0, pxo_cplx *E);
void pxo_bem_d(const pxo_mesh *m, int accurate, double sphere_r, double *D,
               double *D_ief /* may be NULL: D with unrefined diagonal - 2pi */,
               double *SA /* may be NULL: IEF S A^-1 (real) */);
void pxo_bem_rhs(const pxo_mesh *m, int ief, const double *D_ief, double e0,
                 pxo_cplx *rhs);
void pxo_bem_matrix(const pxo_mesh *m, int ief, const double *D,
                    const double *SA, pxo_cplx lambda, pxo_cplx *A);
pxo_cplx pxo_bem_lambda(double eps_solvent, pxo_cplx eps);
void pxo_bem_polar(const pxo_mesh *m, int ief, double e0, const pxo_cplx *var,
                   pxo_cplx *polar);
int pxo_bem_solve_spectrum(const pxo_mesh *m, int ief, int accurate,
                           double sphere_r, double eps_solvent, int n_freq,
                           const double *freq, const pxo_cplx *eps_w,
                           double e0, int charge_constraint, int method,
                           double tol, int mdim, pxo_cplx *variables,
                           pxo_cplx *polar, double *results, int *iters);

#ifdef __cplusplus
}
#endif
#endif
