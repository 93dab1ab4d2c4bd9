/*
 * pxo_static.c -- FP64 CPU oracle: the static FQ / FQFmu models ('what:
 * energy' and 'what: static response').
 *
 * TEST INFRASTRUCTURE ONLY (see pxo.h).  Not linked into the product.
 *
 * Restates:
 *   - the matrix of src/forcefield/fq_module.f90:179-207 and
 *     fqfmu_module.f90:81-140: unknowns [q_1..q_N, lambda_1..lambda_M,
 *     mu_1x, mu_1y, mu_1z, ...], blocks T_qq (matrix_module.f90:58-115),
 *     the Lagrangian columns of ones of each molecule
 *     (add_lagrangian_blocks, matrix_module.f90:351-382, molecules as
 *     contiguous atom runs), T_qmu and T_muq = T_qmu^T (:120-234), T_mumu
 *     with 1/alpha on its diagonal for the static models (:240-345);
 *   - the right-hand sides: ground state -chi on the charge rows
 *     (fq_module.f90:211-240, fqfmu_module.f90:142-168); static field the
 *     potential of rhs_module.f90:43-90 on the charge rows and, on dipole
 *     component k of every atom, 1 when e0 is not zero and 0 when it is
 *     (rhs_module.f90:96-127, as written there);
 *   - the solve of algorithm_module.f90:469-504 (DSYSV there; here the
 *     textbook LU of pxo_zgesv on the same system);
 *   - the energy -1/2 x^T b (fq_module.f90:364-381, fqfmu_module.f90:320-336);
 *   - the FQ static polarizability of fq_module.f90:385-597:
 *     alpha = R J^-1 R^T - (L^T J^-1 R)^T (L^T J^-1 L)^-1 (L^T J^-1 R) with
 *     J = T_qq, R the atomic coordinates, by an explicit J^-1 as there.
 *
 * Pinned by the reference's fq_energy, fq_polarizability, fqfmu_energy and
 * fqfmu_rq_rmu_energy logs (matrix E11.4, J^-1 E11.4, charges and dipoles
 * E14.6, energy F14.8, polarizability tensor E11.4 and F19.10), and by the
 * closed form alpha = -R^T q(E0 = 1) of the constrained solve.
 */
#include "pxo.h"

#include <stdlib.h>
#include <string.h>

int pxo_static_n_var(const pxo_atoms *a, int n_mol) {
    return a->n_atoms + n_mol + (a->model == PXO_WFQFMU ? 3 * a->n_atoms : 0);
}

void pxo_static_matrix(const pxo_atoms *a, const double *alpha, int n_mol, const int *mol_count, double *A) {
    const int n = a->n_atoms, nv = pxo_static_n_var(a, n_mol), m0 = n + n_mol;
    memset(A, 0, sizeof(double) * (size_t)nv * nv);
    double *T = malloc(sizeof(double) * (size_t)n * n);
    pxo_t_qq(a, T);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) A[i + (size_t)nv * j] = T[i + (size_t)n * j];
    free(T);
    int first = 0;
    for (int k = 0; k < n_mol; ++k) {
        for (int i = first; i < first + mol_count[k]; ++i) {
            A[i + (size_t)nv * (n + k)] = 1.0;
            A[(n + k) + (size_t)nv * i] = 1.0;
        }
        first += mol_count[k];
    }
    if (a->model != PXO_WFQFMU) return;
    const int nm = 3 * n;
    double *Tq = malloc(sizeof(double) * (size_t)n * nm);
    pxo_t_qmu(a, Tq);
    for (int j = 0; j < nm; ++j)
        for (int i = 0; i < n; ++i) {
            A[i + (size_t)nv * (m0 + j)] = Tq[i + (size_t)n * j];
            A[(m0 + j) + (size_t)nv * i] = Tq[i + (size_t)n * j];
        }
    free(Tq);
    double *Tm = malloc(sizeof(double) * (size_t)nm * nm);
    pxo_t_mumu(a, Tm);
    for (int j = 0; j < nm; ++j)
        for (int i = 0; i < nm; ++i) A[(m0 + i) + (size_t)nv * (m0 + j)] = Tm[i + (size_t)nm * j];
    free(Tm);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) A[(m0 + 3 * i + k) * (size_t)(nv + 1)] = 1.0 / alpha[a->type[i]];
}

void pxo_static_rhs(const pxo_atoms *a, const double *chi, int n_mol, int what, double e0, int rhs_form,
                    double *rhs) {
    const int n = a->n_atoms, nv = pxo_static_n_var(a, n_mol), m0 = n + n_mol;
    const int ncol = what == 0 ? 1 : 3;
    memset(rhs, 0, sizeof(double) * (size_t)nv * ncol);
    if (what == 0) {
        for (int i = 0; i < n; ++i) rhs[i] = -chi[a->type[i]];
        return;
    }
    double *V = malloc(sizeof(double) * (size_t)n * 3);
    pxo_static_rhs_q(a, e0, rhs_form, V);
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < n; ++i) rhs[i + (size_t)nv * k] = V[i + (size_t)n * k];
    free(V);
    if (a->model == PXO_WFQFMU)
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) rhs[(m0 + 3 * i + k) + (size_t)nv * k] = e0 == 0.0 ? 0.0 : 1.0;
}

/* Solve A X = B for real A (nv x nv) and ncol columns through pxo_zgesv. */
static int real_solve(int nv, const double *A, int ncol, const double *B, double *X) {
    pxo_cplx *Ac = malloc(sizeof(pxo_cplx) * (size_t)nv * nv);
    pxo_cplx *Bc = malloc(sizeof(pxo_cplx) * (size_t)nv * ncol);
    for (size_t e = 0; e < (size_t)nv * nv; ++e) Ac[e] = A[e];
    for (size_t e = 0; e < (size_t)nv * ncol; ++e) Bc[e] = B[e];
    const int info = pxo_zgesv(nv, ncol, Ac, Bc);
    for (size_t e = 0; e < (size_t)nv * ncol; ++e) X[e] = creal(Bc[e]);
    free(Ac);
    free(Bc);
    return info;
}

int pxo_static_solve(const pxo_atoms *a, const double *chi, const double *alpha, int n_mol, const int *mol_count,
                     int what, double e0, int rhs_form, double *variables, double *polar, double *energy) {
    const int n = a->n_atoms, nv = pxo_static_n_var(a, n_mol);
    const int ncol = what == 0 ? 1 : 3;
    double *A = malloc(sizeof(double) * (size_t)nv * nv);
    double *B = malloc(sizeof(double) * (size_t)nv * ncol);
    double *X = malloc(sizeof(double) * (size_t)nv * ncol);
    pxo_static_matrix(a, alpha, n_mol, mol_count, A);
    pxo_static_rhs(a, chi, n_mol, what, e0, rhs_form, B);
    int status = real_solve(nv, A, ncol, B, X) != 0 ? 3 : 0;
    if (variables) memcpy(variables, X, sizeof(double) * (size_t)nv * ncol);
    if (energy) {
        double e = 0.0;
        for (int i = 0; i < nv; ++i) e += X[i] * B[i];
        *energy = -0.5 * e;
    }
    if (polar && what == 1 && a->model != PXO_WFQFMU && status == 0) {
        /* J^-1 explicitly (J = T_qq), then the two terms of the formula */
        double *J = malloc(sizeof(double) * (size_t)n * n);
        double *Id = calloc((size_t)n * n, sizeof(double));
        double *Ji = malloc(sizeof(double) * (size_t)n * n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) J[i + (size_t)n * j] = A[i + (size_t)nv * j];
        for (int i = 0; i < n; ++i) Id[i * (size_t)(n + 1)] = 1.0;
        if (real_solve(n, J, n, Id, Ji) != 0) status = 3;
        /* JR = J^-1 R^T (n x 3), JL = J^-1 L (n x M) */
        double *JR = calloc((size_t)n * 3, sizeof(double));
        double *JL = calloc((size_t)n * (n_mol > 0 ? n_mol : 1), sizeof(double));
        for (int b = 0; b < 3; ++b)
            for (int j = 0; j < n; ++j) {
                const double r = a->coord[3 * j + b];
                for (int i = 0; i < n; ++i) JR[i + (size_t)n * b] += Ji[i + (size_t)n * j] * r;
            }
        int first = 0;
        for (int k = 0; k < n_mol; ++k) {
            for (int j = first; j < first + mol_count[k]; ++j)
                for (int i = 0; i < n; ++i) JL[i + (size_t)n * k] += Ji[i + (size_t)n * j];
            first += mol_count[k];
        }
        double P[9];
        for (int aa = 0; aa < 3; ++aa)
            for (int b = 0; b < 3; ++b) {
                double s = 0.0;
                for (int i = 0; i < n; ++i) s += a->coord[3 * i + aa] * JR[i + (size_t)n * b];
                P[aa + 3 * b] = s;
            }
        if (n_mol > 0) {
            /* C = L^T J^-1 R (M x 3), G = L^T J^-1 L (M x M), polar -= C^T G^-1 C */
            double *C = calloc((size_t)n_mol * 3, sizeof(double));
            double *G = calloc((size_t)n_mol * n_mol, sizeof(double));
            double *GC = malloc(sizeof(double) * (size_t)n_mol * 3);
            first = 0;
            for (int k = 0; k < n_mol; ++k) {
                for (int i = first; i < first + mol_count[k]; ++i) {
                    for (int b = 0; b < 3; ++b) C[k + (size_t)n_mol * b] += JR[i + (size_t)n * b];
                    for (int l = 0; l < n_mol; ++l) G[k + (size_t)n_mol * l] += JL[i + (size_t)n * l];
                }
                first += mol_count[k];
            }
            if (real_solve(n_mol, G, 3, C, GC) != 0) status = 3;
            for (int aa = 0; aa < 3; ++aa)
                for (int b = 0; b < 3; ++b) {
                    double s = 0.0;
                    for (int k = 0; k < n_mol; ++k) s += C[k + (size_t)n_mol * aa] * GC[k + (size_t)n_mol * b];
                    P[aa + 3 * b] -= s;
                }
            free(C);
            free(G);
            free(GC);
        }
        memcpy(polar, P, sizeof P);
        free(J);
        free(Id);
        free(Ji);
        free(JR);
        free(JL);
    }
    free(A);
    free(B);
    free(X);
    return status;
}
