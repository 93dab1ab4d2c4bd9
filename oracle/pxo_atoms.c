/*
 * pxo_atoms.c -- FP64 CPU oracle: omegaFQ / omegaFQFmu operators, RHS,
 * polarizability, cross sections and the per-frequency spectrum driver.
 *
 * TEST INFRASTRUCTURE ONLY (see pxo.h).  Not linked into the product.
 *
 * What pins these functions (tests/test_oracle_*.py):
 *   - the reference's regression logs, parsed into the JSON fixtures under tests/golden:
 *     printed constant / dynamic matrices (E11.4), RHS (E10.3), charges and
 *     dipoles (E10.3), polarizability tensors and the 20 result columns
 *     (E14.6) of wfq_* and wfqfmu_* cases;
 *   - invariants: Fermi f(s*r0) = 1/2, K rows sum to zero, T symmetric,
 *     sum of charges = 0 for every solution, T -> point-charge / point-dipole
 *     limits for r >> R, on-the-fly matvec == assembled matrix times x;
 *   - cross sections recomputed from the polarizability (closed form);
 *   - induced field and plasmon density at points: the cubes (E14.6), plane
 *     CSVs (E25.16), .plt grids and effective volumes the reference's
 *     analysis tests printed; the point-dipole limit of the field far from a
 *     +-q pair; the density's zeroth and first moments (q, q r_i + mu_i).
 */
#include "pxo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.14159265358979323846
#define TOBOHR 1.8897261254578281
#define AU_TO_EV 27.211386245981
#define AU_TO_CM 219474.63136320
#define LIGHT 137.035227

static inline int tt(const pxo_atoms *a, int ti, int tj) { return ti + a->n_types * tj; }

static inline double dist3(const double *c, int i, int j, double *d) {
    d[0] = c[3 * i + 0] - c[3 * j + 0];
    d[1] = c[3 * i + 1] - c[3 * j + 1];
    d[2] = c[3 * i + 2] - c[3 * j + 2];
    return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

int pxo_n_var(const pxo_atoms *a) {
    return a->model == PXO_WFQFMU ? 4 * a->n_atoms : a->n_atoms;
}

/* Fermi damping between atoms i and j at distance r.
 * src/matrix/matrix_module.f90:694-715 (fermi_function); r0 per element pair
 * from src/target/target_module.f90:472-518 (nearest_distance). */
double pxo_fermi(const pxo_atoms *a, int i, int j, double r) {
    int k = tt(a, a->type[i], a->type[j]);
    double d = a->fermi_d[k], s = a->fermi_s[k], r0 = a->r0[k];
    return 1.0 / (1.0 + exp(-d * ((r / (s * r0)) - 1.0)));
}

/* K-bar = K - P.  src/matrix/matrix_module.f90:393-426. */
void pxo_kbar(const pxo_atoms *a, double *K) {
    int n = a->n_atoms;
    memset(K, 0, sizeof(double) * (size_t)n * n);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            double d[3];
            double r = dist3(a->coord, i, j, d);
            K[i + (size_t)n * j] = (1.0 - pxo_fermi(a, i, j, r)) * a->a_ij[a->type[i]] / r;
            K[j + (size_t)n * i] = (1.0 - pxo_fermi(a, j, i, r)) * a->a_ij[a->type[j]] / r;
        }
    }
#pragma omp parallel for
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += K[i + (size_t)n * j];
        K[i + (size_t)n * i] = -s;
    }
}

/* Drude weight w_t(omega) = 2 n_t / (1/tau_t - i omega).
 * src/matrix/matrix_module.f90:455-466. */
static pxo_cplx drude_w(const pxo_atoms *a, int t, double freq) {
    return 2.0 * a->density[t] / (1.0 / a->tau[t] - I * freq);
}

/* H-bar(omega) = H - L (alloys).  src/matrix/matrix_module.f90:435-482. */
void pxo_hbar(const pxo_atoms *a, double freq, pxo_cplx *H) {
    int n = a->n_atoms;
    memset(H, 0, sizeof(pxo_cplx) * (size_t)n * n);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        pxo_cplx wi = drude_w(a, a->type[i], freq);
        for (int j = i + 1; j < n; ++j) {
            pxo_cplx wj = drude_w(a, a->type[j], freq);
            double d[3];
            double r = dist3(a->coord, i, j, d);
            H[i + (size_t)n * j] = (wj / wi) * (1.0 - pxo_fermi(a, i, j, r)) * a->a_ij[a->type[j]] / r;
            H[j + (size_t)n * i] = (wi / wj) * (1.0 - pxo_fermi(a, j, i, r)) * a->a_ij[a->type[i]] / r;
        }
    }
#pragma omp parallel for
    for (int i = 0; i < n; ++i) {
        pxo_cplx s = 0.0;
        for (int j = 0; j < n; ++j) s += H[i + (size_t)n * j];
        H[i + (size_t)n * i] = -s;
    }
}

/* Charge-charge interaction element.  src/matrix/matrix_module.f90:58-115
 * (dense) and :740-820 (on the fly: the diagonal is eta_i). */
static double tqq_elem(const pxo_atoms *a, int i, int j, double r) {
    int ti = a->type[i], tj = a->type[j];
    if (a->kernel == PXO_OHNO) {
        double e = 0.5 * (a->eta[ti] + a->eta[tj]);
        return e / sqrt(1.0 + e * e * r * r);
    }
    if (i == j) return 0.5 * (a->eta[ti] + a->eta[tj]);
    if (a->kernel == PXO_COULOMB) return 1.0 / r;
    double R = sqrt(a->r_q[ti] * a->r_q[ti] + a->r_q[tj] * a->r_q[tj]);
    return erf(r / R) / r;
}

/* Shared radial factor of the damped charge-dipole and dipole-dipole
 * kernels: erf(r/R) - 2r/(sqrt(pi) R) exp(-r^2/R^2).
 * src/matrix/matrix_module.f90:155-163, 297-300. */
static inline double fd1(double r, double R) {
    return erf(r / R) - ((2.0 * r) / (sqrt(PI) * R)) * exp(-(r * r) / (R * R));
}

/* T_qmu(i, 3j+k) = (r_i - r_j)_k / r^3 * fd1(r, sqrt(Rq_i^2 + Rmu_j^2)).
 * src/matrix/matrix_module.f90:120-173. */
static void tqmu_vec(const pxo_atoms *a, int i, int j, const double *d, double r, double *v) {
    int ti = a->type[i], tj = a->type[j];
    double R = sqrt(a->r_q[ti] * a->r_q[ti] + a->r_mu[tj] * a->r_mu[tj]);
    double f = fd1(r, R) / (r * r * r);
    v[0] = d[0] * f;
    v[1] = d[1] * f;
    v[2] = d[2] * f;
}

/* Dipole-dipole 3x3 block (symmetric), row-major t[3*a+b].
 * src/matrix/matrix_module.f90:290-318. */
static void tmumu_block(const pxo_atoms *a, int i, int j, const double *d, double r, double *t) {
    int ti = a->type[i], tj = a->type[j];
    double R = sqrt(a->r_mu[ti] * a->r_mu[ti] + a->r_mu[tj] * a->r_mu[tj]);
    double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
    double e = exp(-r2 / (R * R));
    double f1 = erf(r / R) - ((2.0 * r) / (sqrt(PI) * R)) * e;
    double f2 = (4.0 / (sqrt(PI) * R * R * R)) * e;
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q)
            t[3 * p + q] = ((-3.0 * d[p] * d[q]) / r5 + (p == q ? 1.0 / r3 : 0.0)) * f1 +
                           f2 * (d[p] * d[q]) / r2;
}

void pxo_t_qq(const pxo_atoms *a, double *T) {
    int n = a->n_atoms;
#pragma omp parallel for
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double d[3];
            double r = dist3(a->coord, i, j, d);
            T[i + (size_t)n * j] = tqq_elem(a, i, j, r);
        }
}

/* T_qmu: n x 3n.  src/matrix/matrix_module.f90:120-173. */
void pxo_t_qmu(const pxo_atoms *a, double *T) {
    int n = a->n_atoms;
    memset(T, 0, sizeof(double) * (size_t)n * 3 * n);
#pragma omp parallel for
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            double d[3], v[3];
            double r = dist3(a->coord, i, j, d);
            tqmu_vec(a, i, j, d, r, v);
            for (int k = 0; k < 3; ++k) T[i + (size_t)n * (3 * j + k)] = v[k];
        }
}

/* T_muq: 3n x n, T_muq(3j+k, i) = T_qmu(i, 3j+k).
 * src/matrix/matrix_module.f90:179-234. */
void pxo_t_muq(const pxo_atoms *a, double *T) {
    int n = a->n_atoms, nm = 3 * n;
    memset(T, 0, sizeof(double) * (size_t)n * nm);
#pragma omp parallel for
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            double d[3], v[3];
            double r = dist3(a->coord, i, j, d);
            tqmu_vec(a, i, j, d, r, v);
            for (int k = 0; k < 3; ++k) T[(3 * j + k) + (size_t)nm * i] = v[k];
        }
}

/* T_mumu: 3n x 3n, zero diagonal blocks for the w-models.
 * src/matrix/matrix_module.f90:240-345. */
void pxo_t_mumu(const pxo_atoms *a, double *T) {
    int n = a->n_atoms, nm = 3 * n;
    memset(T, 0, sizeof(double) * (size_t)nm * nm);
#pragma omp parallel for
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            double d[3], t[9];
            double r = dist3(a->coord, i, j, d);
            tmumu_block(a, i, j, d, r, t);
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) T[(3 * i + p) + (size_t)nm * (3 * j + q)] = t[3 * p + q];
        }
}

/* C = A * B, A m x k, B k x n, all column-major (own blocked product). */
static void dgemm_nn(int m, int n, int k, const double *A, int lda, const double *B, int ldb,
                     double *C, int ldc) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j) {
        double *c = C + (size_t)ldc * j;
        for (int i = 0; i < m; ++i) c[i] = 0.0;
        for (int p = 0; p < k; ++p) {
            double b = B[p + (size_t)ldb * j];
            if (b == 0.0) continue;
            const double *ap = A + (size_t)lda * p;
            for (int i = 0; i < m; ++i) c[i] += ap[i] * b;
        }
    }
}

/* Constant (frequency-independent) matrix of the homogeneous models:
 *   wfq:    M = Kbar * T_qq                   (wfq_module.f90:192-206)
 *   wfqfmu: M = [Kbar T_qq, Kbar T_muq^T; T_muq, T_mumu]
 *                                             (wfqfmu_module.f90:275-323)
 * For alloys the q rows depend on omega and are left at zero here, as in the
 * reference (they are built per frequency by pxo_dynamic_matrix). */
void pxo_constant_matrix(const pxo_atoms *a, double *M) {
    int n = a->n_atoms, nv = pxo_n_var(a);
    memset(M, 0, sizeof(double) * (size_t)nv * nv);
    double *K = malloc(sizeof(double) * (size_t)n * n);
    pxo_kbar(a, K);
    if (!a->heterogeneous) {
        double *T = malloc(sizeof(double) * (size_t)n * n);
        double *A = malloc(sizeof(double) * (size_t)n * n);
        pxo_t_qq(a, T);
        dgemm_nn(n, n, n, K, n, T, n, A, n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) M[i + (size_t)nv * j] = A[i + (size_t)n * j];
        free(T);
        free(A);
    }
    if (a->model == PXO_WFQFMU) {
        int nm = 3 * n;
        double *Tmq = malloc(sizeof(double) * (size_t)nm * n);
        pxo_t_muq(a, Tmq);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < nm; ++i) M[(n + i) + (size_t)nv * j] = Tmq[i + (size_t)nm * j];
        if (!a->heterogeneous) {
            /* Kbar * T_muq^T: T_muq^T is n x 3n with element (i, c) = Tmq(c, i) */
            double *Tqm = malloc(sizeof(double) * (size_t)n * nm);
            for (int c = 0; c < nm; ++c)
                for (int i = 0; i < n; ++i) Tqm[i + (size_t)n * c] = Tmq[c + (size_t)nm * i];
            double *A = malloc(sizeof(double) * (size_t)n * nm);
            dgemm_nn(n, nm, n, K, n, Tqm, n, A, n);
            for (int c = 0; c < nm; ++c)
                for (int i = 0; i < n; ++i) M[i + (size_t)nv * (n + c)] = A[i + (size_t)n * c];
            free(Tqm);
            free(A);
        }
        free(Tmq);
        double *Tmm = malloc(sizeof(double) * (size_t)nm * nm);
        pxo_t_mumu(a, Tmm);
        for (int j = 0; j < nm; ++j)
            for (int i = 0; i < nm; ++i) M[(n + i) + (size_t)nv * (n + j)] = Tmm[i + (size_t)nm * j];
        free(Tmm);
    }
    free(K);
}

/* z_q(omega) = -(omega / (2 n0 tau)) (omega tau + i), per atom; the
 * q-diagonal of A(omega) is M_ii - z_q.  wfq_module.f90:239-242. */
void pxo_q_shift(const pxo_atoms *a, double freq, pxo_cplx *zq) {
    for (int i = 0; i < a->n_atoms; ++i) {
        int t = a->type[i];
        zq[i] = -(freq / (2.0 * a->density[t] * a->tau[t])) * (freq * a->tau[t] + I);
    }
}

/* Lambda_i(omega) = sum_t (nb_t,i / nb_i) / alpha_t(omega): the mu-diagonal.
 * wfqfmu_module.f90:516-529 and :729-745. */
void pxo_mu_shift(const pxo_atoms *a, const pxo_cplx *alpha_t, pxo_cplx *lam) {
    for (int i = 0; i < a->n_atoms; ++i) {
        pxo_cplx z = 0.0;
        for (int t = 0; t < a->n_types; ++t)
            z += ((double)a->nb_counts[t + a->n_types * i] / (double)a->nb_total[i]) / alpha_t[t];
        lam[i] = z;
    }
}

/* A(omega).  wfq_module.f90:214-254; wfqfmu_module.f90:436-538 (including
 * the alloy q rows 1/2 (Kbar + Hbar(omega)) [T_qq, T_muq^T]). */
void pxo_dynamic_matrix(const pxo_atoms *a, const double *M, double freq, const pxo_cplx *alpha_t,
                        pxo_cplx *A) {
    int n = a->n_atoms, nv = pxo_n_var(a);
    for (size_t e = 0; e < (size_t)nv * nv; ++e) A[e] = M[e];
    if (a->heterogeneous) {
        double *K = malloc(sizeof(double) * (size_t)n * n);
        pxo_cplx *H = malloc(sizeof(pxo_cplx) * (size_t)n * n);
        pxo_kbar(a, K);
        pxo_hbar(a, freq, H);
        for (size_t e = 0; e < (size_t)n * n; ++e) H[e] += K[e];
        int nc = a->model == PXO_WFQFMU ? 4 * n : n;
        /* columns: T_qq then (for wfqfmu) T_muq^T */
        double *Tc = malloc(sizeof(double) * (size_t)n * nc);
        double *T = malloc(sizeof(double) * (size_t)n * n);
        pxo_t_qq(a, T);
        memcpy(Tc, T, sizeof(double) * (size_t)n * n);
        free(T);
        if (a->model == PXO_WFQFMU) {
            double *Tmq = malloc(sizeof(double) * (size_t)3 * n * n);
            pxo_t_muq(a, Tmq);
            for (int c = 0; c < 3 * n; ++c)
                for (int i = 0; i < n; ++i) Tc[i + (size_t)n * (n + c)] = Tmq[c + (size_t)3 * n * i];
            free(Tmq);
        }
#pragma omp parallel for
        for (int c = 0; c < nc; ++c) {
            for (int i = 0; i < n; ++i) {
                pxo_cplx s = 0.0;
                for (int p = 0; p < n; ++p) s += H[i + (size_t)n * p] * Tc[p + (size_t)n * c];
                A[i + (size_t)nv * c] = 0.5 * s;
            }
        }
        free(Tc);
        free(H);
        free(K);
    }
    pxo_cplx *zq = malloc(sizeof(pxo_cplx) * n);
    pxo_q_shift(a, freq, zq);
    for (int i = 0; i < n; ++i) A[i + (size_t)nv * i] -= zq[i];
    free(zq);
    if (a->model == PXO_WFQFMU) {
        pxo_cplx *lam = malloc(sizeof(pxo_cplx) * n);
        pxo_mu_shift(a, alpha_t, lam);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) {
                size_t r = n + 3 * i + k;
                A[r + (size_t)nv * r] = lam[i];
            }
        free(lam);
    }
}

/* Static potential of the external field, V(i,k) = -E0 x_k(i) ('field') or
 * -E0 |x_k(i) - min_k| ('potential').  src/rhs/rhs_module.f90:43-90. */
void pxo_static_rhs_q(const pxo_atoms *a, double e0, int rhs_form, double *V) {
    int n = a->n_atoms;
    double mn[3] = {0, 0, 0};
    if (rhs_form == PXO_RHS_POTENTIAL) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = a->coord[k];
            for (int i = 0; i < n; ++i)
                if (a->coord[3 * i + k] < mn[k]) mn[k] = a->coord[3 * i + k];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            double x = a->coord[3 * i + k];
            V[i + (size_t)n * k] = rhs_form == PXO_RHS_POTENTIAL ? -e0 * fabs(x - mn[k]) : -e0 * x;
        }
}

/* Dynamic RHS, all three field directions (n_var x 3):
 *   q rows: -Kbar V (rhs_module.f90:140-186 dense, :195-250 on the fly) or,
 *           for alloys, -1/2 (Kbar + Hbar(omega)) V (:261-305, :316-390);
 *   mu rows: E0 on component k of every atom (rhs_module.f90:398-430). */
void pxo_rhs(const pxo_atoms *a, double e0, int rhs_form, double freq, int on_the_fly,
             pxo_cplx *rhs) {
    int n = a->n_atoms, nv = pxo_n_var(a);
    double *V = malloc(sizeof(double) * (size_t)n * 3);
    pxo_static_rhs_q(a, e0, rhs_form, V);
    memset(rhs, 0, sizeof(pxo_cplx) * (size_t)nv * 3);
    (void)on_the_fly; /* the dense and pair-sweep forms are the same sums */
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        pxo_cplx acc[3] = {0, 0, 0};
        int ti = a->type[i];
        pxo_cplx wi = drude_w(a, ti, freq);
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double d[3];
            double r = dist3(a->coord, i, j, d);
            double kij = (1.0 - pxo_fermi(a, i, j, r)) * a->a_ij[ti] / r;
            pxo_cplx c = kij;
            if (a->heterogeneous) {
                int tj = a->type[j];
                double kji = (1.0 - pxo_fermi(a, j, i, r)) * a->a_ij[tj] / r;
                pxo_cplx wj = drude_w(a, tj, freq);
                c = 0.5 * (kij + (wj / wi) * kji);
            }
            for (int k = 0; k < 3; ++k) acc[k] -= c * (V[j + (size_t)n * k] - V[i + (size_t)n * k]);
        }
        for (int k = 0; k < 3; ++k) rhs[i + (size_t)nv * k] = acc[k];
    }
    if (a->model == PXO_WFQFMU)
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) rhs[(n + 3 * i + k) + (size_t)nv * k] = e0;
    free(V);
}

/* y = A(omega) x computed on the fly:
 *   t = T x           (product_q_block_x matrix_module.f90:727-825 and
 *                      product_mu_block_x :834-1001)
 *   y_q = Kbar t_q    (update_y_q :1011-1098, 1/2(Kbar+Hbar) for alloys)
 *   y_mu = t_mu       (update_y_mu :1107-1125)
 *   then the diagonal shift (new_gmres_diagonal_shift_wfq
 *   wfq_module.f90:398-424, _wfqfmu wfqfmu_module.f90:698-746). */
void pxo_matvec_otf(const pxo_atoms *a, double freq, const pxo_cplx *alpha_t, const pxo_cplx *x,
                    pxo_cplx *y) {
    int n = a->n_atoms, nv = pxo_n_var(a);
    int fmu = a->model == PXO_WFQFMU;
    pxo_cplx *t = malloc(sizeof(pxo_cplx) * (size_t)nv);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        pxo_cplx tq = 0.0, tm[3] = {0, 0, 0};
        for (int j = 0; j < n; ++j) {
            double d[3];
            double r = dist3(a->coord, i, j, d);
            tq += tqq_elem(a, i, j, r) * x[j];
            if (!fmu || i == j) continue;
            double v[3], w[3], b[9];
            tqmu_vec(a, i, j, d, r, v); /* T_qmu(i, j) */
            double dm[3] = {-d[0], -d[1], -d[2]};
            tqmu_vec(a, j, i, dm, r, w); /* T_qmu(j, i) = T_muq(i, j) */
            tmumu_block(a, i, j, d, r, b);
            const pxo_cplx *mj = x + n + 3 * j;
            tq += v[0] * mj[0] + v[1] * mj[1] + v[2] * mj[2];
            for (int p = 0; p < 3; ++p)
                tm[p] += w[p] * x[j] + b[3 * p + 0] * mj[0] + b[3 * p + 1] * mj[1] + b[3 * p + 2] * mj[2];
        }
        t[i] = tq;
        if (fmu)
            for (int p = 0; p < 3; ++p) t[n + 3 * i + p] = tm[p];
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        int ti = a->type[i];
        pxo_cplx wi = drude_w(a, ti, freq);
        pxo_cplx acc = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double d[3];
            double r = dist3(a->coord, i, j, d);
            double kij = (1.0 - pxo_fermi(a, i, j, r)) * a->a_ij[ti] / r;
            if (a->heterogeneous) {
                int tj = a->type[j];
                double kji = (1.0 - pxo_fermi(a, j, i, r)) * a->a_ij[tj] / r;
                pxo_cplx wj = drude_w(a, tj, freq);
                acc += 0.5 * (kij + (wj / wi) * kji) * (t[j] - t[i]);
            } else {
                acc += kij * (t[j] - t[i]);
            }
        }
        y[i] = acc;
    }
    pxo_cplx *zq = malloc(sizeof(pxo_cplx) * n);
    pxo_q_shift(a, freq, zq);
    for (int i = 0; i < n; ++i) y[i] -= zq[i] * x[i];
    free(zq);
    if (fmu) {
        pxo_cplx *lam = malloc(sizeof(pxo_cplx) * n);
        pxo_mu_shift(a, alpha_t, lam);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) y[n + 3 * i + k] = t[n + 3 * i + k] + lam[i] * x[n + 3 * i + k];
        free(lam);
    }
    free(t);
}

/* alpha_ab = sum_i q_i^(b) x_a(i) / E0 (+ sum_i mu_a,i^(b) / E0 for wfqfmu).
 * wfq_module.f90:523-583; wfqfmu_module.f90:879-964.  polar is 3x3
 * column-major (polar[a + 3 b]). */
void pxo_polar(const pxo_atoms *a, double e0, const pxo_cplx *var, pxo_cplx *polar) {
    int n = a->n_atoms, nv = pxo_n_var(a);
    for (int k = 0; k < 9; ++k) polar[k] = 0.0;
    for (int b = 0; b < 3; ++b)
        for (int i = 0; i < n; ++i)
            for (int p = 0; p < 3; ++p) polar[p + 3 * b] += var[i + (size_t)nv * b] * a->coord[3 * i + p] / e0;
    if (a->model == PXO_WFQFMU)
        for (int b = 0; b < 3; ++b)
            for (int i = 0; i < n; ++i)
                for (int p = 0; p < 3; ++p) polar[p + 3 * b] += var[(n + 3 * i + p) + (size_t)nv * b] / e0;
}

/* The 20 result columns: isotropic / longitudinal polarizabilities and the
 * absorption, scattering and extinction cross sections.
 * wfq_module.f90:561-581 and target_module.f90:1989-2054. */
void pxo_results_row(double freq, const pxo_cplx *polar, double *r) {
    r[0] = (creal(polar[0]) + creal(polar[4]) + creal(polar[8])) / 3.0;
    r[1] = (cimag(polar[0]) + cimag(polar[4]) + cimag(polar[8])) / 3.0;
    r[2] = creal(polar[0]);
    r[3] = creal(polar[4]);
    r[4] = creal(polar[8]);
    r[5] = cimag(polar[0]);
    r[6] = cimag(polar[4]);
    r[7] = cimag(polar[8]);
    double m_iso = r[0] * r[0] + r[1] * r[1];
    double m_x = r[2] * r[2] + r[5] * r[5];
    double m_y = r[3] * r[3] + r[6] * r[6];
    double m_z = r[4] * r[4] + r[7] * r[7];
    r[8] = 4.0 * PI * freq * r[1] / LIGHT;
    r[9] = 4.0 * PI * freq * r[5] / LIGHT;
    r[10] = 4.0 * PI * freq * r[6] / LIGHT;
    r[11] = 4.0 * PI * freq * r[7] / LIGHT;
    double c4 = 3.0 * LIGHT * LIGHT * LIGHT * LIGHT;
    double f4 = 8.0 * PI * freq * freq * freq * freq;
    r[12] = f4 * m_iso / c4;
    r[13] = f4 * m_x / c4;
    r[14] = f4 * m_y / c4;
    r[15] = f4 * m_z / c4;
    for (int k = 0; k < 4; ++k) r[16 + k] = r[8 + k] + r[12 + k];
}

/* Neighbour counts by type within r0 (self included).
 * src/target/target_module.f90:417-464. */
void pxo_neighbours(int n, int nt, const double *coord, const int *type, const double *r0,
                    int *nb_counts, int *nb_total) {
#pragma omp parallel for
    for (int i = 0; i < n; ++i) {
        int tot = 1;
        for (int t = 0; t < nt; ++t) nb_counts[t + nt * i] = 0;
        nb_counts[type[i] + nt * i] = 1;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double d[3];
            double r = dist3(coord, i, j, d);
            if (r <= r0[type[i] + nt * type[j]]) {
                ++tot;
                nb_counts[type[j] + nt * i] += 1;
            }
        }
        nb_total[i] = tot;
    }
}

/* ----------------------------------------------------------------------- */
/* Spectrum driver: src/solvers/algorithm_module.f90:150-465.               */
/* method 0 = inversion (ZGETRF/ZGETRS), 1 = GMRES on the stored matrix,    */
/* 2 = GMRES on the fly.  tol is the absolute threshold (already scaled by  */
/* sqrt(n_var) as input_module.f90:251 does).                               */
/* ----------------------------------------------------------------------- */
int pxo_solve_spectrum(const pxo_atoms *a, int n_freq, const double *freq, const pxo_cplx *alpha_w,
                       double e0, int rhs_form, int pol_mask, int method, double tol, int m,
                       pxo_cplx *variables, pxo_cplx *polar, double *results, int *iters) {
    int nv = pxo_n_var(a), nt = a->n_types;
    int status = 0;
    double *M = NULL;
    if (method != 2) {
        M = malloc(sizeof(double) * (size_t)nv * nv);
        pxo_constant_matrix(a, M);
    }
    pxo_cplx *rhs = malloc(sizeof(pxo_cplx) * (size_t)nv * 3);
    pxo_cplx *var = malloc(sizeof(pxo_cplx) * (size_t)nv * 3);
    pxo_cplx *A = method == 2 ? NULL : malloc(sizeof(pxo_cplx) * (size_t)nv * nv);
    pxo_cplx *alpha_t = malloc(sizeof(pxo_cplx) * (nt > 0 ? nt : 1));
    for (int f = 0; f < n_freq && status == 0; ++f) {
        for (int t = 0; t < nt; ++t) alpha_t[t] = alpha_w ? alpha_w[t + (size_t)nt * f] : 1.0;
        pxo_rhs(a, e0, rhs_form, freq[f], method == 2, rhs);
        memset(var, 0, sizeof(pxo_cplx) * (size_t)nv * 3);
        if (A) pxo_dynamic_matrix(a, M, freq[f], alpha_t, A);
        if (method == 0) {
            pxo_cplx *rr = malloc(sizeof(pxo_cplx) * (size_t)nv * 3);
            memcpy(rr, rhs, sizeof(pxo_cplx) * (size_t)nv * 3);
            for (int k = 0; k < 3; ++k)
                if (!(pol_mask & (1 << k)))
                    for (int i = 0; i < nv; ++i) rr[i + (size_t)nv * k] = 0.0;
            if (pxo_zgesv(nv, 3, A, rr) != 0) status = 3;
            memcpy(var, rr, sizeof(pxo_cplx) * (size_t)nv * 3);
            free(rr);
            if (iters)
                for (int k = 0; k < 3; ++k) iters[k + 3 * f] = 0;
        } else {
            for (int k = 0; k < 3; ++k) {
                int it = 0;
                if (pol_mask & (1 << k)) {
                    double tr;
                    int st = method == 1
                                 ? pxo_gmres_dense(nv, A, rhs + (size_t)nv * k, var + (size_t)nv * k, tol, m, &it,
                                                   NULL, &tr)
                                 : pxo_gmres_otf(a, freq[f], alpha_t, rhs + (size_t)nv * k, var + (size_t)nv * k,
                                                 tol, m, &it, NULL, &tr);
                    if (st != 0) status = 2;
                }
                if (iters) iters[k + 3 * f] = it;
            }
        }
        pxo_cplx pol[9];
        pxo_polar(a, e0, var, pol);
        if (polar) memcpy(polar + 9 * (size_t)f, pol, sizeof(pol));
        if (results) {
            double r[20];
            pxo_results_row(freq[f], pol, r);
            for (int c = 0; c < 20; ++c) results[f + (size_t)n_freq * c] = r[c];
        }
        if (variables) memcpy(variables + (size_t)nv * 3 * f, var, sizeof(pxo_cplx) * (size_t)nv * 3);
    }
    free(M);
    free(rhs);
    free(var);
    free(A);
    free(alpha_t);
    return status;
}

/* ----------------------------------------------------------------------- */
/* Etchegoin permittivity (polar = 0) or interband polarizability          */
/* (polar = 1) of silver and gold.  src/parameters/parameters_module.f90:   */
/* 107-194 (Ag) and 295-375 (Au).  The numbers are the published           */
/* Etchegoin-Le Ru-Meyer fit parameters.                                    */
/* ----------------------------------------------------------------------- */
static double au_to_nm(double f) { return 10000000.0 / (f * AU_TO_CM); }

static pxo_cplx cp_term(double aj, double lj, double gj, double phi, double f3) {
    pxo_cplx e_p = cexp(I * phi), e_m = cexp(-I * phi);
    return aj * lj * (e_p / ((lj - f3) - I * gj) + e_m / ((lj + f3) + I * gj));
}

void pxo_silver_etchegoin(int polar, int n, const double *freq, pxo_cplx *out) {
    const double eps_inf_re = 1.469322, eps_inf_im = 0.2450325;
    const double A[3] = {0.1632212, 0.9672775, 0.6150938};
    const double PH[3] = {-1.709221, 0.1546664, -1.782307};
    const double lam_p = 137.9643, gam_p = 1.0 / 74484.21;
    const double L[3] = {1.0 / 307.9550, 1.0 / 260.7301, 1.0 / 236.2128};
    const double G[3] = {1.0 / 4277.721, 1.0 / 1051.474, 1.0 / 1008.622};
    double n0 = 10.49 / 107.8682 * 6.022e23 / pow(1.0e8 * TOBOHR, 3);
    double wp2 = 4.0 * PI * n0;
    for (int i = 0; i < n; ++i) {
        double f3 = 1.0 / au_to_nm(freq[i]);
        pxo_cplx s = 0.0;
        for (int k = 0; k < 3; ++k) s += cp_term(A[k], L[k], G[k], PH[k], f3);
        if (polar)
            out[i] = (I * eps_inf_im + s) / wp2;
        else
            out[i] = (eps_inf_re + I * eps_inf_im) - 1.0 / (lam_p * lam_p * f3 * (f3 + I * gam_p)) + s;
    }
}

void pxo_gold_etchegoin(int polar, int n, const double *freq, pxo_cplx *out) {
    const double eps_inf = 1.54;
    const double A[2] = {1.27, 1.10};
    const double PH[2] = {-PI / 4.0, -PI / 4.0};
    const double lam_p = 143.0, gam_p = 1.0 / 14500.0;
    const double L[2] = {1.0 / 470.0, 1.0 / 325.0};
    const double G[2] = {1.0 / 1900.0, 1.0 / 1060.0};
    double n0 = 19.320 / 196.96657 * 6.022e23 / pow(1.0e8 * TOBOHR, 3);
    double wp2 = 4.0 * PI * n0;
    for (int i = 0; i < n; ++i) {
        double f3 = 1.0 / au_to_nm(freq[i]);
        pxo_cplx s = 0.0;
        for (int k = 0; k < 2; ++k) s += cp_term(A[k], L[k], G[k], PH[k], f3);
        if (polar) {
            pxo_cplx v = s / wp2;
            if (cimag(v) <= 0.0) v = creal(v) + I * 1.0e-10;
            out[i] = v;
        } else {
            out[i] = eps_inf - 1.0 / (lam_p * lam_p * f3 * (f3 + I * gam_p)) + s;
        }
    }
}

/* Field at n_points points (bohr, 3 per point) for the variables var of one
 * field direction: the incident e0 along i_pol plus, for every atom i with
 * d = r_i - p and |d| > 0.1 bohr,
 *   the field of its Gaussian charge (width R_q of atom i alone)
 *     E -= (erf(|d|/R) - 2|d|/(sqrt(pi) R) exp(-|d|^2/R^2)) q_i d / |d|^3
 *   (src/matrix/matrix_module.f90:1140-1172), and for omegaFQFmu the field
 *   of its Gaussian dipole (width R_mu of atom i), E += T mu_i with
 *     T_ab = (-3 d_a d_b / |d|^5 + delta_ab / |d|^3) fD1 + fD2 d_a d_b / |d|^2,
 *     fD1 = erf(|d|/R) - 2|d|/(sqrt(pi) R) exp(-|d|^2/R^2),
 *     fD2 = 4 / (sqrt(pi) R^3) exp(-|d|^2/R^2)
 *   (:1174-1246),
 * summed over the atoms as calculate_induced_field_at_point of
 * src/wfq/wfq_module.f90:592-617 and src/wfq/wfqfmu_module.f90:973-1003
 * (dipole of atom i at var[n + 3i + c]).  E is 3 x n_points, point-major. */
void pxo_induced_field(const pxo_atoms *a, int n_points, const double *points, const pxo_cplx *var, int i_pol,
                       double e0, pxo_cplx *E) {
    const int n = a->n_atoms;
#pragma omp parallel for
    for (int p = 0; p < n_points; ++p) {
        pxo_cplx e[3] = {0, 0, 0};
        e[i_pol] = e0;
        for (int i = 0; i < n; ++i) {
            double d[3];
            for (int c = 0; c < 3; ++c) d[c] = a->coord[3 * i + c] - points[3 * p + c];
            const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (!(r > 1.0e-1)) continue;
            const double r2 = r * r, r3 = r * r2;
            const int t = a->type[i];
            const double Rq = a->r_q[t];
            const double fq = erf(r / Rq) - (2.0 / (sqrt(PI) * Rq)) * r * exp(-r2 / (Rq * Rq));
            for (int c = 0; c < 3; ++c) e[c] -= fq * d[c] * var[i] / r3;
            if (a->model == PXO_WFQFMU) {
                const double R = a->r_mu[t], r5 = r2 * r3;
                const double ex = exp(-r2 / (R * R));
                const double f1 = erf(r / R) - ((2.0 * r) / (sqrt(PI) * R)) * ex;
                const double f2 = (4.0 / (sqrt(PI) * R * R * R)) * ex;
                const pxo_cplx *mu = var + n + 3 * i;
                for (int u = 0; u < 3; ++u) {
                    pxo_cplx s = 0;
                    for (int v = 0; v < 3; ++v) {
                        const double tuv = ((-3.0 * d[u] * d[v]) / r5 + (u == v ? 1.0 / r3 : 0.0)) * f1 +
                                           f2 * (d[u] * d[v]) / r2;
                        s += mu[v] * tuv;
                    }
                    e[u] += s;
                }
            }
        }
        for (int c = 0; c < 3; ++c) E[3 * p + c] = e[c];
    }
}

/* Plasmon density at n_points points (bohr, 3 per point): for every atom i
 * with d = r_i - p,
 *   the normalised Gaussian of its charge,
 *     rho_q += q_i exp(-|d|^2/R_q^2) / (pi^{3/2} R_q^3)
 *   (src/matrix/matrix_module.f90:1258-1284), and for omegaFQFmu minus the
 *   derivative of the Gaussian of R_mu along its dipole,
 *     rho_mu -= 2 exp(-|d|^2/R_mu^2) / (pi^{3/2} R_mu^5) (mu_i . d)
 *   (:1295-1324),
 * summed over the atoms as src/wfq/wfq_module.f90:625-648 and
 * src/wfq/wfqfmu_module.f90:1012-1081.  rho = rho_q + rho_mu; rho_q and
 * rho_mu may be NULL. */
void pxo_plasmon_density(const pxo_atoms *a, int n_points, const double *points, const pxo_cplx *var, pxo_cplx *rho,
                         pxo_cplx *rho_q, pxo_cplx *rho_mu) {
    const int n = a->n_atoms;
    const double pi32 = PI * sqrt(PI);
#pragma omp parallel for
    for (int p = 0; p < n_points; ++p) {
        pxo_cplx sq = 0, smu = 0;
        for (int i = 0; i < n; ++i) {
            double d[3];
            for (int c = 0; c < 3; ++c) d[c] = a->coord[3 * i + c] - points[3 * p + c];
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const int t = a->type[i];
            const double Rq = a->r_q[t];
            sq += var[i] * (exp(-r2 / (Rq * Rq)) / (pi32 * Rq * Rq * Rq));
            if (a->model == PXO_WFQFMU) {
                const double R = a->r_mu[t];
                const double g = exp(-r2 / (R * R)) / (pi32 * R * R * R);
                const pxo_cplx *mu = var + n + 3 * i;
                smu -= (2.0 * g / (R * R)) * (mu[0] * d[0] + mu[1] * d[1] + mu[2] * d[2]);
            }
        }
        rho[p] = sq + smu;
        if (rho_q) rho_q[p] = sq;
        if (rho_mu) rho_mu[p] = smu;
    }
}
