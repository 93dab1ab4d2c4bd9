/*
 * pxo_bem.c -- FP64 CPU oracle: BEM tesserae, DPCM/IEFPCM operators
 * (approximate and accurate Green-function derivative), RHS, polarizability
 * and the per-frequency driver.
 *
 * TEST INFRASTRUCTURE ONLY (see pxo.h).  Not linked into the product.
 *
 * What pins these functions (tests/test_oracle_golden.py,
 * tests/test_oracle_invariants.py):
 *   - bem_dpcm_accurate_inversion and bem_ief_*_inversion reference logs:
 *     RHS (E10.3), D + lambda I real/imag (E11.4), charges, polarizabilities
 *     and cross sections (E14.6) on the 284-tessera matlab.msh mesh; the
 *     quadrature (Legendre-Gauss-Lobatto polar rule, 28-point Strang-Fix
 *     near-field rule) enters the printed D and is pinned through it;
 *   - the accurate-mode sum rule sum_k A_k D_ki = 2 pi A_i (exact by
 *     construction of the diagonal) and the quasi-static sphere limit
 *     alpha -> R^3 (eps - 1)/(eps + 2) on an icosphere (within 2 %);
 *   - the field of the surface charges at points (tests/test_analysis_bem.py):
 *     the analysis_field_3d_y_bem cube and analysis_field_2d_y_xy_bem CSV,
 *     zeros inside included, and the Coulomb field of one charge.
 */
#include "pxo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.14159265358979323846
#define DIAG_PARAM 1.0694
#define LIGHT 137.035227

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Centroid, area, normal and tangents of each flat triangle.
 * src/bem/bem_module.f90:420-491 (create_tesserae). */
void pxo_tesserae(int n_tess, const double *v9, double normal_factor, double *coord, double *normal, double *tang1,
                  double *tang2, double *area) {
    for (int i = 0; i < n_tess; ++i) {
        const double *p1 = v9 + 9 * i, *p2 = p1 + 3, *p3 = p1 + 6;
        double a[3], b[3], c[3];
        for (int k = 0; k < 3; ++k) {
            coord[3 * i + k] = (p1[k] + p2[k] + p3[k]) / 3.0;
            a[k] = p2[k] - p1[k];
            b[k] = p3[k] - p1[k];
        }
        double na = sqrt(dot3(a, a));
        for (int k = 0; k < 3; ++k) tang1[3 * i + k] = -a[k] / na;
        c[0] = a[1] * b[2] - a[2] * b[1];
        c[1] = a[2] * b[0] - a[0] * b[2];
        c[2] = a[0] * b[1] - a[1] * b[0];
        double nc = sqrt(dot3(c, c));
        area[i] = nc / 2.0;
        for (int k = 0; k < 3; ++k) normal[3 * i + k] = normal_factor * c[k] / nc;
        const double *nn = normal + 3 * i, *t1 = tang1 + 3 * i;
        tang2[3 * i + 0] = nn[1] * t1[2] - nn[2] * t1[1];
        tang2[3 * i + 1] = nn[2] * t1[0] - nn[0] * t1[2];
        tang2[3 * i + 2] = nn[0] * t1[1] - nn[1] * t1[0];
    }
}

/* Legendre-Gauss-Lobatto nodes (descending, x_0 = 1) and weights for
 * npol points: Newton iteration on (1-x^2) P'_{N}, N = npol-1, from the
 * Chebyshev-Gauss-Lobatto guess.  Same nodes as
 * src/bem/accurate_green_function.f90:383-477. */
void pxo_lgl(int npol, double *x, double *w) {
    int N = npol - 1;
    double *P = malloc(sizeof(double) * npol * npol);
    double *xold = malloc(sizeof(double) * npol);
    for (int i = 0; i < npol; ++i) x[i] = cos(PI * i / N);
    for (int it = 0; it < 100; ++it) {
        double diff = 0.0;
        for (int i = 0; i < npol; ++i) {
            xold[i] = x[i];
            P[i] = 1.0;
            P[i + npol] = x[i];
            for (int k = 2; k <= N; ++k)
                P[i + npol * k] = ((2.0 * k - 1.0) * x[i] * P[i + npol * (k - 1)] - (k - 1.0) * P[i + npol * (k - 2)]) / k;
            x[i] = xold[i] - (x[i] * P[i + npol * N] - P[i + npol * (N - 1)]) / (npol * P[i + npol * N]);
            double d = fabs(x[i] - xold[i]);
            if (d > diff) diff = d;
        }
        if (diff <= 2.2204e-16) break;
    }
    for (int i = 0; i < npol; ++i) {
        double p0 = 1.0, p1 = x[i];
        for (int k = 2; k <= N; ++k) {
            double p2 = ((2.0 * k - 1.0) * x[i] * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        w[i] = 2.0 / (N * (N + 1.0) * p1 * p1);
    }
    free(P);
    free(xold);
}

/* 28-point degree-11 rule on the unit triangle (Strang & Fix 1973, as used
 * at src/bem/accurate_green_function.f90:138-238): barycentric (x, y) and
 * weights summing to 1.  Orbits: centroid; 3 x {a, b, b}; 6 x {a, b, c}. */
static void strang_fix28(double *xt, double *yt, double *wt) {
    static const double s3[6][3] = {
        /* a, b (b repeated), weight */
        {0.9480217181434233, 0.02598914092828833, 0.008744311553736190},
        {0.8114249947041546, 0.09428750264792270, 0.03808157199393533},
        {0.01072644996557060, 0.4946367750172147, 0.01885544805613125},
        {0.5853132347709715, 0.2073433826145142, 0.07215969754474100},
        {0.1221843885990187, 0.4389078057004907, 0.06932913870553720},
        {0, 0, 0}};
    static const double s6[2][4] = {
        {0.6779376548825902, 0.04484167758913055, 0.27722066752827925, 0.04105631542928860},
        {0.8588702812826364, 0.0, 0.1411297187173636, 0.007362383783300573}};
    int k = 0;
    xt[k] = yt[k] = 1.0 / 3.0;
    wt[k++] = 0.08797730116222190;
    for (int o = 0; o < 5; ++o) {
        double a = s3[o][0], b = s3[o][1], w = s3[o][2];
        xt[k] = a; yt[k] = b; wt[k++] = w;
        xt[k] = b; yt[k] = a; wt[k++] = w;
        xt[k] = b; yt[k] = b; wt[k++] = w;
    }
    for (int o = 0; o < 2; ++o) {
        double a = s6[o][0], b = s6[o][1], c = s6[o][2], w = s6[o][3];
        /* the six permutations in the order of the published table:
         * (a,b) (a,c) (b,a) (b,c) (c,a) (c,b) */
        double px[6] = {a, a, b, b, c, c};
        double py[6] = {b, c, a, c, a, b};
        for (int p = 0; p < 6; ++p) {
            xt[k] = px[p];
            yt[k] = py[p];
            wt[k++] = w;
        }
    }
}

/* Polar-coordinate rule around the centroid of a flat triangle (8 x 6 LGL
 * points in each of the 3 sub-triangles), barycentric (ix, iy) and weights
 * normalised to 1.  src/bem/accurate_green_function.f90:245-376. */
static void polar_rule(double *ix, double *iy, double *w) {
    const int n1 = 8, n2 = 6, nt = n1 * n2;
    double p1[8], w1[8], p2[6], w2[6];
    pxo_lgl(n1, p1, w1);
    pxo_lgl(n2, p2, w2);
    const double phi0 = (120.0 / 180.0) * PI;
    double tot = 0.0;
    for (int i2 = 0; i2 < n2; ++i2)
        for (int i1 = 0; i1 < n1; ++i1) {
            int i = i1 + n1 * i2;
            double rho = 0.5 * (p1[i1] + 1.0 + 1.0e-06);
            double phi = PI * (270.0 + 60.0 * p2[i2]) / 180.0;
            double sc = 1.0 / fabs(2.0 * sin(phi));
            for (int c = 0; c < 3; ++c) {
                double tx = cos(phi + c * phi0) * rho * sc;
                double ty = sin(phi + c * phi0) * rho * sc;
                ix[i + c * nt] = (1.0 - sqrt(3.0) * tx - ty) / 3.0;
                iy[i + c * nt] = (1.0 + sqrt(3.0) * tx - ty) / 3.0;
                w[i + c * nt] = w1[i1] * w2[i2] * rho * sc * sc;
                tot += w[i + c * nt];
            }
        }
    for (int i = 0; i < 3 * nt; ++i) w[i] /= tot;
}

/* D (the normal derivative of the Green function, n x n column-major) in
 * approximate (centroid) or accurate (near-field quadrature + sum rule)
 * form; optionally the IEF matrices D_ief = D(unrefined diagonal) - 2 pi I
 * and S A^-1.  src/bem/bem_module.f90:678-782 and
 * src/bem/accurate_green_function.f90:25-838. */
void pxo_bem_d(const pxo_mesh *m, int accurate, double sphere_r, double *D, double *D_ief, double *SA) {
    int n = m->n;
    double *gdiag = NULL, *radius = NULL, *icoord = NULL;
    double wt[28], xt[28], yt[28];
    if (accurate) {
        /* diagonal: polar rule around the own centroid */
        double ix[144], iy[144], iw[144];
        polar_rule(ix, iy, iw);
        double *dist = malloc(sizeof(double) * 144 * (size_t)n);
        double *pts = malloc(sizeof(double) * 3 * 144 * (size_t)n);
        double dmax = 0.0;
        for (int i = 0; i < n; ++i) {
            const double *v = m->verts + 9 * i;
            for (int k = 0; k < 144; ++k) {
                double *p = pts + 3 * (144 * (size_t)i + k);
                for (int c = 0; c < 3; ++c) p[c] = ix[k] * v[c] + iy[k] * v[3 + c] + (1.0 - ix[k] - iy[k]) * v[6 + c];
                double d[3] = {m->coord[3 * i] - p[0], m->coord[3 * i + 1] - p[1], m->coord[3 * i + 2] - p[2]};
                dist[144 * (size_t)i + k] = sqrt(dot3(d, d));
                if (dist[144 * (size_t)i + k] > dmax) dmax = dist[144 * (size_t)i + k];
            }
        }
        double clamp = dmax * 1.0e-4;
        gdiag = malloc(sizeof(double) * 3 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            double gn = 0.0, g1 = 0.0, g2 = 0.0;
            const double *ni = m->normal + 3 * i, *t1 = m->tang1 + 3 * i, *t2 = m->tang2 + 3 * i;
            for (int k = 0; k < 144; ++k) {
                const double *p = pts + 3 * (144 * (size_t)i + k);
                double v[3] = {m->coord[3 * i] - p[0], m->coord[3 * i + 1] - p[1], m->coord[3 * i + 2] - p[2]};
                double dk = dist[144 * (size_t)i + k];
                gn -= iw[k] * m->area[i] * dot3(v, ni) / (dk * dk * dk);
                if (dk < 1.0e-04) dk = clamp;
                g1 -= iw[k] * m->area[i] * dot3(v, t1) / (dk * dk * dk);
                g2 -= iw[k] * m->area[i] * dot3(v, t2) / (dk * dk * dk);
            }
            for (int c = 0; c < 3; ++c) gdiag[3 * i + c] = ni[c] * gn + t1[c] * g1 + t2[c] * g2;
        }
        free(dist);
        free(pts);
        /* off-diagonal quadrature points and the near-field radii */
        strang_fix28(xt, yt, wt);
        icoord = malloc(sizeof(double) * 3 * 28 * (size_t)n);
        radius = malloc(sizeof(double) * (size_t)n);
        for (int j = 0; j < n; ++j) {
            const double *v = m->verts + 9 * j;
            for (int q = 0; q < 28; ++q)
                for (int c = 0; c < 3; ++c)
                    icoord[3 * (28 * (size_t)j + q) + c] = xt[q] * v[c] + yt[q] * v[3 + c] + (1.0 - xt[q] - yt[q]) * v[6 + c];
            double rmax = 0.0;
            for (int c = 0; c < 3; ++c) {
                double d[3] = {m->coord[3 * j] - v[3 * c], m->coord[3 * j + 1] - v[3 * c + 1],
                               m->coord[3 * j + 2] - v[3 * c + 2]};
                double r = sqrt(dot3(d, d));
                if (r > rmax) rmax = r;
            }
            radius[j] = rmax;
        }
    }
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < n; ++i) {
        const double *ci = m->coord + 3 * i, *ni = m->normal + 3 * i;
        double ni2 = dot3(ci, ci);
        for (int j = 0; j < n; ++j) {
            const double *cj = m->coord + 3 * j;
            double d[3] = {ci[0] - cj[0], ci[1] - cj[1], ci[2] - cj[2]};
            double r = sqrt(dot3(d, d));
            double val;
            if (accurate) {
                /* near-field selection: |c_i - c_j| (from the norm
                 * expansion) minus the radius of tessera j */
                double t1 = ni2 + dot3(cj, cj) - 2.0 * dot3(ci, cj);
                if (t1 < 1.0e-10f) t1 = 0.0;
                double sel = sqrt(t1) - radius[j];
                if (sel >= 0.0) {
                    val = m->area[j] * dot3(d, ni) / (r * r * r);
                } else if (i == j) {
                    val = -dot3(gdiag + 3 * i, ni);
                } else {
                    double ox = 0.0, oy = 0.0, oz = 0.0;
                    for (int q = 0; q < 28; ++q) {
                        const double *p = icoord + 3 * (28 * (size_t)j + q);
                        double tx = ci[0] - p[0], ty = ci[1] - p[1], tz = ci[2] - p[2];
                        double tn = sqrt(tx * tx + ty * ty + tz * tz);
                        double t3 = tn * tn * tn;
                        ox -= wt[q] * m->area[j] * tx / t3;
                        oy -= wt[q] * m->area[j] * ty / t3;
                        oz -= wt[q] * m->area[j] * tz / t3;
                    }
                    val = -(ox * ni[0] + oy * ni[1] + oz * ni[2]);
                }
            } else {
                val = i != j ? m->area[j] * dot3(d, ni) / (r * r * r)
                             : DIAG_PARAM * sqrt(4.0 * PI * m->area[i]) / (2.0 * sphere_r);
            }
            D[i + (size_t)n * j] = val;
            if (D_ief) D_ief[i + (size_t)n * j] = i == j ? val - 2.0 * PI : val;
            if (SA)
                SA[i + (size_t)n * j] = i == j ? DIAG_PARAM * sqrt(4.0 * PI * m->area[i]) * (1.0 / m->area[i])
                                              : (m->area[j] / r) * (1.0 / m->area[j]);
        }
    }
    if (accurate) {
        /* Gauss sum rule on the diagonal: accurate_green_function.f90:796-838 */
        double *colsum = malloc(sizeof(double) * (size_t)n);
#pragma omp parallel for
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += (D[k + (size_t)n * i] / m->area[i]) * m->area[k];
            colsum[i] = s;
        }
        for (int i = 0; i < n; ++i) {
            const double *ni = m->normal + 3 * i;
            double g[3];
            for (int c = 0; c < 3; ++c) g[c] = ni[c] * (-2.0 * PI + colsum[i]) + gdiag[3 * i + c];
            D[i + (size_t)n * i] = -dot3(g, ni);
        }
        free(colsum);
        free(gdiag);
        free(radius);
        free(icoord);
    }
}

/* lambda(omega) = 2 pi (eps_s + eps) / (eps_s - eps).
 * src/bem/bem_module.f90:898-906. */
pxo_cplx pxo_bem_lambda(double eps_solvent, pxo_cplx eps) {
    return 2.0 * PI * (eps_solvent + eps) / (eps_solvent - eps);
}

/* DPCM: rhs(:,k) = -n . E0 e_k; IEF: rhs(:,k) = D_ief (-E0 e_k . c).
 * src/bem/bem_module.f90:788-873. */
void pxo_bem_rhs(const pxo_mesh *m, int ief, const double *D_ief, double e0, pxo_cplx *rhs) {
    int n = m->n;
    for (int k = 0; k < 3; ++k) {
        if (!ief) {
            for (int i = 0; i < n; ++i) rhs[i + (size_t)n * k] = -m->normal[3 * i + k] * e0;
        } else {
#pragma omp parallel for
            for (int i = 0; i < n; ++i) {
                double s = 0.0;
                for (int j = 0; j < n; ++j) s += D_ief[i + (size_t)n * j] * (-e0 * m->coord[3 * j + k]);
                rhs[i + (size_t)n * k] = s;
            }
        }
    }
}

/* DPCM: A = D + lambda I; IEF: A = -(D + lambda I) S A^-1.
 * src/bem/bem_module.f90:881-933. */
void pxo_bem_matrix(const pxo_mesh *m, int ief, const double *D, const double *SA, pxo_cplx lambda, pxo_cplx *A) {
    int n = m->n;
    if (!ief) {
        for (size_t e = 0; e < (size_t)n * n; ++e) A[e] = D[e];
        for (int i = 0; i < n; ++i) A[i + (size_t)n * i] += lambda;
        return;
    }
#pragma omp parallel for
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            pxo_cplx s = 0.0;
            for (int p = 0; p < n; ++p) {
                pxo_cplx dip = D[i + (size_t)n * p] + (i == p ? lambda : 0.0);
                s += dip * SA[p + (size_t)n * j];
            }
            A[i + (size_t)n * j] = -s;
        }
}

/* alpha_ab = sum_i A_i sigma_i c_a(i) / E0 (DPCM) or sum_i q_i c_a(i) / E0
 * (IEF).  src/bem/bem_module.f90:1051-1142. */
void pxo_bem_polar(const pxo_mesh *m, int ief, double e0, const pxo_cplx *var, pxo_cplx *polar) {
    int n = m->n;
    for (int k = 0; k < 9; ++k) polar[k] = 0.0;
    for (int b = 0; b < 3; ++b)
        for (int i = 0; i < n; ++i) {
            pxo_cplx v = var[i + (size_t)n * b] * (ief ? 1.0 : m->area[i]);
            for (int p = 0; p < 3; ++p) polar[p + 3 * b] += v * m->coord[3 * i + p] / e0;
        }
}

/* Per-frequency BEM loop: src/solvers/algorithm_module.f90:189-410, with
 * the charge constraint of :266-319 (inversion only).  method 1 (GMRES on
 * the stored operator) solves the same system as inversion, for IEF too. */
int pxo_bem_solve_spectrum(const pxo_mesh *m, int ief, int accurate, double sphere_r, double eps_solvent, int n_freq,
                           const double *freq, const pxo_cplx *eps_w, double e0, int charge_constraint, int method,
                           double tol, int mdim, pxo_cplx *variables, pxo_cplx *polar, double *results, int *iters) {
    int n = m->n, status = 0;
    double *D = malloc(sizeof(double) * (size_t)n * n);
    double *Dief = ief ? malloc(sizeof(double) * (size_t)n * n) : NULL;
    double *SA = ief ? malloc(sizeof(double) * (size_t)n * n) : NULL;
    pxo_bem_d(m, accurate, sphere_r, D, Dief, SA);
    pxo_cplx *rhs = malloc(sizeof(pxo_cplx) * (size_t)n * 3);
    pxo_bem_rhs(m, ief, Dief, e0, rhs);
    pxo_cplx *A = malloc(sizeof(pxo_cplx) * (size_t)n * n);
    pxo_cplx *Acp = malloc(sizeof(pxo_cplx) * (size_t)n * n);
    pxo_cplx *var = malloc(sizeof(pxo_cplx) * (size_t)n * 3);
    pxo_cplx *one = malloc(sizeof(pxo_cplx) * (size_t)n);
    for (int f = 0; f < n_freq && status == 0; ++f) {
        pxo_cplx lam = pxo_bem_lambda(eps_solvent, eps_w[f]);
        pxo_bem_matrix(m, ief, D, SA, lam, A);
        if (method == 0) {
            memcpy(Acp, A, sizeof(pxo_cplx) * (size_t)n * n);
            memcpy(var, rhs, sizeof(pxo_cplx) * (size_t)n * 3);
            if (pxo_zgesv(n, 3, Acp, var) != 0) status = 3;
            if (charge_constraint && status == 0) {
                for (int i = 0; i < n; ++i) one[i] = 1.0;
                memcpy(Acp, A, sizeof(pxo_cplx) * (size_t)n * n);
                if (pxo_zgesv(n, 1, Acp, one) != 0) status = 3;
                pxo_cplx s1 = 0.0;
                for (int i = 0; i < n; ++i) s1 += one[i];
                for (int k = 0; k < 3; ++k) {
                    pxo_cplx sk = 0.0;
                    for (int i = 0; i < n; ++i) sk += var[i + (size_t)n * k];
                    for (int i = 0; i < n; ++i) var[i + (size_t)n * k] -= one[i] * sk / s1;
                }
            }
            if (iters)
                for (int k = 0; k < 3; ++k) iters[k + 3 * f] = 0;
        } else {
            for (int k = 0; k < 3; ++k) {
                int it = 0;
                double tr;
                if (pxo_gmres_dense(n, A, rhs + (size_t)n * k, var + (size_t)n * k, tol, mdim, &it, NULL, &tr) != 0)
                    status = 2;
                if (iters) iters[k + 3 * f] = it;
            }
        }
        pxo_cplx pol[9];
        pxo_bem_polar(m, ief, e0, var, pol);
        if (polar) memcpy(polar + 9 * (size_t)f, pol, sizeof pol);
        if (results) {
            double r[20];
            pxo_results_row(freq[f], pol, r);
            for (int c = 0; c < 20; ++c) results[f + (size_t)n_freq * c] = r[c];
        }
        if (variables) memcpy(variables + (size_t)n * 3 * f, var, sizeof(pxo_cplx) * (size_t)n * 3);
    }
    free(D);
    free(Dief);
    free(SA);
    free(rhs);
    free(A);
    free(Acp);
    free(var);
    free(one);
    return status;
}

/* Field at n_points points (bohr) of the BEM surface charges q (n_tess
 * complex) at the tessera centroids c (bohr), as
 * calculate_induced_field_at_point_bem (src/bem/bem_module.f90:1151-1213):
 * with m the mean of the centroids and l the first centroid nearest to the
 * point p, p is outside when |p - m| > |c_l - m| + 0.5 Angstrom; outside,
 * E = e0 along i_pol minus sum_i q_i d / |d|^3 over |d| > 0.1 bohr,
 * d = c_i - p (src/matrix/matrix_module.f90:1444-1464); inside, E = 0. */
void pxo_bem_induced_field(int n_tess, const double *c, const pxo_cplx *q, int n_points, const double *points,
                           int i_pol, double e0, pxo_cplx *E) {
    double m[3] = {0, 0, 0};
    for (int i = 0; i < n_tess; ++i)
        for (int k = 0; k < 3; ++k) m[k] += c[3 * i + k] / n_tess;
#pragma omp parallel for
    for (int p = 0; p < n_points; ++p) {
        const double *x = points + 3 * p;
        int l = 0;
        double best = 1.0e6 * 1.8897261254578281;
        for (int i = 0; i < n_tess; ++i) {
            const double dx = c[3 * i] - x[0], dy = c[3 * i + 1] - x[1], dz = c[3 * i + 2] - x[2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (best > r) {
                best = r;
                l = i;
            }
        }
        const double rl = sqrt((c[3 * l] - m[0]) * (c[3 * l] - m[0]) + (c[3 * l + 1] - m[1]) * (c[3 * l + 1] - m[1]) +
                               (c[3 * l + 2] - m[2]) * (c[3 * l + 2] - m[2])) +
                          0.5 * 1.8897261254578281;
        const double rp = sqrt((x[0] - m[0]) * (x[0] - m[0]) + (x[1] - m[1]) * (x[1] - m[1]) +
                               (x[2] - m[2]) * (x[2] - m[2]));
        pxo_cplx e[3] = {0, 0, 0};
        if (rp > rl) {
            e[i_pol] = e0;
            for (int i = 0; i < n_tess; ++i) {
                double d[3];
                for (int k = 0; k < 3; ++k) d[k] = c[3 * i + k] - x[k];
                const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (!(r > 1.0e-1)) continue;
                const double r3 = r * r * r;
                for (int k = 0; k < 3; ++k) e[k] -= d[k] * q[i] / r3;
            }
        }
        for (int k = 0; k < 3; ++k) E[3 * p + k] = e[k];
    }
}
