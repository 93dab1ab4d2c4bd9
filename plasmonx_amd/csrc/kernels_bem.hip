// BEM (DPCM / IEFPCM) operators on gfx950.
//
// Restates src/bem/bem_module.f90 (tesserae :420-491, constant matrix
// :678-782, RHS :788-873, dynamic matrix :881-933, shift :977-997) and the
// accurate Green-function derivative of src/bem/accurate_green_function.f90
// (polar diagonal quadrature :43-130,245-376, LGL nodes :383-477, 28-point
// Strang-Fix rule :138-238, near-field selection :547-657, diagonal
// refinement :662-737, off-diagonal refinement :745-788, sum rule :796-838).
//
// D, D_ief and S A^-1 are n x n column-major FP64 in HBM (plus FP32 copies for
// the mixed-precision products); every element (i, j) is one work-item.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pxh_internal.h"

namespace pxh {

namespace {

constexpr double kDiagParam = 1.0694;

struct Quad {
    double ix[144], iy[144], iw[144];  // polar rule (barycentric x, y; weights summing to 1)
    double xt[28], yt[28], wt[28];     // 28-point rule
};

void lgl(int npol, double *x, double *w) {
    const int N = npol - 1;
    for (int i = 0; i < npol; ++i) x[i] = std::cos(M_PI * i / N);
    for (int it = 0; it < 100; ++it) {
        double diff = 0.0;
        for (int i = 0; i < npol; ++i) {
            double p0 = 1.0, p1 = x[i];
            for (int k = 2; k <= N; ++k) {
                const double p2 = ((2.0 * k - 1.0) * x[i] * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double xn = x[i] - (x[i] * p1 - p0) / (npol * p1);
            diff = std::max(diff, std::fabs(xn - x[i]));
            x[i] = xn;
        }
        if (diff <= 2.2204e-16) break;
    }
    for (int i = 0; i < npol; ++i) {
        double p0 = 1.0, p1 = x[i];
        for (int k = 2; k <= N; ++k) {
            const double p2 = ((2.0 * k - 1.0) * x[i] * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        w[i] = 2.0 / (N * (N + 1.0) * p1 * p1);
    }
}

Quad make_quad() {
    Quad q{};
    double p1[8], w1[8], p2[6], w2[6];
    lgl(8, p1, w1);
    lgl(6, p2, w2);
    const double phi0 = (120.0 / 180.0) * M_PI;
    double tot = 0.0;
    for (int i2 = 0; i2 < 6; ++i2)
        for (int i1 = 0; i1 < 8; ++i1) {
            const int i = i1 + 8 * i2;
            const double rho = 0.5 * (p1[i1] + 1.0 + 1.0e-06);
            const double phi = M_PI * (270.0 + 60.0 * p2[i2]) / 180.0;
            const double sc = 1.0 / std::fabs(2.0 * std::sin(phi));
            for (int c = 0; c < 3; ++c) {
                const double tx = std::cos(phi + c * phi0) * rho * sc;
                const double ty = std::sin(phi + c * phi0) * rho * sc;
                q.ix[i + 48 * c] = (1.0 - std::sqrt(3.0) * tx - ty) / 3.0;
                q.iy[i + 48 * c] = (1.0 + std::sqrt(3.0) * tx - ty) / 3.0;
                q.iw[i + 48 * c] = w1[i1] * w2[i2] * rho * sc * sc;
                tot += q.iw[i + 48 * c];
            }
        }
    for (double &w : q.iw) w /= tot;
    // Strang & Fix 28-point rule: centroid, five 3-point orbits, two 6-point orbits
    const double s3[5][3] = {{0.9480217181434233, 0.02598914092828833, 0.008744311553736190},
                             {0.8114249947041546, 0.09428750264792270, 0.03808157199393533},
                             {0.01072644996557060, 0.4946367750172147, 0.01885544805613125},
                             {0.5853132347709715, 0.2073433826145142, 0.07215969754474100},
                             {0.1221843885990187, 0.4389078057004907, 0.06932913870553720}};
    const double s6[2][4] = {{0.6779376548825902, 0.04484167758913055, 0.27722066752827925, 0.04105631542928860},
                             {0.8588702812826364, 0.0, 0.1411297187173636, 0.007362383783300573}};
    int k = 0;
    q.xt[k] = q.yt[k] = 1.0 / 3.0;
    q.wt[k++] = 0.08797730116222190;
    for (int o = 0; o < 5; ++o) {
        const double a = s3[o][0], b = s3[o][1], w = s3[o][2];
        const double px[3] = {a, b, b}, py[3] = {b, a, b};
        for (int p = 0; p < 3; ++p) {
            q.xt[k] = px[p];
            q.yt[k] = py[p];
            q.wt[k++] = w;
        }
    }
    for (int o = 0; o < 2; ++o) {
        const double a = s6[o][0], b = s6[o][1], c = s6[o][2], w = s6[o][3];
        const double px[6] = {a, a, b, b, c, c}, py[6] = {b, c, a, c, a, b};
        for (int p = 0; p < 6; ++p) {
            q.xt[k] = px[p];
            q.yt[k] = py[p];
            q.wt[k++] = w;
        }
    }
    return q;
}

struct MeshDev {
    const double *coord, *normal, *t1, *t2, *area, *verts, *radius;
};

__device__ __forceinline__ double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// per-tessera maximum distance centroid -> polar quadrature point
__global__ void k_polar_dmax(int n, MeshDev m, const Quad *__restrict__ q, double *__restrict__ dmax) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *v = m.verts + 9 * (size_t)i;
    double best = 0.0;
    for (int k = 0; k < 144; ++k) {
        double d2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double p = q->ix[k] * v[c] + q->iy[k] * v[3 + c] + (1.0 - q->ix[k] - q->iy[k]) * v[6 + c];
            const double d = m.coord[3 * i + c] - p;
            d2 += d * d;
        }
        best = fmax(best, sqrt(d2));
    }
    dmax[i] = best;
}

// green_diagonal_accurate (refinement_diagonal_elements :662-737)
// The normal component sums (c_i - p).n / d^3 over points p in the plane of
// tessera i: exactly zero in exact arithmetic, so its value is rounding
// noise amplified by 1/d^3 (d ~ 1e-6 near the centroid).  The reference
// keeps it in the unrefined IEF diagonal; products are not fused here so
// that the noise is the one of plain IEEE multiply-then-add evaluation.
__global__ void k_polar_diag(int n, MeshDev m, const Quad *__restrict__ q, double clamp, double *__restrict__ gdiag) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *v = m.verts + 9 * (size_t)i;
    const double *ci = m.coord + 3 * i, *ni = m.normal + 3 * i, *t1 = m.t1 + 3 * i, *t2 = m.t2 + 3 * i;
    double gn = 0.0, g1 = 0.0, g2 = 0.0;
    for (int k = 0; k < 144; ++k) {
        double vec[3];
        for (int c = 0; c < 3; ++c) {
            const double p = q->ix[k] * v[c] + q->iy[k] * v[3 + c] + (1.0 - q->ix[k] - q->iy[k]) * v[6 + c];
            vec[c] = ci[c] - p;
        }
        double dk = sqrt(dot3(vec, vec));
        gn -= q->iw[k] * m.area[i] * dot3(vec, ni) / (dk * dk * dk);
        if (dk < 1.0e-04) dk = clamp;
        g1 -= q->iw[k] * m.area[i] * dot3(vec, t1) / (dk * dk * dk);
        g2 -= q->iw[k] * m.area[i] * dot3(vec, t2) / (dk * dk * dk);
    }
    for (int c = 0; c < 3; ++c) gdiag[3 * i + c] = ni[c] * gn + t1[c] * g1 + t2[c] * g2;
}

// D, D_ief, S A^-1 element by element.
__global__ void k_bem_d(int n, MeshDev m, const Quad *__restrict__ q, int accurate, double sphere_r,
                        const double *__restrict__ gdiag, int ief, double *__restrict__ D, double *__restrict__ Dief,
                        double *__restrict__ SA) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= n) return;
    const double *ci = m.coord + 3 * i, *cj = m.coord + 3 * j, *ni = m.normal + 3 * i;
    const double d[3] = {ci[0] - cj[0], ci[1] - cj[1], ci[2] - cj[2]};
    const double r = sqrt(dot3(d, d));
    double val;
    if (accurate) {
        double t1 = dot3(ci, ci) + dot3(cj, cj) - 2.0 * dot3(ci, cj);
        if (t1 < (double)1.0e-10f) t1 = 0.0;
        const double sel = sqrt(t1) - m.radius[j];
        if (sel >= 0.0) {
            val = m.area[j] * dot3(d, ni) / (r * r * r);
        } else if (i == j) {
            val = -dot3(gdiag + 3 * i, ni);
        } else {
            const double *v = m.verts + 9 * (size_t)j;
            double ox = 0.0, oy = 0.0, oz = 0.0;
            for (int k = 0; k < 28; ++k) {
                double p[3];
                for (int c = 0; c < 3; ++c)
                    p[c] = q->xt[k] * v[c] + q->yt[k] * v[3 + c] + (1.0 - q->xt[k] - q->yt[k]) * v[6 + c];
                const double tx = ci[0] - p[0], ty = ci[1] - p[1], tz = ci[2] - p[2];
                const double tn = sqrt(tx * tx + ty * ty + tz * tz);
                const double t3 = tn * tn * tn;
                ox -= q->wt[k] * m.area[j] * tx / t3;
                oy -= q->wt[k] * m.area[j] * ty / t3;
                oz -= q->wt[k] * m.area[j] * tz / t3;
            }
            val = -(ox * ni[0] + oy * ni[1] + oz * ni[2]);
        }
    } else {
        val = i != j ? m.area[j] * dot3(d, ni) / (r * r * r) : kDiagParam * sqrt(4.0 * M_PI * m.area[i]) / (2.0 * sphere_r);
    }
    const size_t o = i + (size_t)n * j;
    D[o] = val;
    if (ief) {
        Dief[o] = i == j ? val - 2.0 * M_PI : val;
        SA[o] = i == j ? kDiagParam * sqrt(4.0 * M_PI * m.area[i]) * (1.0 / m.area[i])
                       : (m.area[j] / r) * (1.0 / m.area[j]);
    }
}

// Sum rule on the diagonal (final_refinement_green_function :796-838).
__global__ __launch_bounds__(256) void k_sum_rule(int n, MeshDev m, const double *__restrict__ gdiag,
                                                  double *__restrict__ D) {
    const int i = blockIdx.x;
    __shared__ double sh[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) s += (D[k + (size_t)n * i] / m.area[i]) * m.area[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double *ni = m.normal + 3 * i;
        double g[3];
        for (int c = 0; c < 3; ++c) g[c] = ni[c] * (-2.0 * M_PI + sh[0]) + gdiag[3 * i + c];
        D[i + (size_t)n * i] = -dot3(g, ni);
    }
}

__global__ void k_to_f32(size_t n, const double *__restrict__ a, float *__restrict__ b) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// y_s = Mat x_s (real n x n matrix, complex vectors), optional FP32 matrix.
template <typename T>
__global__ void k_real_matvec(int n, const T *__restrict__ Mat, const int *__restrict__ slots,
                              const double2 *__restrict__ x, int ld, double2 *__restrict__ y, int ldy, int y_by_k) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const int s = slots[k];
    if (r >= n) return;
    const double2 *xs = x + (size_t)ld * s;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < n; ++c) {
        const double a = (double)Mat[r + (size_t)n * c];
        re += a * xs[c].x;
        im += a * xs[c].y;
    }
    y[(size_t)ldy * (y_by_k ? k : s) + r] = make_double2(re, im);
}

// DPCM: y = Dx + lambda x ;  IEF (second stage): y = -(D u + lambda u), u = S A^-1 x
__global__ void k_bem_shift(int n, const int *__restrict__ slots, const double2 *__restrict__ lam,
                            const double2 *__restrict__ dx, int lddx, int dx_by_k, const double2 *__restrict__ u,
                            int ldu, int u_by_k, double sign, double2 *__restrict__ y, int ld) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const int s = slots[k];
    if (r >= n) return;
    const double2 a = dx[(size_t)lddx * (dx_by_k ? k : s) + r];
    const double2 b = cmul(lam[s], u[(size_t)ldu * (u_by_k ? k : s) + r]);
    y[(size_t)ld * s + r] = make_double2(sign * (a.x + b.x), sign * (a.y + b.y));
}

__global__ void k_bem_rhs_dpcm(int n, const double *__restrict__ normal, double e0, double2 *__restrict__ rhs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; ++k) rhs[(size_t)n * k + i] = make_double2(-normal[3 * i + k] * e0, 0.0);
}

__global__ void k_bem_rhs_ief(int n, const double *__restrict__ Dief, const double *__restrict__ coord, double e0,
                              double2 *__restrict__ rhs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s[3] = {0, 0, 0};
    for (int j = 0; j < n; ++j) {
        const double d = Dief[i + (size_t)n * j];
        for (int k = 0; k < 3; ++k) s[k] += d * (-e0 * coord[3 * j + k]);
    }
    for (int k = 0; k < 3; ++k) rhs[(size_t)n * k + i] = make_double2(s[k], 0.0);
}

}  // namespace

void build_mesh_device(pxh_ctx *c) {
    Mesh &M = c->mesh;
    const int n = M.n;
    hipStream_t st = c->stream;
    // tesserae (bem_module.f90:420-491), host side: O(n)
    M.h_coord.assign(3 * (size_t)n, 0.0);
    M.h_normal.assign(3 * (size_t)n, 0.0);
    M.h_t1.assign(3 * (size_t)n, 0.0);
    M.h_t2.assign(3 * (size_t)n, 0.0);
    M.h_area.assign(n, 0.0);
    std::vector<double> radius(n);
    for (int i = 0; i < n; ++i) {
        const double *p1 = &M.h_verts[9 * (size_t)i], *p2 = p1 + 3, *p3 = p1 + 6;
        double a[3], b[3], cr[3];
        for (int k = 0; k < 3; ++k) {
            M.h_coord[3 * i + k] = (p1[k] + p2[k] + p3[k]) / 3.0;
            a[k] = p2[k] - p1[k];
            b[k] = p3[k] - p1[k];
        }
        const double na = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        for (int k = 0; k < 3; ++k) M.h_t1[3 * i + k] = -a[k] / na;
        cr[0] = a[1] * b[2] - a[2] * b[1];
        cr[1] = a[2] * b[0] - a[0] * b[2];
        cr[2] = a[0] * b[1] - a[1] * b[0];
        const double nc = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        M.h_area[i] = nc / 2.0;
        for (int k = 0; k < 3; ++k) M.h_normal[3 * i + k] = M.normal_factor * cr[k] / nc;
        const double *nn = &M.h_normal[3 * i], *t1 = &M.h_t1[3 * i];
        M.h_t2[3 * i + 0] = nn[1] * t1[2] - nn[2] * t1[1];
        M.h_t2[3 * i + 1] = nn[2] * t1[0] - nn[0] * t1[2];
        M.h_t2[3 * i + 2] = nn[0] * t1[1] - nn[1] * t1[0];
        double rmax = 0.0;
        for (int v = 0; v < 3; ++v) {
            double d2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double d = M.h_coord[3 * i + k] - p1[3 * v + k];
                d2 += d * d;
            }
            rmax = std::max(rmax, std::sqrt(d2));
        }
        radius[i] = rmax;
    }
    DevBuf<double> coord, normal, t1, t2, area, verts, rad, gdiag, dmax;
    coord.upload(M.h_coord.data(), M.h_coord.size(), st);
    normal.upload(M.h_normal.data(), M.h_normal.size(), st);
    t1.upload(M.h_t1.data(), M.h_t1.size(), st);
    t2.upload(M.h_t2.data(), M.h_t2.size(), st);
    area.upload(M.h_area.data(), M.h_area.size(), st);
    verts.upload(M.h_verts.data(), M.h_verts.size(), st);
    rad.upload(radius.data(), radius.size(), st);
    Quad hq = make_quad();
    DevBuf<Quad> q;
    q.upload(&hq, 1, st);
    MeshDev md{coord.p, normal.p, t1.p, t2.p, area.p, verts.p, rad.p};
    const bool acc = M.green == PXH_ACCURATE;
    const bool ief = M.variant == PXH_IEFPCM;
    gdiag.alloc(3 * (size_t)n);
    if (acc) {
        dmax.alloc(n);
        k_polar_dmax<<<(n + 127) / 128, 128, 0, st>>>(n, md, q.p, dmax.p);
        std::vector<double> hd(n);
        PXH_HIP(hipMemcpyAsync(hd.data(), dmax.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        PXH_HIP(hipStreamSynchronize(st));
        const double clamp = *std::max_element(hd.begin(), hd.end()) * 1.0e-4;
        k_polar_diag<<<(n + 127) / 128, 128, 0, st>>>(n, md, q.p, clamp, gdiag.p);
    }
    const size_t nn = (size_t)n * n;
    M.D.alloc(nn);
    if (ief) {
        M.Dief.alloc(nn);
        M.SA.alloc(nn);
    }
    k_bem_d<<<dim3((n + 127) / 128, n), 128, 0, st>>>(n, md, q.p, acc ? 1 : 0, M.sphere_r, gdiag.p, ief ? 1 : 0, M.D.p,
                                                     ief ? M.Dief.p : nullptr, ief ? M.SA.p : nullptr);
    if (acc) k_sum_rule<<<n, 256, 0, st>>>(n, md, gdiag.p, M.D.p);
    M.Df.alloc(nn);
    k_to_f32<<<(unsigned)((nn + 255) / 256), 256, 0, st>>>(nn, M.D.p, M.Df.p);
    if (ief) {
        M.SAf.alloc(nn);
        k_to_f32<<<(unsigned)((nn + 255) / 256), 256, 0, st>>>(nn, M.SA.p, M.SAf.p);
    }
    PXH_HIP(hipGetLastError());
    PXH_HIP(hipStreamSynchronize(st));
    M.built = true;
}

void apply_bem(pxh_ctx *c, int nsys, const int *d_sys_list, const double2 *d_lambda, const double2 *x, double2 *y,
               int ld, bool fp32) {
    if (nsys <= 0) return;
    Mesh &M = c->mesh;
    const int n = M.n;
    hipStream_t st = c->stream;
    c->work_part.alloc((size_t)nsys * n);
    c->work_t.alloc((size_t)nsys * n);
    dim3 g((n + 255) / 256, nsys);
    if (M.variant == PXH_DPCM) {
        if (fp32)
            k_real_matvec<float><<<g, 256, 0, st>>>(n, M.Df.p, d_sys_list, x, ld, c->work_part.p, n, 1);
        else
            k_real_matvec<double><<<g, 256, 0, st>>>(n, M.D.p, d_sys_list, x, ld, c->work_part.p, n, 1);
        k_bem_shift<<<g, 256, 0, st>>>(n, d_sys_list, d_lambda, c->work_part.p, n, 1, x, ld, 0, 1.0, y, ld);
    } else {
        // u = S A^-1 x, then y = -(D u + lambda u)
        if (fp32)
            k_real_matvec<float><<<g, 256, 0, st>>>(n, M.SAf.p, d_sys_list, x, ld, c->work_t.p, n, 1);
        else
            k_real_matvec<double><<<g, 256, 0, st>>>(n, M.SA.p, d_sys_list, x, ld, c->work_t.p, n, 1);
        // D u: u is indexed by k (position in the list); use an identity list
        std::vector<int> ident(nsys);
        for (int k = 0; k < nsys; ++k) ident[k] = k;
        DevBuf<int> id;
        id.upload(ident.data(), nsys, st);
        if (fp32)
            k_real_matvec<float><<<g, 256, 0, st>>>(n, M.Df.p, id.p, c->work_t.p, n, c->work_part.p, n, 1);
        else
            k_real_matvec<double><<<g, 256, 0, st>>>(n, M.D.p, id.p, c->work_t.p, n, c->work_part.p, n, 1);
        k_bem_shift<<<g, 256, 0, st>>>(n, d_sys_list, d_lambda, c->work_part.p, n, 1, c->work_t.p, n, 1, -1.0, y, ld);
        PXH_HIP(hipStreamSynchronize(st));
    }
    PXH_HIP(hipGetLastError());
    c->stats.v[6] += (double)n * n * nsys;
}

void dynamic_rhs_bem(pxh_ctx *c, double e0, double2 *d_rhs) {
    Mesh &M = c->mesh;
    const int n = M.n;
    hipStream_t st = c->stream;
    if (M.variant == PXH_DPCM) {
        DevBuf<double> normal;
        normal.upload(M.h_normal.data(), M.h_normal.size(), st);
        k_bem_rhs_dpcm<<<(n + 255) / 256, 256, 0, st>>>(n, normal.p, e0, d_rhs);
        PXH_HIP(hipStreamSynchronize(st));
    } else {
        DevBuf<double> coord;
        coord.upload(M.h_coord.data(), M.h_coord.size(), st);
        k_bem_rhs_ief<<<(n + 255) / 256, 256, 0, st>>>(n, M.Dief.p, coord.p, e0, d_rhs);
        PXH_HIP(hipStreamSynchronize(st));
    }
    PXH_HIP(hipGetLastError());
}

}  // namespace pxh

namespace pxh {

namespace {
// dinv = 1 / (D_ii + lambda) (DPCM diagonal)
__global__ void k_precond_bem(int n, const double *__restrict__ D, const double2 *__restrict__ lam,
                              double2 *__restrict__ dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 d = make_double2(D[i + (size_t)n * i] + lam->x, lam->y);
    const double den = d.x * d.x + d.y * d.y;
    dinv[i] = make_double2(d.x / den, -d.y / den);
}
}  // namespace

void precond_diag_bem(pxh_ctx *c, const double2 *d_lambda_one, double2 *dinv) {
    const int n = c->mesh.n;
    k_precond_bem<<<(n + 255) / 256, 256, 0, c->stream>>>(n, c->mesh.D.p, d_lambda_one, dinv);
    PXH_HIP(hipGetLastError());
}

}  // namespace pxh
