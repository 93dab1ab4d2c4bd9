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

// Operator products Y = Mat X on the matrix cores.  Mat is a real n x n
// column-major matrix (FP64, or FP32 widened on load); X holds nsys complex
// vectors, i.e. 2 nsys real columns.  The arithmetic is FP64 throughout
// (v_mfma_f64_16x16x4_f64, FP64 accumulation).
//
// X is first packed by k_gemm_pack into tiles of 16 systems (32 real
// columns): Xp[tile][c][j][h] with real column h*16 + j, so one 16-byte load
// gives a lane both of its B operands.  Rows c >= n are zero up to cpad.
// A workgroup computes a 32-row x 32-column tile of Y; its four waves split
// the columns of Mat (K) in equal ranges and sum through LDS at the end.
// Each wave holds 2 x 2 16x16 accumulators: row fragment h covers tile rows
// 2i + h (so a lane reads two adjacent rows of one matrix column), column
// fragment h covers real columns h*16 + j.
constexpr int kGemmRows = 32;
constexpr int kGemmSys = 16;

typedef double d4v __attribute__((ext_vector_type(4)));

__global__ void k_gemm_pack(int n, int cpad, int nsys, const int *__restrict__ slots, const double2 *__restrict__ x,
                            int ld, int x_by_k, int ntiles, double *__restrict__ Xp) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)ntiles * cpad * kGemmSys) return;
    // one work-item per (tile, row c, system within the tile)
    const int q = (int)(e % kGemmSys);
    const size_t tc = e / kGemmSys;
    const int c = (int)(tc % cpad);
    const int t = (int)(tc / cpad);
    const int k = t * kGemmSys + q;
    double2 v = make_double2(0.0, 0.0);
    if (k < nsys && c < n) v = x[(size_t)ld * (x_by_k ? k : slots[k]) + c];
    // real columns 2q (re) and 2q+1 (im) -> (j, h) = (col % 16, col / 16)
    double *row = Xp + ((size_t)t * cpad + c) * 32;
    const int c0 = 2 * q, c1 = 2 * q + 1;
    row[(c0 % 16) * 2 + c0 / 16] = v.x;
    row[(c1 % 16) * 2 + c1 / 16] = v.y;
}

struct GemmOut {
    const double2 *lam;  // per slot; null: no shift
    const double2 *u;    // shifted vectors (lam u), ldu, by k or by slot
    int ldu, u_by_k;
    double sign;
    double2 *y;
    int ldy, y_by_k;
};

// Y tile = sign * (Mat X + lam u) for the listed systems.  Grid: 8 * per
// workgroups, per = ceil(tiles / 8); workgroup b takes tile
// (b % 8) * per + b / 8, so each XCD (b % 8) works on a contiguous range of
// row tiles and reads their matrix rows through its own L2 for every column
// tile.
//
// With ksplit > 1 the columns of Mat are also split over ksplit workgroups
// per tile (products of few systems, which would leave the chip idle); each
// writes its raw sums to part[(ks coltiles + ct) rowpad + row][32] and
// k_gemm_reduce adds them in a fixed order and applies the epilogue.
template <typename T>
__global__ __launch_bounds__(256) void k_bem_gemm(int n, int cpad, int steps_per_wave, int rowtiles, int coltiles,
                                                  int ksplit, const T *__restrict__ Mat,
                                                  const double *__restrict__ Xp, const int *__restrict__ slots,
                                                  int nsys, GemmOut o, double *__restrict__ part) {
    __shared__ double red[4][16 * 64];
    const int total = rowtiles * coltiles * ksplit;
    const int per = (total + 7) / 8;
    const int L = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (L >= total) return;
    const int ks = L % ksplit;
    const int rt = L / ksplit / coltiles, ct = (L / ksplit) % coltiles;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = lane & 15, kk = lane >> 4;
    const int r0 = rt * kGemmRows;
    const int ra = min(r0 + 2 * i, n - 1), rb = min(r0 + 2 * i + 1, n - 1);
    const int s0 = (ks * 4 + w) * steps_per_wave;
    const double *xb = Xp + (size_t)ct * cpad * 32 + 2 * i;
    d4v acc00 = {0, 0, 0, 0}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
    // four K steps (16 matrix columns) per iteration, the next four loaded
    // while the current ones are multiplied
    double a0[4], a1[4];
    double2 bv[4];
    auto load = [&](int s, double *pa0, double *pa1, double2 *pb) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = 4 * (s + u) + kk;
            const T *mc = Mat + (size_t)n * min(c, n - 1);
            pa0[u] = (double)mc[ra];
            pa1[u] = (double)mc[rb];
            pb[u] = *reinterpret_cast<const double2 *>(xb + (size_t)c * 32);
        }
    };
    load(s0, a0, a1, bv);
    for (int s = s0; s < s0 + steps_per_wave; s += 4) {
        double na0[4], na1[4];
        double2 nb[4];
        if (s + 4 < s0 + steps_per_wave) load(s + 4, na0, na1, nb);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], bv[u].x, acc00, 0, 0, 0);
            acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], bv[u].y, acc01, 0, 0, 0);
            acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], bv[u].x, acc10, 0, 0, 0);
            acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], bv[u].y, acc11, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a0[u] = na0[u];
            a1[u] = na1[u];
            bv[u] = nb[u];
        }
    }
    // accumulator element v of a lane (f64 16x16 layout): fragment row
    // lane / 16 + 4 v, fragment column lane % 16
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        red[w][(0 * 4 + v) * 64 + lane] = acc00[v];
        red[w][(1 * 4 + v) * 64 + lane] = acc01[v];
        red[w][(2 * 4 + v) * 64 + lane] = acc10[v];
        red[w][(3 * 4 + v) * 64 + lane] = acc11[v];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = threadIdx.x + 256 * q;
        const double sum = red[0][e] + red[1][e] + red[2][e] + red[3][e];
        const int l = e & 63, v = (e >> 6) & 3, fc = (e >> 8) & 1, fr = e >> 9;
        const int r = r0 + 2 * ((l >> 4) + 4 * v) + fr;
        const int col = fc * 16 + (l & 15);
        if (ksplit > 1) {
            part[(((size_t)ks * coltiles + ct) * rowtiles * kGemmRows + r) * 32 + col] = sum;
            continue;
        }
        const int k = ct * kGemmSys + col / 2, part = col & 1;
        if (r >= n || k >= nsys) continue;
        const int slot = slots[k];
        double val = sum;
        if (o.lam) {
            const double2 l2 = o.lam[slot];
            const double2 uu = o.u[(size_t)o.ldu * (o.u_by_k ? k : slot) + r];
            val += part ? l2.x * uu.y + l2.y * uu.x : l2.x * uu.x - l2.y * uu.y;
        }
        reinterpret_cast<double *>(o.y)[2 * ((size_t)o.ldy * (o.y_by_k ? k : slot) + r) + part] = o.sign * val;
    }
}

// Y = M X for up to kGemvSys systems (the multi-shift spectra: one vector per
// field direction): a streaming matrix-vector pass on the FP64 VALU, as the
// few columns leave the matrix cores nothing to gain and their padded
// 16-system tiles cost MFMA time that does not overlap the loads.  One
// work-item per row (pair), the columns split over blockIdx.y (their slice of X
// staged in LDS, read as broadcasts), row pairs by 16-byte loads when n is
// even; partials in k_gemm_reduce's layout
// (one column tile), which adds them in order and applies the epilogue.
constexpr int kGemvSys = 4;
constexpr int kGemvMaxCols = 256;  // columns per workgroup (LDS slice of X)

template <int RPT>
__global__ __launch_bounds__(256) void k_bem_gemv(int n, int cols_per, const double *__restrict__ Mat,
                                                  const double2 *__restrict__ x, int ldx, int x_by_k,
                                                  const int *__restrict__ slots, int nsys, int rowpad,
                                                  double *__restrict__ part) {
    __shared__ double2 xs[kGemvSys][kGemvMaxCols];
    const int r = RPT * (blockIdx.x * 256 + threadIdx.x);  // first of the work-item's RPT rows
    const int ks = blockIdx.y;
    const int c0 = ks * cols_per, c1 = min(n, c0 + cols_per);
    for (int e = threadIdx.x; e < kGemvSys * cols_per; e += 256) {
        const int k = e / cols_per, cc = e % cols_per;
        const int cidx = c0 + cc;
        xs[k][cc] = (k < nsys && cidx < n) ? x[(size_t)(x_by_k ? k : slots[k]) * ldx + cidx] : make_double2(0, 0);
    }
    __syncthreads();
    if (r >= n) return;
    double re[RPT][kGemvSys], im[RPT][kGemvSys];
#pragma unroll
    for (int q = 0; q < RPT; ++q)
#pragma unroll
        for (int k = 0; k < kGemvSys; ++k) re[q][k] = im[q][k] = 0.0;
    const double *m = Mat + r;
    // one column's RPT rows: a 16-byte load for row pairs (n even, r even)
    auto ld = [&](int cl, double (&a)[RPT]) {
        if constexpr (RPT == 2) {
            const double2 v = *reinterpret_cast<const double2 *>(m + (size_t)n * cl);
            a[0] = v.x;
            a[1] = v.y;
        } else {
            a[0] = m[(size_t)n * cl];
        }
    };
    int col = c0;
    for (; col + 8 <= c1; col += 8) {
        double a[8][RPT];
#pragma unroll
        for (int u = 0; u < 8; ++u) ld(col + u, a[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int k = 0; k < kGemvSys; ++k) {
                const double2 xv = xs[k][col - c0 + u];
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    re[q][k] = fma(a[u][q], xv.x, re[q][k]);
                    im[q][k] = fma(a[u][q], xv.y, im[q][k]);
                }
            }
    }
    for (; col < c1; ++col) {
        double a[RPT];
        ld(col, a);
#pragma unroll
        for (int k = 0; k < kGemvSys; ++k) {
            const double2 xv = xs[k][col - c0];
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                re[q][k] = fma(a[q], xv.x, re[q][k]);
                im[q][k] = fma(a[q], xv.y, im[q][k]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        double *pr = part + ((size_t)ks * rowpad + r + q) * 32;
#pragma unroll
        for (int k = 0; k < kGemvSys; ++k) {
            if (k >= nsys) break;
            pr[2 * k] = re[q][k];
            pr[2 * k + 1] = im[q][k];
        }
    }
}

// Sum of the ksplit partials of each (row, real column) in order, then the
// epilogue of k_bem_gemm.  One work-item per (column tile, row, real column).
__global__ void k_gemm_reduce(int n, int rowpad, int coltiles, int ksplit, const double *__restrict__ part,
                              const int *__restrict__ slots, int nsys, GemmOut o) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)coltiles * n * 32) return;
    const int col = (int)(e % 32);
    const int r = (int)((e / 32) % n);
    const int ct = (int)(e / 32 / n);
    const int k = ct * kGemmSys + col / 2, pt = col & 1;
    if (k >= nsys) return;
    double val = 0.0;
    for (int ks = 0; ks < ksplit; ++ks) val += part[(((size_t)ks * coltiles + ct) * rowpad + r) * 32 + col];
    const int slot = slots[k];
    if (o.lam) {
        const double2 l2 = o.lam[slot];
        const double2 uu = o.u[(size_t)o.ldu * (o.u_by_k ? k : slot) + r];
        val += pt ? l2.x * uu.y + l2.y * uu.x : l2.x * uu.x - l2.y * uu.y;
    }
    reinterpret_cast<double *>(o.y)[2 * ((size_t)o.ldy * (o.y_by_k ? k : slot) + r) + pt] = o.sign * val;
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
    M.ief_z_valid = false;
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
    const int coltiles = (nsys + kGemmSys - 1) / kGemmSys;
    const int rowtiles = (n + kGemmRows - 1) / kGemmRows;
    // K steps of 4 columns over ksplit workgroups x 4 waves, in multiples of
    // 4 per wave; ksplit brings the launch to ~1024 workgroups (at least 64
    // columns per wave), and the packed X is zero-padded to cover every range
    const int ksteps = (n + 3) / 4;
    int ksplit = std::max(1, std::min((1024 + rowtiles * coltiles - 1) / (rowtiles * coltiles), n / 256));
    ksplit = std::min(ksplit, 16);
    const int spw = (ksteps + 16 * ksplit - 1) / (16 * ksplit) * 4;
    const int cpad = 16 * spw * ksplit;
    c->work_pack.alloc((size_t)coltiles * cpad * 32);
    if (ksplit > 1) c->work_gemm.alloc((size_t)ksplit * coltiles * rowtiles * kGemmRows * 32);
    const size_t pack_items = (size_t)coltiles * cpad * kGemmSys;
    const unsigned pack_blocks = (unsigned)((pack_items + 255) / 256);
    const unsigned gemm_blocks = 8u * (unsigned)((rowtiles * coltiles * ksplit + 7) / 8);
    const unsigned red_blocks = (unsigned)(((size_t)coltiles * n * 32 + 255) / 256);
    // few systems in FP64: the streaming pass (k_bem_gemv), the columns split
    // so that the launch has >= ~640 workgroups (at least 16 columns each)
    // and a slice fits in LDS
    const bool gemv = !fp32 && nsys <= kGemvSys && c->opt.bem_gemv;
    // rows per work-item: row pairs (16-byte loads, n even) pay where the
    // matrix streams from HBM; one row where it stays in the last-level cache
    // (measured: n = 10240 201 against 213 us per product, n = 5120 73 against
    // 53-56 us; the MFMA tiles 204 and 60-62 us)
    const int vrpt = (n % 2 == 0 && n >= 8192) ? 2 : 1;
    const int vrows = (n + 256 * vrpt - 1) / (256 * vrpt);
    const int vsplit =
        std::max((n + kGemvMaxCols - 1) / kGemvMaxCols, std::min((640 + vrows - 1) / vrows, std::max(1, n / 16)));
    const int vcols = (n + vsplit - 1) / vsplit;
    const int vsplit_eff = (n + vcols - 1) / vcols;
    const int vrowpad = vrows * 256 * vrpt;
    if (gemv) c->work_gemm.alloc((size_t)vsplit_eff * vrowpad * 32);
    auto gemm = [&](const double *Md, const float *Mf, const double2 *xin, int x_by_k, const GemmOut &o) {
        if (gemv) {
            const int t_gemv = timer_begin(c);
            if (vrpt == 2)
                k_bem_gemv<2><<<dim3(vrows, vsplit_eff), 256, 0, st>>>(n, vcols, Md, xin, x_by_k ? n : ld, x_by_k,
                                                                       d_sys_list, nsys, vrowpad, c->work_gemm.p);
            else
                k_bem_gemv<1><<<dim3(vrows, vsplit_eff), 256, 0, st>>>(n, vcols, Md, xin, x_by_k ? n : ld, x_by_k,
                                                                       d_sys_list, nsys, vrowpad, c->work_gemm.p);
            k_gemm_reduce<<<(unsigned)(((size_t)n * 32 + 255) / 256), 256, 0, st>>>(
                n, vrowpad, 1, vsplit_eff, c->work_gemm.p, d_sys_list, nsys, o);
            timer_end(c, t_gemv, 16);
            if (!c->capturing) {
                c->stats.v[17] += 1;
                c->stats.v[18] += 4.0 * n * (double)n * nsys;
                c->stats.v[22] += (double)n * n * 8.0 + 32.0 * n * (double)nsys;
            }
            return;
        }
        k_gemm_pack<<<pack_blocks, 256, 0, st>>>(n, cpad, nsys, d_sys_list, xin, x_by_k ? n : ld, x_by_k, coltiles,
                                                 c->work_pack.p);
        double *part = ksplit > 1 ? c->work_gemm.p : nullptr;
        // the matrix product alone (with its split-K reduction), timed on the
        // context's stream; algorithmic work: 4 flop per matrix element and
        // system (a real matrix on a complex vector), the matrix read once
        // plus each system's vector read and written
        const int t_gemm = timer_begin(c);
        if (fp32)
            k_bem_gemm<float><<<gemm_blocks, 256, 0, st>>>(n, cpad, spw, rowtiles, coltiles, ksplit, Mf,
                                                           c->work_pack.p, d_sys_list, nsys, o, part);
        else
            k_bem_gemm<double><<<gemm_blocks, 256, 0, st>>>(n, cpad, spw, rowtiles, coltiles, ksplit, Md,
                                                            c->work_pack.p, d_sys_list, nsys, o, part);
        if (ksplit > 1)
            k_gemm_reduce<<<red_blocks, 256, 0, st>>>(n, rowtiles * kGemmRows, coltiles, ksplit, part, d_sys_list,
                                                      nsys, o);
        timer_end(c, t_gemm, 16);
        if (!c->capturing) {  // replays of a captured step are neither timed nor counted here
            c->stats.v[17] += 1;
            c->stats.v[18] += 4.0 * n * (double)n * nsys;
            c->stats.v[22] += (double)n * n * (fp32 ? 4.0 : 8.0) + 32.0 * n * (double)nsys;
        }
    };
    if (M.variant == PXH_DPCM) {
        // y = D x + lambda x
        gemm(M.D.p, M.Df.p, x, 0, GemmOut{d_lambda, x, ld, 0, 1.0, y, ld, 0});
    } else if (c->bem_op == 1) {
        // y = -(D x + lambda x)
        gemm(M.D.p, M.Df.p, x, 0, GemmOut{d_lambda, x, ld, 0, -1.0, y, ld, 0});
    } else if (c->bem_op == 2) {
        // y = S x
        gemm(M.SA.p, M.SAf.p, x, 0, GemmOut{nullptr, nullptr, 0, 0, 1.0, y, ld, 0});
    } else {
        // u = S A^-1 x, then y = -(D u + lambda u)
        c->work_t.alloc((size_t)nsys * n);
        gemm(M.SA.p, M.SAf.p, x, 0, GemmOut{nullptr, nullptr, 0, 0, 1.0, c->work_t.p, n, 1});
        gemm(M.D.p, M.Df.p, c->work_t.p, 1, GemmOut{d_lambda, c->work_t.p, n, 1, -1.0, y, ld, 0});
    }
    PXH_HIP(hipGetLastError());
    if (!c->capturing) c->stats.v[6] += (double)n * n * nsys;
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
