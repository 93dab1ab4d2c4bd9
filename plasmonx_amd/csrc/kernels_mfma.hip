// MFMA form of the omegaFQFmu T sweep for a block of systems (gfx950).
//
// t_k = T x_k for up to 16*NCB systems at once.  T (4N x 4N, atom-interleaved
// rows [q_i, mu_ix, mu_iy, mu_iz]) is the frequency-independent real
// interaction matrix of src/matrix/matrix_module.f90 (T_qq :58-115,
// T_qmu/T_muq :120-234, T_mumu :240-345); its 4 x 4 block for a pair (i, j)
// is generated on the fly in FP32 from the pair's distance vector and
// stored in LDS as fp16 hi + lo (value * 2^kTScale), then contracted with
// the block X (4N x 32*NCB real columns: re / im of each system, also fp16
// hi + lo with a power-of-two scale per column) by
// v_mfma_f32_32x32x16_f16: acc += Ahi Bhi + Ahi Blo + Alo Bhi, a product
// with a 22-bit significand, accumulated in FP32 within a source split and
// in FP64 across splits.
//
// Workgroup: 256 work-items (4 waves), 32 target atoms (128 rows); wave w
// owns rows [32w, 32w+32) for all 32*NCB columns.  Each source step stages
// 16 atoms (K = 64): every work-item evaluates 2 pair blocks into LDS, then
// each wave runs 4 k-steps x NCB column blocks x 3 MFMAs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "pxh_internal.h"

namespace pxh {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTScale = 12;        // T elements are stored as T * 2^12
constexpr int kXTarget = 13;       // columns are scaled so that max |x| * s lies in [2^12, 2^13)
constexpr double kTwoOverSqrtPi = 1.1283791670955126;
constexpr double kFourOverSqrtPi = 2.2567583341910252;
constexpr int kAtomsI = 32;        // target atoms per workgroup
constexpr int kAtomsJ = 16;        // source atoms per step (K = 64)
constexpr int kStepsPerFlush = 8;  // source steps between FP32 -> FP64 flushes

// Per-column scale 2^e with max|x_col| * 2^e in [2^(kXTarget-1), 2^kXTarget).
__global__ __launch_bounds__(256) void k_col_scale(int nv, int nsys, const int *__restrict__ sys_list,
                                                   const double2 *__restrict__ x, int ld, float *__restrict__ scale,
                                                   int ncols) {
    const int c = blockIdx.x;  // column index 0..ncols-1
    const int k = c / 2;
    __shared__ double sh[256];
    double m = 0.0;
    if (k < nsys) {
        const double2 *xs = x + (size_t)sys_list[k] * ld;
        for (int r = threadIdx.x; r < nv; r += 256) m = fmax(m, fabs((c & 1) ? xs[r].y : xs[r].x));
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float s = 0.f;
        if (sh[0] > 0.0) {
            int e;
            frexp(sh[0], &e);  // max = f 2^e, f in [0.5, 1)
            s = ldexpf(1.f, kXTarget - e);
        }
        scale[c] = s;
    }
}

// X in fp16 hi / lo, laid out as the MFMA B fragments of every source step:
// [step jb][k-step s][col block cb][lane][8 halves].
template <int NCB>
__global__ void k_prep_x(int n, int nsys, const int *__restrict__ sys_list, const double2 *__restrict__ x, int ld,
                         const float *__restrict__ scale, _Float16 *__restrict__ xh, _Float16 *__restrict__ xl) {
    constexpr int NC = 32 * NCB;
    const int jb = blockIdx.x;
    for (int e = threadIdx.x; e < 64 * NC; e += blockDim.x) {
        const int k = e / NC, col = e % NC;
        const int jl = k / 4, b = k % 4;
        const int j = jb * kAtomsJ + jl;
        const int sys = col / 2;
        float v = 0.f;
        if (j < n && sys < nsys) {
            const double2 xv = x[(size_t)sys_list[sys] * ld + (b == 0 ? j : n + 3 * j + b - 1)];
            v = (float)((col & 1) ? xv.y : xv.x) * scale[col];
        }
        const _Float16 hi = (_Float16)v;
        const _Float16 lo = (_Float16)(v - (float)hi);
        const int s = k / 16, kk = k % 16, cb = col / 32, nl = col % 32;
        const int lane = (kk / 8) * 32 + nl;
        const size_t off = ((((size_t)jb * 4 + s) * NCB + cb) * 64 + lane) * 8 + (kk % 8);
        xh[off] = hi;
        xl[off] = lo;
    }
}

template <int NCB>
__global__ __launch_bounds__(256) void k_tsweep_mfma(int n, int nt, int hetero, const double4 *__restrict__ pos,
                                                     const PairConst *__restrict__ pc, int nsys,
                                                     const float *__restrict__ scale,
                                                     const _Float16 *__restrict__ xh, const _Float16 *__restrict__ xl,
                                                     double2 *__restrict__ part, int nv, int jsteps_per_split,
                                                     int nsys_total_in_part, int sys_offset) {
    constexpr int NC = 32 * NCB;
    __shared__ half8 sAh[4 * 4 * 64];  // [wave][k-step][lane] (16 KB)
    __shared__ half8 sAl[4 * 4 * 64];
    __shared__ half8 sBh[4 * NCB * 64];  // [k-step][cb][lane]
    __shared__ half8 sBl[4 * NCB * 64];
    __shared__ float spc[kMaxTypes * kMaxTypes][4];
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const int i0 = blockIdx.x * kAtomsI;
    const int split = blockIdx.y;
    const int jb0 = split * jsteps_per_split;
    const int jb_end = min((n + kAtomsJ - 1) / kAtomsJ, jb0 + jsteps_per_split);
    for (int e = tid; e < nt * nt; e += 256) {
        spc[e][0] = (float)pc[e].inv_rqq;
        spc[e][1] = (float)pc[e].inv_rqmu;
        spc[e][2] = (float)pc[e].inv_rmumu;
        spc[e][3] = (float)pc[e].eta_bar;
    }
    // the pair blocks this work-item evaluates: target il, sources jl0 and jl0 + 8
    const int il = tid % 32;
    const int jl0 = tid / 32;
    const int gi = i0 + il;
    const double4 pi = gi < n ? pos[gi] : make_double4(0, 0, 0, 0);
    const int ti = (int)pi.w;
    f32x16 acc[NCB];
    double accd[NCB][16];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            acc[cb][v] = 0.f;
            accd[cb][v] = 0.0;
        }
    const float tsc = ldexpf(1.f, kTScale);
    int since_flush = 0;
    for (int jb = jb0; jb < jb_end; ++jb) {
        __syncthreads();
        // ---- stage X fragments of this source step (contiguous in HBM)
        {
            const size_t base = (size_t)jb * 4 * NCB * 64;
            for (int e = tid; e < 4 * NCB * 64; e += 256) {
                sBh[e] = *reinterpret_cast<const half8 *>(xh + (base + e) * 8);
                sBl[e] = *reinterpret_cast<const half8 *>(xl + (base + e) * 8);
            }
        }
        // ---- evaluate two 4 x 4 pair blocks into the A fragments
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int jl = jl0 + 8 * p;
            const int gj = jb * kAtomsJ + jl;
            float blk[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) blk[a][b] = 0.f;
            if (gi < n && gj < n && gi != gj) {
                const double4 pj = pos[gj];
                const int tj = (int)pj.w;
                const float dx = (float)(pi.x - pj.x), dy = (float)(pi.y - pj.y), dz = (float)(pi.z - pj.z);
                const float r2 = dx * dx + dy * dy + dz * dz;
                const float inv_r = rsqrtf(r2);
                const float r = r2 * inv_r;
                const float *pp = spc[ti + nt * tj];
                const float tqq = erff(r * pp[0]) * inv_r;
                const float inv_r2 = inv_r * inv_r, inv_r3 = inv_r2 * inv_r;
                const float a1 = r * pp[1];
                const float g1 = (erff(a1) - (float)kTwoOverSqrtPi * a1 * __expf(-a1 * a1)) * inv_r3;
                float g2 = g1;
                if (hetero) {
                    const float a2 = r * spc[tj + nt * ti][1];
                    g2 = (erff(a2) - (float)kTwoOverSqrtPi * a2 * __expf(-a2 * a2)) * inv_r3;
                }
                const float a3 = r * pp[2];
                const float e3 = __expf(-a3 * a3);
                const float fd1 = erff(a3) - (float)kTwoOverSqrtPi * a3 * e3;
                const float fd2 = (float)kFourOverSqrtPi * pp[2] * pp[2] * pp[2] * e3;
                const float c0 = fd1 * inv_r3;
                const float c1 = fd2 * inv_r2 - 3.f * fd1 * inv_r3 * inv_r2;
                const float d[3] = {dx, dy, dz};
                blk[0][0] = tqq;
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    blk[0][1 + b] = g1 * d[b];   // T_qmu(i, j)
                    blk[1 + b][0] = -g2 * d[b];  // T_muq(i, j) = T_qmu(j, i)
#pragma unroll
                    for (int a = 0; a < 3; ++a) blk[1 + a][1 + b] = c1 * d[a] * d[b] + (a == b ? c0 : 0.f);
                }
            }
            // rows 4 il + a, columns 4 jl + b of the 128 x 64 A tile
            const int s = jl / 4, h0 = 4 * (jl % 4) % 8, khalf = (4 * (jl % 4)) / 8;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int r = 4 * il + a;
                const int w = r / 32, lr = (khalf * 32) + (r % 32);
                half4 hv, lv;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const float v = blk[a][b] * tsc;
                    const _Float16 hi = (_Float16)v;
                    hv[b] = hi;
                    lv[b] = (_Float16)(v - (float)hi);
                }
                const int idx = (w * 4 + s) * 64 + lr;
                *reinterpret_cast<half4 *>(reinterpret_cast<_Float16 *>(&sAh[idx]) + h0) = hv;
                *reinterpret_cast<half4 *>(reinterpret_cast<_Float16 *>(&sAl[idx]) + h0) = lv;
            }
        }
        __syncthreads();
        // ---- contraction: 4 k-steps x NCB column blocks x 3 products
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const half8 ah = sAh[(wave * 4 + s) * 64 + lane];
            const half8 al = sAl[(wave * 4 + s) * 64 + lane];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const half8 bh = sBh[(s * NCB + cb) * 64 + lane];
                const half8 bl = sBl[(s * NCB + cb) * 64 + lane];
                acc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[cb], 0, 0, 0);
                acc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[cb], 0, 0, 0);
                acc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[cb], 0, 0, 0);
            }
        }
        if (++since_flush == kStepsPerFlush) {
            since_flush = 0;
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    accd[cb][v] += (double)acc[cb][v];
                    acc[cb][v] = 0.f;
                }
        }
    }
    // ---- write partials: acc[cb][v] -> row m = (v/4)*8 + (lane/32)*4 + v%4 of the wave, column cb*32 + lane%32
    const double tinv = ldexp(1.0, -kTScale);
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const int col = cb * 32 + (lane % 32);
        const int sys = col / 2;
        if (sys >= nsys) continue;
        const double cs = scale[col] > 0.f ? tinv / (double)scale[col] : 0.0;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int m = (v / 4) * 8 + (lane / 32) * 4 + (v % 4);
            const int rr = wave * 32 + m;  // row in the 128-row tile
            const int i = i0 + rr / 4, a = rr % 4;
            if (i >= n) continue;
            const int row = a == 0 ? i : n + 3 * i + a - 1;
            const double val = (accd[cb][v] + (double)acc[cb][v]) * cs;
            double *dst = reinterpret_cast<double *>(
                part + ((size_t)split * nsys_total_in_part + sys_offset + sys) * nv + row);
            dst[col & 1] = val;
        }
    }
}

}  // namespace

void finish_product(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x, double2 *y,
                    int ld, int splits);

// MFMA path of the FP32-class product for omegaFQFmu (Gaussian kernel).
void apply_atoms_mfma(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x,
                      double2 *y, int ld) {
    constexpr int NCB = 3;  // 48 systems per pass; LDS 57 KB per workgroup
    constexpr int SYS_PER_GROUP = 16 * NCB;
    Atoms &A = c->atoms;
    const int n = A.n, nv = A.n_var();
    const int jsteps = (n + kAtomsJ - 1) / kAtomsJ;
    const int iblocks = (n + kAtomsI - 1) / kAtomsI;
    const int groups = (nsys + SYS_PER_GROUP - 1) / SYS_PER_GROUP;
    // splits so that a launch has >= ~1500 workgroups, each split >= 16 source steps
    int splits = c->opt.splits > 0 ? c->opt.splits : std::max(1, 1536 / std::max(1, iblocks));
    splits = std::min(splits, std::max(1, jsteps / 16));
    const int per = (jsteps + splits - 1) / splits;
    splits = (jsteps + per - 1) / per;
    c->work_part.alloc((size_t)splits * nsys * nv);
    c->work_t.alloc((size_t)nsys * nv);
    const size_t xelems = (size_t)jsteps * 4 * NCB * 64 * 8;
    c->work_xh.alloc(xelems);
    c->work_xl.alloc(xelems);
    c->work_scale.alloc(32 * NCB);
    PXH_HIP(hipEventRecord(c->ev0, c->stream));
    for (int g = 0; g < groups; ++g) {
        const int ns = std::min(SYS_PER_GROUP, nsys - g * SYS_PER_GROUP);
        const int *lst = d_sys_list + g * SYS_PER_GROUP;
        k_col_scale<<<32 * NCB, 256, 0, c->stream>>>(nv, ns, lst, x, ld, c->work_scale.p, 32 * NCB);
        k_prep_x<NCB><<<jsteps, 256, 0, c->stream>>>(n, ns, lst, x, ld, c->work_scale.p, c->work_xh.p,
                                                     c->work_xl.p);
        k_tsweep_mfma<NCB><<<dim3(iblocks, splits), 256, 0, c->stream>>>(
            n, A.nt, A.hetero ? 1 : 0, A.pos.p, A.pair.p, ns, c->work_scale.p, c->work_xh.p, c->work_xl.p,
            c->work_part.p, nv, per, nsys, g * SYS_PER_GROUP);
    }
    PXH_HIP(hipGetLastError());
    PXH_HIP(hipEventRecord(c->ev1, c->stream));
    finish_product(c, nsys, d_sys_list, d_par, x, y, ld, splits);
    PXH_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    PXH_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    // algorithmic flops: the contraction of the 4N x 4N real T with 2 nsys real
    // columns (2 flops per element and column) plus the pair evaluation
    const double nn = (double)n * (double)(n - 1);
    c->stats.v[3] += ms;
    c->stats.v[8] += groups;
    c->stats.v[10] += nn * 16.0 * 2.0 * 2.0 * nsys + nn * 42.0 * groups;
    c->stats.v[12] += ms;
    c->stats.v[6] += (double)n * n * groups;
}

}  // namespace pxh
