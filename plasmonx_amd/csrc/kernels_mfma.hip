// MFMA form of the omegaFQFmu T sweep for a block of systems (gfx950).
//
// t_k = T x_k for up to 16*NCB systems at once.  T (4N x 4N, atom-interleaved
// rows [q_i, mu_ix, mu_iy, mu_iz]) is the frequency-independent real
// interaction matrix of src/matrix/matrix_module.f90 (T_qq :58-115,
// T_qmu/T_muq :120-234, T_mumu :240-345).  Its 4 x 4 block for a pair (i, j)
// is a function of five pair scalars (T_qq, the two charge-dipole radial
// factors, the two dipole-dipole coefficients) and the distance vector d:
//     row q:    [ T_qq,      g1 d_x,                g1 d_y,   g1 d_z ]
//     row mu_a: [ -g2 d_a,   c1 d_a d_b + c0 delta_ab  (b = x, y, z) ]
// The scalars are generated on the fly in FP32 by the lane that uses them,
// the elements are built in registers directly in the MFMA A-operand
// layout, split into fp16 hi + lo (value * 2^tscale), and contracted with
// the block X (4N x 32*NCB real columns: re / im of each system, also fp16
// hi + lo with a power-of-two scale per column) by
// v_mfma_f32_32x32x16_f16: acc += Ahi Bhi + Ahi Blo + Alo Bhi, a product
// with a 22-bit significand, accumulated in FP32 in two levels.
//
// Beyond the far-field radius the Gaussian damping (erf -> 1, exp -> 0) is
// below FP32 resolution and the scalars take their Coulomb limits, so the
// erf / exp evaluations run only for the near pairs of a wavefront.
//
// What bounds the kernel on gfx950: a 32x32x16 MFMA issues every 32 cycles
// per SIMD and hides only a few VALU instructions; the rest of the element
// generation adds to it, and every X fragment read from LDS costs LDS
// bandwidth.  Hence: no LDS round trip for the T elements (each lane builds
// the rows it multiplies), each X fragment read feeds two row blocks, and
// the X fragments reach LDS by DMA instead of through registers.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include <algorithm>
#include <cmath>

#include "pxh_internal.h"

namespace pxh {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kXTarget = 13;       // columns are scaled so that max |x| * s lies in [2^12, 2^13)
constexpr double kTwoOverSqrtPi = 1.1283791670955126;
constexpr double kFourOverSqrtPi = 2.2567583341910252;
constexpr int kAtomsI = 64;        // target atoms per workgroup
constexpr int kAtomsJ = 16;        // source atoms per step (K = 64)

// Per-column scale 2^e with max|x_col| * 2^e in [2^(kXTarget-1), 2^kXTarget),
// in two stages: max |re| and max |im| of each system over chunks of
// kScaleRows rows (one workgroup per chunk and system), then the scales.
constexpr int kScaleRows = 4096;

__global__ __launch_bounds__(256) void k_col_max(int nv, const int *__restrict__ sys_list,
                                                 const double2 *__restrict__ x, int ld, double *__restrict__ part) {
    __shared__ double sre[4], sim[4];
    const int ch = blockIdx.x, k = blockIdx.y;
    const double2 *xs = x + (size_t)sys_list[k] * ld;
    double mre = 0.0, mim = 0.0;
    const int r1 = min(nv, (ch + 1) * kScaleRows);
    for (int r = ch * kScaleRows + threadIdx.x; r < r1; r += 256) {
        const double2 v = xs[r];
        mre = fmax(mre, fabs(v.x));
        mim = fmax(mim, fabs(v.y));
    }
    for (int off = 32; off > 0; off >>= 1) {
        mre = fmax(mre, __shfl_xor(mre, off, 64));
        mim = fmax(mim, __shfl_xor(mim, off, 64));
    }
    if (threadIdx.x % 64 == 0) {
        sre[threadIdx.x / 64] = mre;
        sim[threadIdx.x / 64] = mim;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[((size_t)k * gridDim.x + ch) * 2] = fmax(fmax(sre[0], sre[1]), fmax(sre[2], sre[3]));
        part[((size_t)k * gridDim.x + ch) * 2 + 1] = fmax(fmax(sim[0], sim[1]), fmax(sim[2], sim[3]));
    }
}

__global__ void k_col_scale(int nsys, int nchunks, const double *__restrict__ part, float *__restrict__ scale,
                            int ncols) {
    const int c = threadIdx.x;  // column: re / im of system c / 2
    if (c >= ncols) return;
    const int k = c / 2;
    double m = 0.0;
    if (k < nsys)
        for (int ch = 0; ch < nchunks; ++ch) m = fmax(m, part[((size_t)k * nchunks + ch) * 2 + (c & 1)]);
    float s = 0.f;
    if (m > 0.0) {
        int e;
        frexp(m, &e);  // max = f 2^e, f in [0.5, 1)
        s = ldexpf(1.f, kXTarget - e);
    }
    scale[c] = s;
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

// Charge-dipole radial factor (erf(a) - 2a/sqrt(pi) exp(-a^2)) / r^3.
__device__ __forceinline__ float qmu_factor(float a, float inv_r3) {
    return (erff(a) - (float)kTwoOverSqrtPi * a * __expf(-a * a)) * inv_r3;
}

// erf(x) for x >= 0 given e = exp(-x^2), Abramowitz & Stegun 7.1.26:
// 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) e with t = 1 / (1 + p x),
// absolute error <= 1.5e-7.  Used where every argument of the wave is >= 1,
// so that the charge-dipole and dipole-dipole differences
// erf(a) - 2a/sqrt(pi) e keep a relative error of the same order.
__device__ __forceinline__ float erf_as(float x, float e) {
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, x, 1.0f));
    const float p = fmaf(fmaf(fmaf(fmaf(1.061405429f, t, -1.453152027f), t, 1.421413741f), t, -0.284496736f), t,
                         0.254829592f);
    return fmaf(-p * t, e, 1.0f);
}

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));

// Split of a float pair into fp16 hi + lo: hi = RN16(v), lo = RN16(v - hi).
// v - hi is exact in FP32; v_fma_mix{lo,hi}_f16 read hi as fp16 and write
// the rounded difference straight into the low / high half of lo.
__device__ __forceinline__ void split2(float v0, float v1, half2v &hi, half2v &lo) {
    hi = __builtin_convertvector((float2v){v0, v1}, half2v);
    unsigned r;
    asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(hi), "v"(v0));
    asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(r) : "v"(hi), "v"(v1));
    lo = __builtin_bit_cast(half2v, r);
}

__device__ __forceinline__ void split4(float e0, float e1, float e2, float e3, half4 &hi, half4 &lo) {
    half2v h0, l0, h1, l1;
    split2(e0, e1, h0, l0);
    split2(e2, e3, h1, l1);
    hi = (half4){h0[0], h0[1], h1[0], h1[1]};
    lo = (half4){l0[0], l0[1], l1[0], l1[1]};
}

// The pair scalars of the 4 x 4 block of (i, j), scaled by tsc:
// o = [T_qq, g1, -g2, c0, c1, dx, dy, dz] (rows q: [T_qq, g1 d];
// mu_a: [-g2 d_a, c1 d_a d_b + c0 delta_ab]).  The pair i = j (r^2 = 0) gets
// r^2 = 3e38, whose scaled elements round to zero in fp16; the padding atoms
// past n (the position array is padded to a multiple of 128 with atoms 1e15
// bohr away) give elements that round to zero as well, and their X is zero.
// Beyond the far-field radius the Coulomb limits share one scaled 1/r^3.
template <bool HET>
__device__ __forceinline__ void pair_scalars(const double4 &pi, const double4 &pj, int ti, int nt, float rfar2,
                                             float tsc, const float (*spc)[4], float *o) {
    const float dx = (float)(pi.x - pj.x);
    const float dy = (float)(pi.y - pj.y);
    const float dz = (float)(pi.z - pj.z);
    const float r2a = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float r2 = r2a > 0.f ? r2a : 3.0e38f;  // i = j (also padding against padding)
    const float inv_r = __builtin_amdgcn_rsqf(r2);  // r2 is never denormal here
    const float inv_r2 = inv_r * inv_r;
    float tqq = inv_r * tsc;
    float c0 = tqq * inv_r2;
    float g1 = c0, g2 = -c0;
    float c1 = -3.f * c0 * inv_r2;
    if (r2 < rfar2) {
        const float inv_r3 = inv_r2 * inv_r;
        const int tj = (int)pj.w;
        const float r = r2 * inv_r;
        const float *pp = spc[ti + nt * tj];
        const float a1 = r * pp[0], a2 = r * pp[1], a3 = r * pp[2];
        const float a2h = HET ? r * spc[tj + nt * ti][1] : a2;
        const float e3 = __expf(-a3 * a3);
        float fd1;
        if (__all(fminf(fminf(a1, a2), fminf(a3, a2h)) >= 1.0f)) {
            const float e1 = __expf(-a1 * a1), e2 = __expf(-a2 * a2);
            tqq = erf_as(a1, e1) * inv_r * tsc;
            g1 = (erf_as(a2, e2) - (float)kTwoOverSqrtPi * a2 * e2) * inv_r3 * tsc;
            if constexpr (HET) {
                const float e2h = __expf(-a2h * a2h);
                g2 = -(erf_as(a2h, e2h) - (float)kTwoOverSqrtPi * a2h * e2h) * inv_r3 * tsc;
            } else {
                g2 = -g1;
            }
            fd1 = erf_as(a3, e3) - (float)kTwoOverSqrtPi * a3 * e3;
        } else {
            tqq = erff(a1) * inv_r * tsc;
            g1 = qmu_factor(a2, inv_r3) * tsc;
            g2 = -(HET ? qmu_factor(a2h, inv_r3) * tsc : g1);
            fd1 = erff(a3) - (float)kTwoOverSqrtPi * a3 * e3;
        }
        const float fd2 = (float)kFourOverSqrtPi * pp[2] * pp[2] * pp[2] * e3;
        c0 = fd1 * inv_r3 * tsc;
        c1 = (fd2 * inv_r2 - 3.f * fd1 * inv_r3 * inv_r2) * tsc;
    }
    o[0] = tqq;
    o[1] = g1;
    o[2] = g2;
    o[3] = c0;
    o[4] = c1;
    o[5] = dx;
    o[6] = dy;
    o[7] = dz;
}

// Workgroup: 256 work-items (4 waves), kAtomsI = 64 target atoms.  Wave w
// owns targets 32 (w / 2) .. 32 (w / 2) + 31 and components 2 (w % 2) and
// 2 (w % 2) + 1: two MFMA row blocks (row = target, block = component) for
// all 32 * NCB columns.  Per source step (16 atoms, K = 64) the X fragments
// and the source positions reach LDS through the DMA path of the vector
// memory unit (double-buffered, issued one step ahead, one barrier per
// step).  In k-step s lane (target r, k half h) evaluates the pairs
// (r, source 4s + 2h + p), p = 0, 1, builds its two rows of both 4 x 4
// blocks in registers as the A fragments (fp16 hi / lo) and runs
// 2 x NCB x 3 MFMAs; each X fragment read from LDS feeds both row blocks.
// Accumulation: FP32 in the MFMA registers for flush_steps source
// steps, then into a second FP32 accumulator (blocked summation).
// Blocks are placed so that all workgroups of one source split share an XCD
// (split = linear block id mod splits): with splits a multiple of 8 an XCD's
// L2 holds the X fragments of its split.
template <int NCB, int LV, bool HET, bool R4>
__global__ __launch_bounds__(256, LV == 1 ? 3 : 2) void k_tsweep_mfma(int n, int nt, int flush_steps, float rfar2,
                                                                      float tsc,
                                                        const double4 *__restrict__ pos,
                                                        const PairConst *__restrict__ pc, int nsys,
                                                        const float *__restrict__ scale,
                                                        const _Float16 *__restrict__ xh,
                                                        const _Float16 *__restrict__ xl, float2 *__restrict__ part,
                                                        int nv, int splits, int jsteps_per_split,
                                                        int nsys_total_in_part, int sys_offset) {
    __shared__ half8 sBh[2][4 * NCB * 64];  // [buffer][k-step][cb][lane]
    __shared__ half8 sBl[2][4 * NCB * 64];
    __shared__ double4 spos[2][kAtomsJ];
    __shared__ float spc[kMaxTypes * kMaxTypes][4];
    constexpr int NRB = R4 ? 4 : 2;  // row blocks (components) per wave
    const int tid = threadIdx.x, lane = tid % 64;
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const int cp = R4 ? 0 : wave % 2;  // two row blocks: rows 2 cp and 2 cp + 1
    const int split = blockIdx.x % splits;
    const int i0 = R4 ? (blockIdx.x / splits) * 4 * 32 + 32 * wave : (blockIdx.x / splits) * kAtomsI + 32 * (wave / 2);
    const int jb0 = split * jsteps_per_split;
    const int jb_end = min((n + kAtomsJ - 1) / kAtomsJ, jb0 + jsteps_per_split);
    for (int e = tid; e < nt * nt; e += 256) {
        spc[e][0] = (float)pc[e].inv_rqq;
        spc[e][1] = (float)pc[e].inv_rqmu;
        spc[e][2] = (float)pc[e].inv_rmumu;
        spc[e][3] = (float)pc[e].eta_bar;
    }
    const int kh = lane / 32;
    const int gi = i0 + lane % 32;
    const double4 pi = pos[gi];  // padded: targets past n sit far away
    const int ti = (int)pi.w;
    f32x16 acc[NRB][NCB];
    float acc2[NRB][NCB][16];
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                acc[rb][cb][v] = 0.f;
                acc2[rb][cb][v] = 0.f;
            }
    // LDS DMA (global_load_lds): lane L's 16 bytes land at the wave's LDS
    // base + 16 L.  X fragments: element tid + 256 r of the step; positions:
    // 32 lanes of wave 0 copy the 16 double4 of the step.
    auto issue = [&](int jb, int b) {
        const size_t base = (size_t)jb * 4 * NCB * 64 + tid;
#pragma unroll
        for (int r = 0; r < NCB; ++r) {
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)(xh + (base + 256 * r) * 8),
                (__attribute__((address_space(3))) void *)(&sBh[b][64 * wave + 256 * r]), 16, 0, 0);
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)(xl + (base + 256 * r) * 8),
                (__attribute__((address_space(3))) void *)(&sBl[b][64 * wave + 256 * r]), 16, 0, 0);
        }
        if (wave == 0 && lane < 2 * kAtomsJ) {
            const int e = jb * kAtomsJ + lane / 2;  // padded: sources past n sit far away, with zero X
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)(reinterpret_cast<const char *>(pos + e) +
                                                                 16 * (lane % 2)),
                (__attribute__((address_space(3))) void *)(&spos[b][0]), 16, 0, 0);
        }
    };
    if (jb0 < jb_end) issue(jb0, 0);
    int buf = 0;
    for (int jbb = jb0; jbb < jb_end; jbb += flush_steps) {
        const int jbe = min(jb_end, jbb + flush_steps);
        for (int jb = jbb; jb < jbe; ++jb) {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();  // step jb is in LDS; every wave is done with step jb - 1's buffer
            if (jb + 1 < jb_end) issue(jb + 1, buf ^ 1);
            const half8 *bh_ = sBh[buf];
            const half8 *bl_ = sBl[buf];
#pragma unroll 1
            for (int s = 0; s < 4; ++s) {
                half4 h[NRB][2], l[NRB][2];  // [row block][pair]
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const int jl = 4 * s + 2 * kh + p;
                    const int gj = jb * kAtomsJ + jl;
                    float sc[8];
                    pair_scalars<HET>(pi, spos[buf][jl], ti, nt, rfar2, tsc, spc, sc);
                    const float tqq = sc[0], g1 = sc[1], g2 = sc[2], c0 = sc[3], c1 = sc[4];
                    const float dx = sc[5], dy = sc[6], dz = sc[7];
                    // the same element expressions in both layouts, so that
                    // their products agree bit for bit
                    if (R4 || cp == 0) {  // rows q and mu_x
                        const float ux = c1 * dx;
                        split4(tqq, g1 * dx, g1 * dy, g1 * dz, h[0][p], l[0][p]);
                        split4(g2 * dx, fmaf(ux, dx, c0), ux * dy, ux * dz, h[1][p], l[1][p]);
                    }
                    if (R4 || cp == 1) {  // rows mu_y and mu_z
                        constexpr int o = R4 ? 2 : 0;
                        const float uy = c1 * dy, uz = c1 * dz;
                        split4(g2 * dy, uy * dx, fmaf(uy, dy, c0), uy * dz, h[o][p], l[o][p]);
                        split4(g2 * dz, uz * dx, uz * dy, fmaf(uz, dz, c0), h[o + 1][p], l[o + 1][p]);
                    }
                }
                // A row = target lane % 32, k = 8 (lane / 32) + e = component e % 4 of
                // source 4s + 2 (lane / 32) + e / 4
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) {
                    const half8 bh = bh_[(s * NCB + cb) * 64 + lane];
                    const half8 bl = bl_[(s * NCB + cb) * 64 + lane];
#pragma unroll
                    for (int rb = 0; rb < NRB; ++rb) {
                        const half8 ah = (half8){h[rb][0][0], h[rb][0][1], h[rb][0][2], h[rb][0][3],
                                                 h[rb][1][0], h[rb][1][1], h[rb][1][2], h[rb][1][3]};
                        const half8 al = (half8){l[rb][0][0], l[rb][0][1], l[rb][0][2], l[rb][0][3],
                                                 l[rb][1][0], l[rb][1][1], l[rb][1][2], l[rb][1][3]};
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[rb][cb], 0, 0, 0);
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[rb][cb], 0, 0, 0);
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[rb][cb], 0, 0, 0);
                    }
                }
            }
            buf ^= 1;
        }
        if constexpr (LV == 2) {
#pragma unroll
            for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        acc2[rb][cb][v] += acc[rb][cb][v];
                        acc[rb][cb][v] = 0.f;
                    }
        }
    }
    // ---- write partials: acc2[rb][cb][v] -> target i0 + 8 (v / 4) + 4 (lane / 32) + v % 4,
    // component 2 cp + rb, column cb * 32 + lane % 32
    const double tinv = 1.0 / (double)tsc;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const int col = cb * 32 + (lane % 32);
        const int sys = col / 2;
        if (sys >= nsys) continue;
        const double cs = scale[col] > 0.f ? tinv / (double)scale[col] : 0.0;
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int i = i0 + (v / 4) * 8 + (lane / 32) * 4 + (v % 4);
                if (i >= n) continue;
                const int c = R4 ? rb : 2 * cp + rb;
                const int row = c == 0 ? i : n + 3 * i + c - 1;
                float *dst = reinterpret_cast<float *>(
                    part + ((size_t)split * nv + row) * nsys_total_in_part + sys_offset + sys);
                dst[col & 1] = (float)((double)(LV == 2 ? acc2[rb][cb][v] : acc[rb][cb][v]) * cs);
            }
    }
}

}  // namespace


// Power-of-two scale of the T elements: the largest Gaussian-damped element
// of any type pair (T_qq <= 2 kappa_qq / sqrt(pi), |g d| <= 0.45 kappa_qmu^2,
// |c1 d d + c0| <= kappa_mumu^3) times the scale stays below 2^14, so that
// fp16 holds it with margin and small far-field elements stay normal.
float mfma_t_scale(const Atoms &A) {
    double bound = 1e-30;
    for (const PairConst &p : A.h_pair)
        bound = std::max({bound, 1.13 * p.inv_rqq, 0.45 * p.inv_rqmu * p.inv_rqmu,
                          p.inv_rmumu * p.inv_rmumu * p.inv_rmumu});
    const int e = std::min(20, std::max(0, (int)std::floor(std::log2(16384.0 / bound))));
    return std::ldexp(1.f, e);
}

namespace {

void col_scales(pxh_ctx *c, int nv, int ns, const int *lst, const double2 *x, int ld, int ncols) {
    const int nchunks = (nv + kScaleRows - 1) / kScaleRows;
    c->work_colmax.alloc((size_t)2 * nchunks * std::max(ns, 1));
    k_col_max<<<dim3(nchunks, ns), 256, 0, c->stream>>>(nv, lst, x, ld, c->work_colmax.p);
    k_col_scale<<<1, 256, 0, c->stream>>>(ns, nchunks, c->work_colmax.p, c->work_scale.p, ncols);
}

// Source splits chosen by timing, per device and atom count, shared by every
// context of the process (so that one target always sums its partials the
// same way).
std::mutex g_tune_mu;
std::map<std::pair<int, int>, int> g_tuned_splits;

template <int K>
void launch_sweep(pxh_ctx *c, int iblocks, int splits, int per, float rfar2, float tsc, int ns, int nsys, int goff) {
    const Atoms &A = c->atoms;
    // one column block: four row blocks per wave (each pair's elements built
    // once per workgroup instead of twice), 128 targets per workgroup
    constexpr bool R4c = K == 1;
    const bool R4 = R4c && c->opt.mfma_r4;
    const int blocks = R4 ? (A.n + 127) / 128 : iblocks;
    auto go = [&](auto kern) {
        kern<<<blocks * splits, 256, 0, c->stream>>>(A.n, A.nt, c->opt.mfma_flush, rfar2, tsc, A.pos.p, A.pair.p, ns,
                                                     c->work_scale.p, c->work_xh.p, c->work_xl.p,
                                                     reinterpret_cast<float2 *>(c->work_part.p), A.n_var(), splits,
                                                     per, nsys, goff);
    };
    auto pick = [&](auto r4) {
        constexpr bool F = decltype(r4)::value;
        if (c->opt.mfma_levels == 1)
            A.hetero ? go(k_tsweep_mfma<K, 1, true, F>) : go(k_tsweep_mfma<K, 1, false, F>);
        else
            A.hetero ? go(k_tsweep_mfma<K, 2, true, F>) : go(k_tsweep_mfma<K, 2, false, F>);
    };
    if constexpr (R4c) {
        if (R4)
            pick(std::true_type{});
        else
            pick(std::false_type{});
    } else {
        pick(std::false_type{});
    }
}

}  // namespace

void finish_product(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x, double2 *y,
                    int ld, int splits, bool fp32_class, double2 *part, double2 *t, bool part_f32);

namespace {

// Time one 48-system product for each candidate split count (the pass over
// the prepared first group plus the reduction of the split partials): a
// warm-up round, then the best of three timed rounds per candidate.  The
// first candidate (the heuristic) is kept unless another is more than 3 %
// faster, so that timing noise does not move it.
template <int K>
int tune_splits(pxh_ctx *c, const std::vector<int> &cand, int iblocks, int jsteps, float rfar2, float tsc, int ns,
                int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x, double2 *y, int ld) {
    hipEvent_t e0, e1;
    PXH_HIP(hipEventCreate(&e0));
    PXH_HIP(hipEventCreate(&e1));
    std::vector<float> t(cand.size(), 1e30f);
    for (int rep = 0; rep < 4; ++rep)  // the first round warms the caches
        for (size_t q = 0; q < cand.size(); ++q) {
            const int sp = cand[q];
            const int per = (jsteps + sp - 1) / sp;
            PXH_HIP(hipEventRecord(e0, c->stream));
            launch_sweep<K>(c, iblocks, sp, per, rfar2, tsc, ns, nsys, 0);
            finish_product(c, nsys, d_sys_list, d_par, x, y, ld, sp, true, c->work_part.p, c->work_t.p, true);
            PXH_HIP(hipEventRecord(e1, c->stream));
            PXH_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            PXH_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) t[q] = std::min(t[q], ms);
        }
    int best = cand[0];
    float best_ms = t[0] * 0.97f;
    for (size_t q = 1; q < cand.size(); ++q)
        if (t[q] < best_ms) {
            best_ms = t[q];
            best = cand[q];
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace

// MFMA path of the FP32-class product for omegaFQFmu (Gaussian kernel).
void apply_atoms_mfma(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x,
                      double2 *y, int ld) {
    constexpr int NCB = 3;  // 48 systems per pass
    constexpr int SYS_PER_GROUP = 16 * NCB;
    Atoms &A = c->atoms;
    const int n = A.n, nv = A.n_var();
    const int jsteps = (n + kAtomsJ - 1) / kAtomsJ;
    const int iblocks = (n + kAtomsI - 1) / kAtomsI;
    const int groups = (nsys + SYS_PER_GROUP - 1) / SYS_PER_GROUP;
    // source splits: a multiple of 8 (one XCD per split residue) once every
    // split still spans >= 16 source steps, and >= ~1024 workgroups (4 per CU)
    int splits = c->opt.splits;
    if (splits <= 0) {
        splits = std::max(1, (1024 + iblocks - 1) / iblocks);
        if (jsteps >= 8 * 16) splits = std::min((splits + 7) / 8 * 8, jsteps / 16 / 8 * 8);
        splits = std::max(1, std::min(splits, jsteps / 16));
    }
    splits = std::min(splits, jsteps);
    const size_t xelems = (size_t)jsteps * 4 * NCB * 64 * 8;
    c->work_xh.alloc(xelems);
    c->work_xl.alloc(xelems);
    c->work_scale.alloc(32 * NCB);
    c->work_t.alloc((size_t)nsys * nv);
    const float tsc = mfma_t_scale(A);
    const float rfar2 = (float)far_field_r2(A, false);
    if (c->opt.splits <= 0 && c->opt.autotune && nsys >= SYS_PER_GROUP && jsteps >= 8 * 16) {
        // autotuner: the heuristic, half and twice it, and 8..32, on the
        // first full pass that this atom count meets in the process
        const std::pair<int, int> key{c->device, n};
        int tuned = 0;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned_splits.find(key);
            if (it != g_tuned_splits.end()) tuned = it->second;
        }
        if (!tuned) {
            std::vector<int> cand{splits};
            for (int sp : {splits / 2 / 8 * 8, splits * 2, 8, 16, 24, 32})
                if (sp >= 8 && sp <= jsteps / 16 && std::find(cand.begin(), cand.end(), sp) == cand.end())
                    cand.push_back(sp);
            c->work_part.alloc((size_t)(*std::max_element(cand.begin(), cand.end())) * nsys * nv);
            col_scales(c, nv, SYS_PER_GROUP, d_sys_list, x, ld, 32 * NCB);
            k_prep_x<NCB><<<jsteps, 256, 0, c->stream>>>(n, SYS_PER_GROUP, d_sys_list, x, ld, c->work_scale.p,
                                                         c->work_xh.p, c->work_xl.p);
            tuned = tune_splits<NCB>(c, cand, iblocks, jsteps, rfar2, tsc, SYS_PER_GROUP, nsys, d_sys_list, d_par, x,
                                     y, ld);
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tuned_splits.emplace(key, tuned);
            tuned = g_tuned_splits[key];
        }
        splits = tuned;
    }
    const int per = (jsteps + splits - 1) / splits;
    c->work_part.alloc((size_t)splits * nsys * nv);
    const int t_product = timer_begin(c);
    for (int g = 0; g < groups; ++g) {
        const int ns = std::min(SYS_PER_GROUP, nsys - g * SYS_PER_GROUP);
        const int *lst = d_sys_list + g * SYS_PER_GROUP;
        // column blocks of this pass: a partly filled batch runs only the
        // blocks it occupies (16 systems each)
        auto pass = [&](auto ncb_tag) {
            constexpr int K = decltype(ncb_tag)::value;
            col_scales(c, nv, ns, lst, x, ld, 32 * K);
            k_prep_x<K><<<jsteps, 256, 0, c->stream>>>(n, ns, lst, x, ld, c->work_scale.p, c->work_xh.p,
                                                       c->work_xl.p);
            // the pair kernel alone, timed on the library's stream
            const int t_kernel = timer_begin(c);
            launch_sweep<K>(c, iblocks, splits, per, rfar2, tsc, ns, nsys, g * SYS_PER_GROUP);
            timer_end(c, t_kernel, 12);
        };
        const int ncb = (ns + 15) / 16;
        if (ncb == 1)
            pass(std::integral_constant<int, 1>{});
        else if (ncb == 2)
            pass(std::integral_constant<int, 2>{});
        else
            pass(std::integral_constant<int, 3>{});
        c->stats.v[13] += 1;
    }
    PXH_HIP(hipGetLastError());
    timer_end(c, t_product, 3);
    finish_product(c, nsys, d_sys_list, d_par, x, y, ld, splits, true, c->work_part.p, c->work_t.p, true);
    // algorithmic flops: the contraction of the 4N x 4N real T with 2 nsys real
    // columns (2 flops per element and column) plus the pair evaluation
    const double nn = (double)n * (double)(n - 1);
    c->stats.v[8] += groups;
    c->stats.v[10] += nn * 16.0 * 2.0 * 2.0 * nsys + nn * 42.0 * groups;
    c->stats.v[23] = splits;
    c->stats.v[6] += (double)n * n * groups;
}

}  // namespace pxh
