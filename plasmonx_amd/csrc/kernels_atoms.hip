// Interaction kernels of the atomistic models (omegaFQ, omegaFQFmu) for
// gfx950: on-the-fly T sweep, sparse K-bar, diagonal shifts, dynamic RHS and
// the dense constant matrix.
//
// The arithmetic restates src/matrix/matrix_module.f90 (T_qq :58-115,
// T_qmu/T_muq :120-234, T_mumu :240-345, K-bar :393-426, H-bar :435-482,
// Fermi :694-715, on-the-fly products :727-1125), src/rhs/rhs_module.f90
// (:43-430) and the diagonal shifts of wfq_module.f90:398-424 and
// wfqfmu_module.f90:698-746.
//
// Layout in HBM: positions as double4 (x, y, z, type); a block of systems is
// [system][n_var] complex (double2), each system in the reference's order
// [q_1..q_N, mu_1x, mu_1y, mu_1z, ...].  The T sweep is a gather: one
// work-item owns a target atom i and walks the source atoms j through LDS
// tiles; the source range is split over blockIdx.y so that a launch has
// thousands of workgroups, and the split partial sums are added in a fixed
// order afterwards (deterministic results, no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <type_traits>

#include "pxh_internal.h"

namespace pxh {

namespace {

constexpr double kTwoOverSqrtPi = 1.1283791670955126;
constexpr double kFourOverSqrtPi = 2.2567583341910252;
constexpr int TB = 256;  // work-items per workgroup = target atoms per tile = source atoms per LDS tile
// Algorithmic flops of the T sweep, counted from k_tsweep with each erf /
// exp / rsqrt counted as one: the interaction elements of one ordered pair
// (omegaFQFmu: distance 10, T_qq 3, charge-dipole 12, dipole-dipole 17;
// omegaFQ: 13) and the contraction with one complex vector (omegaFQFmu:
// d.mu 10, q row 8, dipole coefficient 6, mu rows 24; omegaFQ: 4).
constexpr int kPairEvalFlopsFmu = 42;
constexpr int kPairVecFlopsFmu = 48;
constexpr int kPairEvalFlopsQ = 13;
constexpr int kPairVecFlopsQ = 4;

template <typename R>
struct Cx {
    R re, im;
};

template <typename R>
__device__ __forceinline__ R d_erf(R x);
template <>
__device__ __forceinline__ double d_erf<double>(double x) {
    return erf(x);
}
template <>
__device__ __forceinline__ float d_erf<float>(float x) {
    return erff(x);
}
template <typename R>
__device__ __forceinline__ R d_exp(R x);
template <>
__device__ __forceinline__ double d_exp<double>(double x) {
    return exp(x);
}
template <>
__device__ __forceinline__ float d_exp<float>(float x) {
    return __expf(x);
}
template <typename R>
__device__ __forceinline__ R d_rsqrt(R x);
// 1/sqrt(x) in FP64 from the hardware estimate (v_rsq_f64) and two Newton
// steps y <- y (3/2 - x y^2 / 2), each of which squares the relative error:
// full FP64 precision (within 4e-15 of 1.0 / sqrt(x) over the products of the
// tests) in 6 instructions instead of a square root and a division.
template <>
__device__ __forceinline__ double d_rsqrt<double>(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = y * fma(-hx * y, y, 1.5);
    y = y * fma(-hx * y, y, 1.5);
    return y;
}
template <>
__device__ __forceinline__ float d_rsqrt<float>(float x) {
    return rsqrtf(x);
}

template <typename R>
struct PairR {
    R inv_rqq, inv_rqmu, inv_rmumu, eta_bar;
};

// Source-atom record staged in LDS for KV systems.
template <typename R, int KV>
struct SrcTile {
    double4 pos[TB];
    Cx<R> q[KV][TB];
    Cx<R> m[KV][3][TB];
};

// T sweep: part[split][row][k] = sum_{j in split, j != i} T(i, j) x_k(j).
// KERN: 1 Gaussian, 2 Coulomb, 3 Ohno.  FMU: omegaFQFmu (4 unknowns/atom).
// Work-item t of workgroup b takes target atom tperm[b TB + t]: targets in
// spatially compact blocks (Morton order), so that a wavefront's targets
// meet few sources inside the far-field radius and the erf / exp branch runs
// for fewer wavefronts.
template <typename R, int KV, bool FMU, int KERN>
__global__ __launch_bounds__(TB) void k_tsweep(int n, int nt, int hetero, const double4 *__restrict__ pos,
                                               const PairConst *__restrict__ pc, int nsys,
                                               const int *__restrict__ sys_list, const double2 *__restrict__ x,
                                               int ld, double2 *__restrict__ part, int nv, int j_chunk,
                                               double rfar2, const int *__restrict__ tperm) {
    __shared__ SrcTile<R, KV> tile;
    __shared__ PairR<R> spc[kMaxTypes * kMaxTypes];
    const int tid = threadIdx.x;
    const int it = blockIdx.x * TB + tid;
    const int i = it < n ? tperm[it] : n;
    const int split = blockIdx.y;
    const int g0 = blockIdx.z * KV;  // first system slot of this group
    const int nk = min(KV, nsys - g0);
    for (int e = tid; e < nt * nt; e += TB) {
        spc[e].inv_rqq = (R)pc[e].inv_rqq;
        spc[e].inv_rqmu = (R)pc[e].inv_rqmu;
        spc[e].inv_rmumu = (R)pc[e].inv_rmumu;
        spc[e].eta_bar = (R)pc[e].eta_bar;
    }
    const R rfar2r = (R)rfar2;
    const bool active = i < n;
    double4 pi = active ? pos[i] : make_double4(0, 0, 0, 0);
    const int ti = (int)pi.w;
    double aq_re[KV], aq_im[KV], am_re[KV][3], am_im[KV][3];
#pragma unroll
    for (int v = 0; v < KV; ++v) {
        aq_re[v] = aq_im[v] = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) am_re[v][a] = am_im[v][a] = 0.0;
    }
    const int j_begin = split * j_chunk;
    const int j_end = min(n, j_begin + j_chunk);
    for (int jt = j_begin; jt < j_end; jt += TB) {
        __syncthreads();
        {
            const int j = jt + tid;
            if (j < j_end) {
                tile.pos[tid] = pos[j];
#pragma unroll
                for (int v = 0; v < KV; ++v) {
                    if (v < nk) {
                        const double2 *xs = x + (size_t)sys_list[g0 + v] * ld;
                        double2 q = xs[j];
                        tile.q[v][tid].re = (R)q.x;
                        tile.q[v][tid].im = (R)q.y;
                        if (FMU) {
#pragma unroll
                            for (int a = 0; a < 3; ++a) {
                                double2 m = xs[n + 3 * j + a];
                                tile.m[v][a][tid].re = (R)m.x;
                                tile.m[v][a][tid].im = (R)m.y;
                            }
                        }
                    } else {
                        tile.q[v][tid].re = tile.q[v][tid].im = (R)0;
                        if (FMU) {
#pragma unroll
                            for (int a = 0; a < 3; ++a) tile.m[v][a][tid].re = tile.m[v][a][tid].im = (R)0;
                        }
                    }
                }
            }
        }
        __syncthreads();
        const int cnt = min(TB, j_end - jt);
        // omegaFQ accumulates the charge rows in FP64 even with FP32 elements:
        // K-bar takes differences of neighbouring potentials, which would
        // amplify an FP32 sum's rounding (FP32 and FP64 FMAs issue at the same
        // rate here, packed FP32 being off)
        using QAcc = typename std::conditional<FMU, R, double>::type;
        // FP64 elements: one accumulation level (the tile's sums start from
        // and return to the running totals, in the same registers)
        constexpr bool kOneLevel = std::is_same<R, double>::value;
        QAcc tq_re[KV], tq_im[KV];
        R tm_re[KV][3], tm_im[KV][3];
#pragma unroll
        for (int v = 0; v < KV; ++v) {
            tq_re[v] = kOneLevel ? (QAcc)aq_re[v] : (QAcc)0;
            tq_im[v] = kOneLevel ? (QAcc)aq_im[v] : (QAcc)0;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                tm_re[v][a] = kOneLevel ? (R)am_re[v][a] : (R)0;
                tm_im[v][a] = kOneLevel ? (R)am_im[v][a] : (R)0;
            }
        }
        if (active) {
            for (int jj = 0; jj < cnt; ++jj) {
                const int j = jt + jj;
                if (j == i) continue;
                const double4 pj = tile.pos[jj];
                const R dx = (R)(pi.x - pj.x), dy = (R)(pi.y - pj.y), dz = (R)(pi.z - pj.z);
                const R r2 = dx * dx + dy * dy + dz * dz;
                const R inv_r = d_rsqrt<R>(r2);
                const R r = r2 * inv_r;
                const int tj = (int)pj.w;
                const PairR<R> p = spc[ti + nt * tj];
                const R inv_r2 = inv_r * inv_r;
                const R inv_r3 = inv_r2 * inv_r;
                R tqq, g1 = (R)0, g2 = (R)0, c0 = (R)0, c1 = (R)0;
                if (KERN == 1 && r2 >= rfar2r) {
                    // far field: the Gaussian damping is below the working
                    // precision (erf -> 1, exp -> 0 to rounding)
                    tqq = inv_r;
                    g1 = g2 = c0 = inv_r3;
                    c1 = (R)-3 * inv_r3 * inv_r2;
                } else {
                    if (KERN == 1)
                        tqq = d_erf<R>(r * p.inv_rqq) * inv_r;
                    else if (KERN == 2)
                        tqq = inv_r;
                    else
                        tqq = p.eta_bar * d_rsqrt<R>((R)1 + p.eta_bar * p.eta_bar * r2);
                    if (FMU) {
                        // charge on i, dipole on j
                        const R a1 = r * p.inv_rqmu;
                        const R e1 = d_exp<R>(-a1 * a1);
                        g1 = (d_erf<R>(a1) - (R)kTwoOverSqrtPi * a1 * e1) * inv_r3;
                        g2 = g1;
                        if (hetero) {  // charge on j, dipole on i
                            const R a2 = r * spc[tj + nt * ti].inv_rqmu;
                            const R e2 = d_exp<R>(-a2 * a2);
                            g2 = (d_erf<R>(a2) - (R)kTwoOverSqrtPi * a2 * e2) * inv_r3;
                        }
                        const R a3 = r * p.inv_rmumu;
                        const R e3 = d_exp<R>(-a3 * a3);
                        const R fd1 = d_erf<R>(a3) - (R)kTwoOverSqrtPi * a3 * e3;
                        const R fd2 = (R)kFourOverSqrtPi * p.inv_rmumu * p.inv_rmumu * p.inv_rmumu * e3;
                        c0 = fd1 * inv_r3;
                        c1 = fd2 * inv_r2 - (R)3 * fd1 * inv_r3 * inv_r2;
                    }
                }
                if (!FMU) {
#pragma unroll
                    for (int v = 0; v < KV; ++v) {
                        tq_re[v] += (QAcc)tqq * (QAcc)tile.q[v][jj].re;
                        tq_im[v] += (QAcc)tqq * (QAcc)tile.q[v][jj].im;
                    }
                    continue;
                }
#pragma unroll
                for (int v = 0; v < KV; ++v) {
                    const Cx<R> q = tile.q[v][jj];
                    const Cx<R> mx = tile.m[v][0][jj], my = tile.m[v][1][jj], mz = tile.m[v][2][jj];
                    const R dm_re = fma(dz, mz.re, fma(dy, my.re, dx * mx.re));
                    const R dm_im = fma(dz, mz.im, fma(dy, my.im, dx * mx.im));
                    // every term into the running sums by one FMA each
                    tq_re[v] = fma(g1, dm_re, fma(tqq, q.re, tq_re[v]));
                    tq_im[v] = fma(g1, dm_im, fma(tqq, q.im, tq_im[v]));
                    const R s_re = fma(c1, dm_re, -g2 * q.re);
                    const R s_im = fma(c1, dm_im, -g2 * q.im);
                    tm_re[v][0] = fma(c0, mx.re, fma(s_re, dx, tm_re[v][0]));
                    tm_im[v][0] = fma(c0, mx.im, fma(s_im, dx, tm_im[v][0]));
                    tm_re[v][1] = fma(c0, my.re, fma(s_re, dy, tm_re[v][1]));
                    tm_im[v][1] = fma(c0, my.im, fma(s_im, dy, tm_im[v][1]));
                    tm_re[v][2] = fma(c0, mz.re, fma(s_re, dz, tm_re[v][2]));
                    tm_im[v][2] = fma(c0, mz.im, fma(s_im, dz, tm_im[v][2]));
                }
            }
        }
#pragma unroll
        for (int v = 0; v < KV; ++v) {
            aq_re[v] = kOneLevel ? (double)tq_re[v] : aq_re[v] + (double)tq_re[v];
            aq_im[v] = kOneLevel ? (double)tq_im[v] : aq_im[v] + (double)tq_im[v];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                am_re[v][a] = kOneLevel ? (double)tm_re[v][a] : am_re[v][a] + (double)tm_re[v][a];
                am_im[v][a] = kOneLevel ? (double)tm_im[v][a] : am_im[v][a] + (double)tm_im[v][a];
            }
        }
    }
    if (!active) return;
    // part[split][row][system]
    double2 *out = part + (size_t)split * nv * nsys + g0;
#pragma unroll
    for (int v = 0; v < KV; ++v) {
        if (v >= nk) break;
        out[(size_t)i * nsys + v] = make_double2(aq_re[v], aq_im[v]);
        if (FMU) {
#pragma unroll
            for (int a = 0; a < 3; ++a)
                out[(size_t)(n + 3 * i + a) * nsys + v] = make_double2(am_re[v][a], am_im[v][a]);
        }
    }
}

// t[row][k] = sum_split part[split][row][k] + (q rows) eta_i x_q(i): the
// product intermediates are row-major (all systems of a row adjacent), so
// that the sparse K-bar stage gathers whole rows.
// The partials are FP64 (FP64 products) or FP32 (the MFMA products, whose
// FP32 accumulators they hold); the sum is FP64 either way.
template <typename P>
__global__ void k_reduce_splits(int n, int nv, int nsys, int splits, const P *__restrict__ part,
                                const int *__restrict__ sys_list, const double2 *__restrict__ x, int ld,
                                const double4 *__restrict__ pos, const double *__restrict__ eta,
                                double2 *__restrict__ t) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)nv * nsys) return;
    const int row = (int)(g / nsys), k = (int)(g % nsys);
    double re = 0.0, im = 0.0;
    for (int s = 0; s < splits; ++s) {
        const P p = part[(size_t)s * nv * nsys + g];
        re += (double)p.x;
        im += (double)p.y;
    }
    if (row < n) {
        const double e = eta[(int)pos[row].w];
        const double2 xq = x[(size_t)sys_list[k] * ld + row];
        re += e * xq.x;
        im += e * xq.y;
    }
    t[g] = make_double2(re, im);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// value of v held by lane l (l wave-uniform)
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// q rows: y_k(i) = sum_j c_ij (t_k(j) - t_k(i)) - z_q(t_i) x_k(i), with
// c_ij = K_ij, or 1/2 (K_ij + (w_tj / w_ti) K_ji) for alloys.  One wavefront
// per row and one lane per system: the CSR entries are wave-uniform loads
// and each neighbour's systems are one contiguous row of t.  Workgroups
// with the same blockIdx.x % 8 (one XCD) take one contiguous range of rows,
// so the neighbour rows they gather stay in that XCD's L2.
__global__ __launch_bounds__(256) void k_kbar_q(int n, int nt, int hetero, const int *__restrict__ rowptr,
                                                const int *__restrict__ col, const double *__restrict__ kij,
                                                const double *__restrict__ kji, const double4 *__restrict__ pos,
                                                int nsys, const int *__restrict__ sys_list,
                                                const SysParam *__restrict__ par, const double2 *__restrict__ t,
                                                const double2 *__restrict__ x, int ld, double2 *__restrict__ y) {
    const int nblk = (n + 3) / 4;
    const int per = (nblk + 7) / 8;
    const int blk = (blockIdx.x % 8) * per + blockIdx.x / 8;
    const int i = blk * 4 + threadIdx.x / 64;
    if (blk >= nblk || i >= n) return;
    const int lane = threadIdx.x % 64;
    const int ti = (int)pos[i].w;
    const int b = rowptr[i], e = rowptr[i + 1];
    for (int k0 = 0; k0 < nsys; k0 += 64) {
        const int k = k0 + lane;
        const bool on = k < nsys;
        const int kk = on ? k : 0;
        const int s = sys_list[kk];
        const double2 tii = t[(size_t)i * nsys + kk];
        double re = 0.0, im = 0.0;
        // 64 CSR entries at a time into registers (one per lane), then
        // broadcast with readlane so that 8 row gathers are in flight at once
        for (int zb = b; zb < e; zb += 64) {
            const int zc = min(64, e - zb);
            const bool has = lane < zc;
            const int cj = has ? col[zb + lane] : 0;
            const double kv = has ? kij[zb + lane] : 0.0;
            const double kw = has && hetero ? kji[zb + lane] : 0.0;
            const int tjv = hetero ? (int)pos[cj].w : 0;
            for (int q = 0; q < zc; q += 8) {
                double2 tv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int j = __builtin_amdgcn_readlane(cj, q + u);
                    tv[u] = t[(size_t)j * nsys + kk];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (q + u >= zc) break;
                    const double kz = readlane_f64(kv, q + u);
                    const double2 d = make_double2(tv[u].x - tii.x, tv[u].y - tii.y);
                    if (!hetero) {
                        re += kz * d.x;
                        im += kz * d.y;
                    } else {
                        const double kz2 = readlane_f64(kw, q + u);
                        const int tj = __builtin_amdgcn_readlane(tjv, q + u);
                        const double2 wr = par[s].wratio[ti + nt * tj];
                        const double2 c = make_double2(0.5 * (kz + wr.x * kz2), 0.5 * wr.y * kz2);
                        const double2 p = cmul(c, d);
                        re += p.x;
                        im += p.y;
                    }
                }
            }
        }
        if (on) {
            const double2 xi = x[(size_t)s * ld + i];
            const double2 zx = cmul(par[s].zq[ti], xi);
            y[(size_t)s * ld + i] = make_double2(re - zx.x, im - zx.y);
        }
    }
}

// mu rows: y_k = t_k + Lambda_i x_k, Lambda_i = sum_t nbfrac(t, i) / alpha_t.
__global__ void k_mu_rows(int n, int nt, const double *__restrict__ nbfrac, int nsys,
                          const int *__restrict__ sys_list, const SysParam *__restrict__ par,
                          const double2 *__restrict__ t, const double2 *__restrict__ x, int ld,
                          double2 *__restrict__ y) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)3 * n * nsys) return;
    const int r = (int)(g / nsys), k = (int)(g % nsys);
    const int s = sys_list[k];
    const int i = r / 3;
    double2 lam = make_double2(0, 0);
    for (int tt = 0; tt < nt; ++tt) {
        const double f = nbfrac[tt + (size_t)nt * i];
        lam.x += f * par[s].inv_alpha[tt].x;
        lam.y += f * par[s].inv_alpha[tt].y;
    }
    const double2 xv = x[(size_t)s * ld + n + r];
    const double2 tv = t[(size_t)n * nsys + g];
    const double2 lx = cmul(lam, xv);
    y[(size_t)s * ld + n + r] = make_double2(tv.x + lx.x, tv.y + lx.y);
}

// ---------------------------------------------------------------- K-bar CSR
__device__ __forceinline__ double fermi_one_minus(double d, double s, double r0, double r) {
    // 1 - f with f = 1/(1+exp(-d (r/(s r0) - 1)))  (matrix_module.f90:694-715)
    const double f = 1.0 / (1.0 + exp(-d * ((r / (s * r0)) - 1.0)));
    return 1.0 - f;
}

struct FermiTab {
    double d[kMaxTypes * kMaxTypes], s[kMaxTypes * kMaxTypes], r0[kMaxTypes * kMaxTypes],
        rcut2[kMaxTypes * kMaxTypes];
    double a[kMaxTypes];
};

__global__ void k_kbar_count(int n, int nt, const double4 *__restrict__ pos, FermiTab ft, int *__restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 pi = pos[i];
    const int ti = (int)pi.w;
    int c = 0;
    for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        const double4 pj = pos[j];
        const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const int tj = (int)pj.w;
        if (r2 <= ft.rcut2[ti + nt * tj] || r2 <= ft.rcut2[tj + nt * ti]) ++c;
    }
    cnt[i] = c;
}

__global__ void k_kbar_fill(int n, int nt, const double4 *__restrict__ pos, FermiTab ft,
                            const int *__restrict__ rowptr, int *__restrict__ col, double *__restrict__ kij,
                            double *__restrict__ kji) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 pi = pos[i];
    const int ti = (int)pi.w;
    int z = rowptr[i];
    for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        const double4 pj = pos[j];
        const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const int tj = (int)pj.w;
        if (r2 <= ft.rcut2[ti + nt * tj] || r2 <= ft.rcut2[tj + nt * ti]) {
            const double r = sqrt(r2);
            const int a = ti + nt * tj, b = tj + nt * ti;
            col[z] = j;
            kij[z] = fermi_one_minus(ft.d[a], ft.s[a], ft.r0[a], r) * ft.a[ti] / r;
            kji[z] = fermi_one_minus(ft.d[b], ft.s[b], ft.r0[b], r) * ft.a[tj] / r;
            ++z;
        }
    }
}

// Diagonal of K-bar T_qq per atom (the preconditioner of the batched path):
// (K T)_ii = sum_j K_ij (T_qq(j,i) - T_qq(i,i)) with T_qq(i,i) = eta_i.
__global__ void k_diag_sums(int n, int nt, int kern, const int *__restrict__ rowptr, const int *__restrict__ col,
                            const double *__restrict__ kij, const double *__restrict__ kji,
                            const double4 *__restrict__ pos, const PairConst *__restrict__ pc,
                            const double *__restrict__ eta, double *__restrict__ s1, double *__restrict__ s2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 pi = pos[i];
    const int ti = (int)pi.w;
    double a = 0.0, b[kMaxTypes];
    for (int t = 0; t < nt; ++t) b[t] = 0.0;
    for (int z = rowptr[i]; z < rowptr[i + 1]; ++z) {
        const int j = col[z];
        const double4 pj = pos[j];
        const int tj = (int)pj.w;
        const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
        const PairConst p = pc[tj + nt * ti];
        double t;
        if (kern == PXH_OHNO)
            t = p.eta_bar / sqrt(1.0 + p.eta_bar * p.eta_bar * r2);
        else if (kern == PXH_COULOMB)
            t = 1.0 / r;
        else
            t = erf(r * p.inv_rqq) / r;
        a += kij[z] * (t - eta[ti]);
        b[tj] += kji[z] * (t - eta[ti]);
    }
    s1[i] = a;
    for (int t = 0; t < nt; ++t) s2[t + (size_t)nt * i] = b[t];
}

// dinv = 1 / diag(A(omega)): q rows diag(c T_qq) - z_q, mu rows Lambda.
__global__ void k_precond_diag(int n, int nt, int fmu, int hetero, const double4 *__restrict__ pos,
                               const double *__restrict__ s1, const double *__restrict__ s2,
                               const double *__restrict__ nbfrac, const SysParam *__restrict__ par,
                               double2 *__restrict__ dinv) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int nv = fmu ? 4 * n : n;
    if (r >= nv) return;
    double2 d;
    if (r < n) {
        const int ti = (int)pos[r].w;
        d = make_double2(s1[r], 0.0);
        if (hetero) {
            double re = 0.5 * s1[r], im = 0.0;
            for (int t = 0; t < nt; ++t) {
                const double2 wr = par->wratio[ti + nt * t];
                re += 0.5 * wr.x * s2[t + (size_t)nt * r];
                im += 0.5 * wr.y * s2[t + (size_t)nt * r];
            }
            d = make_double2(re, im);
        }
        d.x -= par->zq[ti].x;
        d.y -= par->zq[ti].y;
    } else {
        const int i = (r - n) / 3;
        d = make_double2(0, 0);
        for (int t = 0; t < nt; ++t) {
            d.x += nbfrac[t + (size_t)nt * i] * par->inv_alpha[t].x;
            d.y += nbfrac[t + (size_t)nt * i] * par->inv_alpha[t].y;
        }
    }
    const double den = d.x * d.x + d.y * d.y;
    dinv[r] = den > 0.0 ? make_double2(d.x / den, -d.y / den) : make_double2(1.0, 0.0);
}

// Dynamic RHS, q rows: rhs(i,k) = -sum_j c_ij (V_k(j) - V_k(i)), c_ij as in k_kbar_q.
__global__ __launch_bounds__(64) void k_rhs_q(int n, int nt, int hetero, const int *__restrict__ rowptr,
                                              const int *__restrict__ col, const double *__restrict__ kij,
                                              const double *__restrict__ kji, const double4 *__restrict__ pos,
                                              SysParam par, double e0, int potential, double3 mn, int nv,
                                              double2 *__restrict__ rhs) {
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const double4 pi = pos[i];
    const int ti = (int)pi.w;
    auto V = [&](const double4 &p, int k) {
        const double c = k == 0 ? p.x : (k == 1 ? p.y : p.z);
        const double m = k == 0 ? mn.x : (k == 1 ? mn.y : mn.z);
        return potential ? -e0 * fabs(c - m) : -e0 * c;
    };
    double acc_re[3] = {0, 0, 0}, acc_im[3] = {0, 0, 0};
    for (int z = rowptr[i] + lane; z < rowptr[i + 1]; z += 64) {
        const int j = col[z];
        const double4 pj = pos[j];
        double2 c = make_double2(kij[z], 0.0);
        if (hetero) {
            const double2 wr = par.wratio[ti + nt * (int)pj.w];
            c = make_double2(0.5 * (kij[z] + wr.x * kji[z]), 0.5 * wr.y * kji[z]);
        }
        for (int k = 0; k < 3; ++k) {
            const double dv = V(pj, k) - V(pi, k);
            acc_re[k] -= c.x * dv;
            acc_im[k] -= c.y * dv;
        }
    }
    for (int k = 0; k < 3; ++k) {
        double re = acc_re[k], im = acc_im[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            re += __shfl_down(re, off, 64);
            im += __shfl_down(im, off, 64);
        }
        if (lane == 0) rhs[(size_t)k * nv + i] = make_double2(re, im);
    }
}

__global__ void k_rhs_mu(int n, int nv, double e0, double2 *__restrict__ rhs) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 3 * n) return;
    const int a = r % 3;
    for (int k = 0; k < 3; ++k) rhs[(size_t)k * nv + n + r] = make_double2(k == a ? e0 : 0.0, 0.0);
}

// Dense T (n_var x n_var, column-major): T_qq, T_qmu, T_muq, T_mumu blocks.
__global__ void k_dense_t(int n, int nt, int fmu, int kern, const double4 *__restrict__ pos,
                          const PairConst *__restrict__ pc, const double *__restrict__ eta, int nv,
                          double *__restrict__ T) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // row atom
    const int j = blockIdx.y;                              // column atom
    if (i >= n) return;
    const double4 pi = pos[i], pj = pos[j];
    const int ti = (int)pi.w, tj = (int)pj.w;
    const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const PairConst p = pc[ti + nt * tj];
    double tqq;
    if (kern == 3)
        tqq = p.eta_bar / sqrt(1.0 + p.eta_bar * p.eta_bar * r2);
    else if (i == j)
        tqq = 0.5 * (eta[ti] + eta[tj]);
    else if (kern == 2)
        tqq = 1.0 / sqrt(r2);
    else
        tqq = erf(sqrt(r2) * p.inv_rqq) / sqrt(r2);
    T[i + (size_t)nv * j] = tqq;
    if (!fmu) return;
    double d[3] = {dx, dy, dz};
    double g1 = 0, g2 = 0, c0 = 0, c1 = 0;
    if (i != j) {
        const double r = sqrt(r2), ir = 1.0 / r, ir3 = ir * ir * ir;
        const double a1 = r * p.inv_rqmu, e1 = exp(-a1 * a1);
        g1 = (erf(a1) - kTwoOverSqrtPi * a1 * e1) * ir3;
        const double a2 = r * pc[tj + nt * ti].inv_rqmu, e2 = exp(-a2 * a2);
        g2 = (erf(a2) - kTwoOverSqrtPi * a2 * e2) * ir3;
        const double a3 = r * p.inv_rmumu, e3 = exp(-a3 * a3);
        const double fd1 = erf(a3) - kTwoOverSqrtPi * a3 * e3;
        const double fd2 = kFourOverSqrtPi * p.inv_rmumu * p.inv_rmumu * p.inv_rmumu * e3;
        c0 = fd1 * ir3;
        c1 = fd2 * ir * ir - 3.0 * fd1 * ir3 * ir * ir;
    }
    for (int a = 0; a < 3; ++a) {
        T[i + (size_t)nv * (n + 3 * j + a)] = g1 * d[a];         // T_qmu(i, j)
        T[(n + 3 * i + a) + (size_t)nv * j] = -g2 * d[a];        // T_muq(i, j) = T_qmu(j, i)
        for (int b = 0; b < 3; ++b)
            T[(n + 3 * i + a) + (size_t)nv * (n + 3 * j + b)] = c1 * d[a] * d[b] + (a == b ? c0 : 0.0);
    }
}

// M(q rows) = K-bar T(q rows):  M(i, c) = sum_j K_ij (T(j, c) - T(i, c)).
// M[i][c] = sum_z K_iz (B[z][c] - B[i][c]) for the q rows i, with B = T_qq
// (c < n) or T_muq^T (c >= n), i.e. B[z][c] = T[c][z].  A workgroup of 256
// computes a 64 x 64 tile: wave w takes rows i0 + w + 4k (its K-bar rows read
// uniformly), its lanes consecutive columns (contiguous reads of T), and the
// tile goes out through LDS with consecutive lanes on consecutive rows
// (contiguous writes of column-major M).
__global__ __launch_bounds__(256) void k_kbar_dense(int n, const int *__restrict__ rowptr,
                                                    const int *__restrict__ col, const double *__restrict__ kij,
                                                    int nv, const double *__restrict__ T, double *__restrict__ M) {
    __shared__ double tile[64][65];  // [row][column]
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int c0 = blockIdx.x * 64, i0 = blockIdx.y * 64;
    const int c = c0 + lane;
    for (int k = 0; k < 16; ++k) {
        const int ii = w + 4 * k, i = i0 + ii;
        double acc = 0.0;
        if (i < n && c < nv) {
            const double tic = T[c + (size_t)nv * i];
            for (int z = rowptr[i]; z < rowptr[i + 1]; ++z) acc += kij[z] * (T[c + (size_t)nv * col[z]] - tic);
        }
        tile[ii][lane] = acc;
    }
    __syncthreads();
    for (int k = 0; k < 16; ++k) {
        const int cc = w + 4 * k, i = i0 + lane;
        if (i < n && c0 + cc < nv) M[i + (size_t)nv * (c0 + cc)] = tile[lane][cc];
    }
}

__global__ void k_copy_mu_rows(int n, int nv, const double *__restrict__ T, double *__restrict__ M) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;  // mu row index 0..3n-1
    if (r >= 3 * n) return;
    for (int c = blockIdx.y; c < nv; c += gridDim.y) M[(n + r) + (size_t)nv * c] = T[(n + r) + (size_t)nv * c];
}

template <typename R, int KV, bool FMU>
void launch_tsweep_k(int kern, dim3 grid, hipStream_t st, int n, int nt, int hetero, const double4 *pos,
                     const PairConst *pc, int nsys, const int *sys_list, const double2 *x, int ld, double2 *part,
                     int nv, int j_chunk, double rfar2, const int *tperm) {
    switch (kern) {
        case PXH_COULOMB:
            k_tsweep<R, KV, FMU, 2><<<grid, TB, 0, st>>>(n, nt, hetero, pos, pc, nsys, sys_list, x, ld, part, nv,
                                                         j_chunk, rfar2, tperm);
            break;
        case PXH_OHNO:
            k_tsweep<R, KV, FMU, 3><<<grid, TB, 0, st>>>(n, nt, hetero, pos, pc, nsys, sys_list, x, ld, part, nv,
                                                         j_chunk, rfar2, tperm);
            break;
        default:
            k_tsweep<R, KV, FMU, 1><<<grid, TB, 0, st>>>(n, nt, hetero, pos, pc, nsys, sys_list, x, ld, part, nv,
                                                         j_chunk, rfar2, tperm);
    }
}

int choose_splits(const pxh_ctx *c, int n, int groups) {
    if (c->opt.splits > 0) return c->opt.splits;
    const int tiles = (n + TB - 1) / TB;
    // aim at >= 2048 workgroups, each split spanning at least two LDS tiles
    int s = std::max(1, 2048 / std::max(1, tiles * groups));
    const int max_s = std::max(1, (n + 2 * TB - 1) / (2 * TB));
    return std::min(s, max_s);
}

}  // namespace

// Second stage of every product: t = sum of the split partials (+ eta x_q),
// then y_q = K-bar t_q - z_q x_q and y_mu = t_mu + Lambda x_mu.
// part: the split partials [split][row][system], FP64 or (part_f32) FP32
// pairs; tt: room for t [row][system].
void finish_product(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x, double2 *y,
                    int ld, int splits, bool fp32_class, double2 *part, double2 *tt, bool part_f32) {
    Atoms &A = c->atoms;
    const int n = A.n, nv = A.n_var();
    const size_t tot = (size_t)nv * nsys;
    if (part_f32)
        k_reduce_splits<<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(
            n, nv, nsys, splits, reinterpret_cast<const float2 *>(part), d_sys_list, x, ld, A.pos.p, A.d_eta.p, tt);
    else
        k_reduce_splits<<<(unsigned)((tot + 255) / 256), 256, 0, c->stream>>>(n, nv, nsys, splits,
                                                                            (const double2 *)part, d_sys_list, x, ld,
                                                                            A.pos.p, A.d_eta.p, tt);
    const int nblk = (n + 3) / 4;
    const bool loose = fp32_class && A.kbm_built;
    k_kbar_q<<<8 * ((nblk + 7) / 8), 256, 0, c->stream>>>(
        n, A.nt, A.hetero ? 1 : 0, loose ? A.kbm_rowptr.p : A.kb_rowptr.p, loose ? A.kbm_col.p : A.kb_col.p,
        loose ? A.kbm_kij.p : A.kb_kij.p, loose ? A.kbm_kji.p : A.kb_kji.p, A.pos.p, nsys, d_sys_list, d_par, tt, x,
        ld, y);
    if (A.model == PXH_WFQFMU) {
        const size_t tm = (size_t)3 * n * nsys;
        k_mu_rows<<<(unsigned)((tm + 255) / 256), 256, 0, c->stream>>>(n, A.nt, A.d_nbfrac.p, nsys, d_sys_list,
                                                                        d_par, tt, x, ld, y);
    }
    PXH_HIP(hipGetLastError());
}

// Squared far-field radius of the Gaussian kernel: beyond it, for every type
// pair, erfc(kappa r) and (kappa r)^3 exp(-(kappa r)^2) are below half an ulp
// of the working precision, so the damped elements equal their Coulomb
// limits to rounding.  kappa r >= 6 (FP64) / 4 (FP32) for T_qq and >= 7 / 5
// for the charge-dipole and dipole-dipole terms.
double far_field_r2(const Atoms &A, bool fp64) {
    if (A.kernel != PXH_GAUSSIAN) return 0.0;
    const double aq = fp64 ? 6.0 : 4.0, ad = fp64 ? 7.0 : 5.0;
    double r = 0.0;
    for (const PairConst &p : A.h_pair) {
        r = std::max(r, aq / p.inv_rqq);
        if (A.model == PXH_WFQFMU) r = std::max({r, ad / p.inv_rqmu, ad / p.inv_rmumu});
    }
    return r * r;
}

namespace {

// Launch shape of the pair sweep: source atoms per split and the number of
// splits that hold atoms.
struct SweepShape {
    int j_chunk, splits;
};

SweepShape sweep_shape(const pxh_ctx *c, int n, int groups) {
    const int splits = choose_splits(c, n, groups);
    const int j_chunk = ((n + splits - 1) / splits + TB - 1) / TB * TB;
    return {j_chunk, (n + j_chunk - 1) / j_chunk};
}

// Systems per pair-kernel pass of the FP64 product (see apply_atoms_fp64).
int fp64_kv(const Atoms &A, int nsys) {
    if (nsys == 1) return 1;
    if (A.model == PXH_WFQ && nsys >= 8) return 8;
    if (A.model == PXH_WFQFMU && nsys >= 4) return 4;
    return 2;
}

// ----------------------------------------------- FP64 products on matrix cores
// omegaFQFmu T sweep with the contraction on v_mfma_f64_16x16x4f64: for
// target components a and source components b, D_a += A_ab B_b with
// A_ab[m][k] = T_ab(target m, source k) (16 targets x 4 sources) and
// B_b[k][c] = x_b(source k) in column c (8 systems x re / im).  Lane l builds
// the 16 elements of its pair (target l % 16, source 4 s + l / 16) once and
// feeds all 16 products; it holds D_a[l / 16 + 4 v][l % 16], v = 0..3.  The
// elements and their FP64 scalars are those of k_tsweep (same far-field
// radius); the sums run in FP64 in the matrix cores, in a different order.
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int kF64Targets = 64;  // targets per workgroup (16 per wave)
constexpr int kF64Src = 32;      // sources per LDS tile
constexpr int kF64Sys = 8;       // systems per pass (16 real columns)

template <int KERN, bool HET>
__global__ __launch_bounds__(256, 2) void k_tsweep_f64mfma(int n, int nt, const double4 *__restrict__ pos,
                                                          const PairConst *__restrict__ pc, int nsys,
                                                          const int *__restrict__ sys_list,
                                                          const double2 *__restrict__ x, int ld,
                                                          double2 *__restrict__ part, int nv, int j_chunk,
                                                          double rfar2, const int *__restrict__ tperm) {
    __shared__ double4 spos[kF64Src];
    __shared__ double sx[4][kF64Src][2 * kF64Sys];  // [component][source][column]
    __shared__ PairR<double> spc[kMaxTypes * kMaxTypes];
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const int tw = blockIdx.x * kF64Targets + wave * 16;  // first target (Morton position) of the wave
    const int m = lane % 16, kq = lane / 16;
    const int i = tw + m < n ? tperm[tw + m] : -1;        // target of this lane's A rows
    const int split = blockIdx.y;
    const int g0 = blockIdx.z * kF64Sys;
    const int nk = min(kF64Sys, nsys - g0);
    for (int e = tid; e < nt * nt; e += 256) {
        spc[e].inv_rqq = pc[e].inv_rqq;
        spc[e].inv_rqmu = pc[e].inv_rqmu;
        spc[e].inv_rmumu = pc[e].inv_rmumu;
        spc[e].eta_bar = pc[e].eta_bar;
    }
    const double4 pi = i >= 0 ? pos[i] : make_double4(0, 0, 0, 0);
    const int ti = (int)pi.w;
    f64x4 acc[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) acc[a] = f64x4{0, 0, 0, 0};
    const int j_begin = split * j_chunk;
    const int j_end = min(n, j_begin + j_chunk);
    for (int jt = j_begin; jt < j_end; jt += kF64Src) {
        const int cnt = min(kF64Src, j_end - jt);
        __syncthreads();
        if (tid < kF64Src) spos[tid] = tid < cnt ? pos[jt + tid] : make_double4(0, 0, 0, 0);
        // x of the tile: per system kF64Src charges then 3 kF64Src dipole
        // components, both contiguous in x; two systems per pass of 256
        for (int e = tid; e < 4 * kF64Src * kF64Sys; e += 256) {
            const int v = e / (4 * kF64Src), r = e % (4 * kF64Src);
            const int src = r < kF64Src ? r : (r - kF64Src) / 3;
            const int comp = r < kF64Src ? 0 : 1 + (r - kF64Src) % 3;
            double2 val = make_double2(0, 0);
            if (v < nk && src < cnt) {
                const double2 *xs = x + (size_t)sys_list[g0 + v] * ld;
                val = r < kF64Src ? xs[jt + r] : xs[n + 3 * jt + (r - kF64Src)];
            }
            sx[comp][src][2 * v] = val.x;
            sx[comp][src][2 * v + 1] = val.y;
        }
        __syncthreads();
        for (int s = 0; 4 * s < cnt; ++s) {
            const int jl = 4 * s + kq;
            const int j = jt + jl;
            const double4 pj = spos[jl];
            const bool valid = i >= 0 && jl < cnt && j != i;
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const double r2v = dx * dx + dy * dy + dz * dz;
            const double r2 = valid ? r2v : 1.0;
            const double inv_r = d_rsqrt<double>(r2);
            const double r = r2 * inv_r;
            const int tj = (int)pj.w;
            const PairR<double> p = spc[ti + nt * tj];
            const double inv_r2 = inv_r * inv_r;
            const double inv_r3 = inv_r2 * inv_r;
            double tqq, g1, g2, c0, c1;
            if (KERN == 1 && r2 >= rfar2) {
                tqq = inv_r;
                g1 = g2 = c0 = inv_r3;
                c1 = -3.0 * inv_r3 * inv_r2;
            } else {
                if (KERN == 1)
                    tqq = erf(r * p.inv_rqq) * inv_r;
                else if (KERN == 2)
                    tqq = inv_r;
                else
                    tqq = p.eta_bar * d_rsqrt<double>(1.0 + p.eta_bar * p.eta_bar * r2);
                const double a1 = r * p.inv_rqmu;
                const double e1 = exp(-a1 * a1);
                g1 = (erf(a1) - kTwoOverSqrtPi * a1 * e1) * inv_r3;
                g2 = g1;
                if (HET) {
                    const double a2 = r * spc[tj + nt * ti].inv_rqmu;
                    const double e2 = exp(-a2 * a2);
                    g2 = (erf(a2) - kTwoOverSqrtPi * a2 * e2) * inv_r3;
                }
                const double a3 = r * p.inv_rmumu;
                const double e3 = exp(-a3 * a3);
                const double fd1 = erf(a3) - kTwoOverSqrtPi * a3 * e3;
                const double fd2 = kFourOverSqrtPi * p.inv_rmumu * p.inv_rmumu * p.inv_rmumu * e3;
                c0 = fd1 * inv_r3;
                c1 = fd2 * inv_r2 - 3.0 * fd1 * inv_r3 * inv_r2;
            }
            if (!valid) {
                tqq = g1 = g2 = c0 = c1 = 0.0;
                dx = dy = dz = 0.0;
            }
            // rows of the 4x4 block: [tqq, g1 d^T; -g2 d, c1 d d^T + c0 I]
            const double d[3] = {dx, dy, dz};
            double t[4][4];
            t[0][0] = tqq;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                t[0][1 + b] = g1 * d[b];
                t[1 + b][0] = -g2 * d[b];
            }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) t[1 + a][1 + b] = c1 * d[a] * d[b] + (a == b ? c0 : 0.0);
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double xb = sx[b][jl][m];
#pragma unroll
                for (int a = 0; a < 4; ++a) acc[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(t[a][b], xb, acc[a], 0, 0, 0);
            }
        }
    }
    // part[split][row][system]: lane l holds targets tw + l / 16 + 4 v, column l % 16
    double *out = reinterpret_cast<double *>(part + (size_t)split * nv * nsys);
    const int col = lane % 16, sv = col / 2;
    if (sv >= nk) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int it = tw + kq + 4 * v;
        if (it >= n) continue;
        const int ia = tperm[it];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const size_t row = a == 0 ? (size_t)ia : (size_t)n + 3 * (size_t)ia + (a - 1);
            out[2 * (row * nsys + g0 + sv) + (col & 1)] = acc[a][v];
        }
    }
}

// The FP64 omegaFQFmu products of five or more systems run on the matrix
// cores (k_tsweep_f64mfma), fewer on k_tsweep: a pass of the matrix-core
// kernel costs the same for 1 to 8 systems.  Measured at 15 683 atoms (C2
// residual products, ~11 systems each): 4.47 against 4.67 ms per launch; the
// counters show the FP64 MFMA time (64 cycles each) and the VALU time of the
// element generation adding up, not overlapping.
constexpr int kF64MfmaMinSys = 5;
bool use_f64mfma(const pxh_ctx *c, int nsys) {
    return c->opt.fp64_mfma && c->atoms.model == PXH_WFQFMU && nsys >= kF64MfmaMinSys;
}

SweepShape f64mfma_shape(const pxh_ctx *c, int n, int nsys) {
    const int groups = (nsys + kF64Sys - 1) / kF64Sys;
    const int xblocks = (n + kF64Targets - 1) / kF64Targets;
    // >= 2048 workgroups, each split spanning at least two LDS tiles
    int s = c->opt.splits > 0 ? c->opt.splits : std::max(1, 2048 / std::max(1, xblocks * groups));
    s = std::min(s, std::max(1, (n + 2 * kF64Src - 1) / (2 * kF64Src)));
    const int j_chunk = ((n + s - 1) / s + kF64Src - 1) / kF64Src * kF64Src;
    return {j_chunk, (n + j_chunk - 1) / j_chunk};
}

void apply_atoms_f64mfma(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x,
                         double2 *y, int ld, double2 *ext_part, double2 *ext_t) {
    Atoms &A = c->atoms;
    const int n = A.n, nv = A.n_var();
    const int groups = (nsys + kF64Sys - 1) / kF64Sys;
    const SweepShape sh = f64mfma_shape(c, n, nsys);
    double2 *part = ext_part, *tt = ext_t;
    if (!part || !tt) {
        c->work_part.alloc((size_t)sh.splits * nsys * nv);
        c->work_t.alloc((size_t)nsys * nv);
        part = c->work_part.p;
        tt = c->work_t.p;
    }
    const dim3 grid((n + kF64Targets - 1) / kF64Targets, sh.splits, groups);
    const double rfar2 = far_field_r2(A, true);
    const int timer = timer_begin(c);
#define PXH_F64MFMA(KERN, HET)                                                                                      \
    k_tsweep_f64mfma<KERN, HET><<<grid, 256, 0, c->stream>>>(n, A.nt, A.pos.p, A.pair.p, nsys, d_sys_list, x, ld, \
                                                             part, nv, sh.j_chunk, rfar2, A.tperm.p)
    const int kern = A.kernel == PXH_COULOMB ? 2 : A.kernel == PXH_OHNO ? 3 : 1;
    if (A.hetero) {
        if (kern == 2)
            PXH_F64MFMA(2, true);
        else if (kern == 3)
            PXH_F64MFMA(3, true);
        else
            PXH_F64MFMA(1, true);
    } else {
        if (kern == 2)
            PXH_F64MFMA(2, false);
        else if (kern == 3)
            PXH_F64MFMA(3, false);
        else
            PXH_F64MFMA(1, false);
    }
#undef PXH_F64MFMA
    PXH_HIP(hipGetLastError());
    timer_end(c, timer, 4);
    finish_product(c, nsys, d_sys_list, d_par, x, y, ld, sh.splits, false, part, tt, false);
    const double pairs = (double)n * (double)(n - 1);
    c->stats.v[9] += 1;
    c->stats.v[11] += pairs * ((double)kPairEvalFlopsFmu * groups + (double)kPairVecFlopsFmu * nsys);
    c->stats.v[6] += (double)n * (double)n * groups;
}

// ext_part / ext_t: caller-provided room for the split partials and t (both
// null: the context's product buffers, grown as needed).
template <typename R, int KV>
void apply_atoms_impl(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x,
                      double2 *y, int ld, double2 *ext_part = nullptr, double2 *ext_t = nullptr) {
    Atoms &A = c->atoms;
    const int n = A.n, nv = A.n_var();
    const bool fmu = A.model == PXH_WFQFMU;
    const int groups = (nsys + KV - 1) / KV;
    const SweepShape sh = sweep_shape(c, n, groups);
    const int j_chunk = sh.j_chunk, splits_eff = sh.splits;
    double2 *part = ext_part, *tt = ext_t;
    if (!part || !tt) {
        c->work_part.alloc((size_t)splits_eff * nsys * nv);
        c->work_t.alloc((size_t)nsys * nv);
        part = c->work_part.p;
        tt = c->work_t.p;
    }
    dim3 grid((n + TB - 1) / TB, splits_eff, groups);
    const double rfar2 = far_field_r2(A, sizeof(R) == 8);
    const bool f32 = sizeof(R) == 4;
    const int timer = timer_begin(c);
    if (fmu)
        launch_tsweep_k<R, KV, true>(A.kernel, grid, c->stream, n, A.nt, A.hetero ? 1 : 0, A.pos.p, A.pair.p, nsys,
                                     d_sys_list, x, ld, part, nv, j_chunk, rfar2, A.tperm.p);
    else
        launch_tsweep_k<R, KV, false>(A.kernel, grid, c->stream, n, A.nt, 0, A.pos.p, A.pair.p, nsys, d_sys_list,
                                      x, ld, part, nv, j_chunk, rfar2, A.tperm.p);
    PXH_HIP(hipGetLastError());
    // time of the pair kernel on the context's stream, and its algorithmic
    // flops: per ordered pair (i != j) kPairEvalFlops for the interaction
    // elements plus kPairVecFlops per complex vector for the contraction
    timer_end(c, timer, f32 ? 3 : 4);
    finish_product(c, nsys, d_sys_list, d_par, x, y, ld, splits_eff, f32, part, tt, false);
    const double pairs = (double)n * (double)(n - 1);
    const double fl = pairs * ((double)(fmu ? kPairEvalFlopsFmu : kPairEvalFlopsQ) * groups +
                               (double)(fmu ? kPairVecFlopsFmu : kPairVecFlopsQ) * nsys);
    c->stats.v[f32 ? 8 : 9] += 1;
    c->stats.v[f32 ? 10 : 11] += fl;
    c->stats.v[6] += (double)n * (double)n * groups;
}

// part[row][k] = sum_c T[row, c] x_k[c] with the stored T (q-q diagonal zero).
__global__ void k_stored_t(int nv, const double *__restrict__ T, int nsys, const int *__restrict__ sys_list,
                           const double2 *__restrict__ x, int ld, double2 *__restrict__ part) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (row >= nv) return;
    const double2 *xs = x + (size_t)sys_list[k] * ld;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < nv; ++c) {
        const double t = T[row + (size_t)nv * c];
        re += t * xs[c].x;
        im += t * xs[c].y;
    }
    part[(size_t)row * nsys + k] = make_double2(re, im);
}

}  // namespace

// Systems per pass: omegaFQ (charges only) has room for 8; omegaFQFmu shares
// each pair block over 4 (two waves per SIMD).
void apply_atoms_fp64(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x,
                      double2 *y, int ld, double2 *ext_part, double2 *ext_t) {
    if (nsys <= 0) return;
    if (use_f64mfma(c, nsys)) {
        apply_atoms_f64mfma(c, nsys, d_sys_list, d_par, x, y, ld, ext_part, ext_t);
        return;
    }
    switch (fp64_kv(c->atoms, nsys)) {
        case 1:
            apply_atoms_impl<double, 1>(c, nsys, d_sys_list, d_par, x, y, ld, ext_part, ext_t);
            break;
        case 8:
            apply_atoms_impl<double, 8>(c, nsys, d_sys_list, d_par, x, y, ld, ext_part, ext_t);
            break;
        case 4:
            apply_atoms_impl<double, 4>(c, nsys, d_sys_list, d_par, x, y, ld, ext_part, ext_t);
            break;
        default:
            apply_atoms_impl<double, 2>(c, nsys, d_sys_list, d_par, x, y, ld, ext_part, ext_t);
    }
}

size_t fp64_product_partials(const pxh_ctx *c, int nsys) {
    const Atoms &A = c->atoms;
    if (use_f64mfma(c, nsys)) return (size_t)f64mfma_shape(c, A.n, nsys).splits * nsys * A.n_var();
    const int kv = fp64_kv(A, nsys);
    const SweepShape sh = sweep_shape(c, A.n, (nsys + kv - 1) / kv);
    return (size_t)sh.splits * nsys * A.n_var();
}

void apply_atoms_fp32(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x,
                      double2 *y, int ld) {
    if (nsys <= 0) return;
    const Atoms &A = c->atoms;
    if (c->opt.mfma && A.model == PXH_WFQFMU && A.kernel == PXH_GAUSSIAN) {
        apply_atoms_mfma(c, nsys, d_sys_list, d_par, x, y, ld);
        return;
    }

    if (nsys == 1)
        apply_atoms_impl<float, 1>(c, nsys, d_sys_list, d_par, x, y, ld);
    else if (nsys == 2)
        apply_atoms_impl<float, 2>(c, nsys, d_sys_list, d_par, x, y, ld);
    else
        apply_atoms_impl<float, 4>(c, nsys, d_sys_list, d_par, x, y, ld);
}

SysParam make_sysparam(const pxh_ctx *c, double freq, int i_freq) {
    const Atoms &A = c->atoms;
    SysParam p{};
    for (int t = 0; t < A.nt; ++t) {
        // z_q(omega) = -(omega / (2 n0 tau)) (omega tau + i)   (wfq_module.f90:239-242)
        const double f = -(freq / (2.0 * A.dens[t] * A.tau[t]));
        p.zq[t] = make_double2(f * freq * A.tau[t], f);
        double2 al = make_double2(1.0, 0.0);
        if (A.model == PXH_WFQFMU && i_freq >= 0 && i_freq < c->alpha_nf) {
            const size_t o = 2 * ((size_t)t + (size_t)A.nt * i_freq);
            al = make_double2(c->alpha[o], c->alpha[o + 1]);
        }
        const double den = al.x * al.x + al.y * al.y;
        p.inv_alpha[t] = make_double2(al.x / den, -al.y / den);
    }
    // w_t = 2 n_t / (1/tau_t - i omega)   (matrix_module.f90:455-466)
    double wre[kMaxTypes], wim[kMaxTypes];
    for (int t = 0; t < A.nt; ++t) {
        const double a = 1.0 / A.tau[t], b = -freq;
        const double den = a * a + b * b;
        wre[t] = 2.0 * A.dens[t] * a / den;
        wim[t] = -2.0 * A.dens[t] * b / den;
    }
    for (int ti = 0; ti < A.nt; ++ti)
        for (int tj = 0; tj < A.nt; ++tj) {
            const double den = wre[ti] * wre[ti] + wim[ti] * wim[ti];
            const double re = (wre[tj] * wre[ti] + wim[tj] * wim[ti]) / den;
            const double im = (wim[tj] * wre[ti] - wre[tj] * wim[ti]) / den;
            p.wratio[ti + A.nt * tj] = make_double2(re, im);
        }
    return p;
}

void build_atoms_device(pxh_ctx *c) {
    Atoms &A = c->atoms;
    A.Mconst_valid = false;
    // padded to a multiple of 128 atoms with atoms 1e15 bohr away (type 0):
    // the MFMA pair kernel reads whole blocks of targets and sources, and the
    // interactions with these atoms round to zero
    const int n_pad = (A.n + 127) / 128 * 128;
    std::vector<double4> hp(n_pad, make_double4(1e15, 1e15, 1e15, 0.0));
    for (int i = 0; i < A.n; ++i)
        hp[i] = make_double4(A.h_coord[3 * i], A.h_coord[3 * i + 1], A.h_coord[3 * i + 2], (double)A.h_type[i]);
    A.pos.upload(hp.data(), hp.size(), c->stream);
    // target order of the VALU pair kernels: Morton order of 2-bohr cells
    {
        double lo[3] = {1e300, 1e300, 1e300};
        for (int i = 0; i < A.n; ++i)
            for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], A.h_coord[3 * i + d]);
        auto spread = [](uint64_t v) {
            uint64_t r = 0;
            for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ull) << (3 * b);
            return r;
        };
        std::vector<std::pair<uint64_t, int>> key(A.n);
        for (int i = 0; i < A.n; ++i) {
            uint64_t code = 0;
            for (int d = 0; d < 3; ++d) {
                const double cell = std::floor((A.h_coord[3 * i + d] - lo[d]) / 2.0);
                code |= spread((uint64_t)std::min(cell, 2097151.0)) << d;
            }
            key[i] = {code, i};
        }
        std::stable_sort(key.begin(), key.end(),
                         [](const std::pair<uint64_t, int> &a, const std::pair<uint64_t, int> &b) {
                             return a.first < b.first;
                         });
        std::vector<int> hperm(A.n);
        for (int i = 0; i < A.n; ++i) hperm[i] = key[i].second;
        A.tperm.upload(hperm.data(), hperm.size(), c->stream);
    }
    A.h_pair.resize((size_t)A.nt * A.nt);
    for (int ti = 0; ti < A.nt; ++ti)
        for (int tj = 0; tj < A.nt; ++tj) {
            PairConst &p = A.h_pair[ti + A.nt * tj];
            p.inv_rqq = 1.0 / std::sqrt(A.rq[ti] * A.rq[ti] + A.rq[tj] * A.rq[tj]);
            p.inv_rqmu = 1.0 / std::sqrt(A.rq[ti] * A.rq[ti] + A.rmu[tj] * A.rmu[tj]);
            p.inv_rmumu = 1.0 / std::sqrt(A.rmu[ti] * A.rmu[ti] + A.rmu[tj] * A.rmu[tj]);
            p.eta_bar = 0.5 * (A.eta[ti] + A.eta[tj]);
        }
    A.pair.upload(A.h_pair.data(), A.h_pair.size(), c->stream);
    A.d_eta.upload(A.eta.data(), A.eta.size(), c->stream);
    if (A.h_nbfrac.empty()) {
        A.h_nbfrac.assign((size_t)A.nt * A.n, 0.0);
        for (int i = 0; i < A.n; ++i) A.h_nbfrac[A.h_type[i] + (size_t)A.nt * i] = 1.0;
    }
    A.d_nbfrac.upload(A.h_nbfrac.data(), A.h_nbfrac.size(), c->stream);
    if (!A.is_static) build_kbar(c);  // the static models have no K-bar
}

namespace {

// CSR of K-bar keeping the pairs with 1 - f_ij >= cut for either direction
// (cut = 0: every pair).  1 - f < cut beyond x = 1 + ln(1/cut)/d, x = r / (s r0).
void build_kbar_csr(pxh_ctx *c, double cut, DevBuf<int> &rowptr, DevBuf<int> &col, DevBuf<double> &kij,
                    DevBuf<double> &kji, size_t &nnz) {
    Atoms &A = c->atoms;
    FermiTab ft{};
    for (int t = 0; t < A.nt; ++t) ft.a[t] = A.aij[t];
    for (int e = 0; e < A.nt * A.nt; ++e) {
        ft.d[e] = A.fd[e];
        ft.s[e] = A.fs[e];
        ft.r0[e] = A.r0[e];
        double rc = 1e300;
        if (cut > 0.0 && A.fd[e] > 0.0) rc = (1.0 + std::log(1.0 / cut) / A.fd[e]) * A.fs[e] * A.r0[e];
        ft.rcut2[e] = rc >= 1e150 ? 1e300 : rc * rc;
    }
    DevBuf<int> cnt;
    cnt.alloc(A.n);
    const int blk = 128;
    k_kbar_count<<<(A.n + blk - 1) / blk, blk, 0, c->stream>>>(A.n, A.nt, A.pos.p, ft, cnt.p);
    PXH_HIP(hipGetLastError());
    std::vector<int> hc(A.n), rp(A.n + 1, 0);
    PXH_HIP(hipMemcpyAsync(hc.data(), cnt.p, sizeof(int) * A.n, hipMemcpyDeviceToHost, c->stream));
    PXH_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < A.n; ++i) rp[i + 1] = rp[i] + hc[i];
    nnz = (size_t)rp[A.n];
    rowptr.upload(rp.data(), rp.size(), c->stream);
    col.alloc(std::max<size_t>(1, nnz));
    kij.alloc(std::max<size_t>(1, nnz));
    kji.alloc(std::max<size_t>(1, nnz));
    k_kbar_fill<<<(A.n + blk - 1) / blk, blk, 0, c->stream>>>(A.n, A.nt, A.pos.p, ft, rowptr.p, col.p, kij.p,
                                                              kji.p);
    PXH_HIP(hipGetLastError());
    PXH_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace

// K-bar for the FP64 products: the pairs with 1 - f < 1e-18 add less than
// 1e-18 of the nearest-neighbour coupling and are dropped.  The FP32-class
// products of the inner solves use a second CSR cut at 2^-30 (below their
// own rounding), about a third of the entries.
void build_kbar(pxh_ctx *c) {
    Atoms &A = c->atoms;
    A.Mconst_valid = false;
    const bool cut = c->opt.kbar_cutoff != 0;
    build_kbar_csr(c, cut ? 1e-18 : 0.0, A.kb_rowptr, A.kb_col, A.kb_kij, A.kb_kji, A.kb_nnz);
    if (cut)
        build_kbar_csr(c, std::ldexp(1.0, -30), A.kbm_rowptr, A.kbm_col, A.kbm_kij, A.kbm_kji, A.kbm_nnz);
    A.kbm_built = cut;
    const int blk = 128;
    A.diag_s1.alloc(A.n);
    A.diag_s2.alloc((size_t)A.n * A.nt);
    k_diag_sums<<<(A.n + blk - 1) / blk, blk, 0, c->stream>>>(A.n, A.nt, A.kernel, A.kb_rowptr.p, A.kb_col.p,
                                                              A.kb_kij.p, A.kb_kji.p, A.pos.p, A.pair.p, A.d_eta.p,
                                                              A.diag_s1.p, A.diag_s2.p);
    PXH_HIP(hipGetLastError());
    PXH_HIP(hipStreamSynchronize(c->stream));
}

void precond_diag_atoms(pxh_ctx *c, const SysParam *d_par_one, double2 *dinv) {
    Atoms &A = c->atoms;
    const int nv = A.n_var();
    k_precond_diag<<<(nv + 255) / 256, 256, 0, c->stream>>>(A.n, A.nt, A.model == PXH_WFQFMU, A.hetero ? 1 : 0,
                                                            A.pos.p, A.diag_s1.p, A.diag_s2.p, A.d_nbfrac.p,
                                                            d_par_one, dinv);
    PXH_HIP(hipGetLastError());
}

void dynamic_rhs_atoms(pxh_ctx *c, double freq, double e0, int rhs_form, double2 *d_rhs) {
    Atoms &A = c->atoms;
    const int n = A.n, nv = A.n_var();
    SysParam p = make_sysparam(c, freq, -1);
    double3 mn = make_double3(0, 0, 0);
    if (rhs_form == PXH_RHS_POTENTIAL) {
        double m[3] = {A.h_coord[0], A.h_coord[1], A.h_coord[2]};
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) m[k] = std::min(m[k], A.h_coord[3 * i + k]);
        mn = make_double3(m[0], m[1], m[2]);
    }
    PXH_HIP(hipMemsetAsync(d_rhs, 0, sizeof(double2) * (size_t)nv * 3, c->stream));
    k_rhs_q<<<n, 64, 0, c->stream>>>(n, A.nt, A.hetero ? 1 : 0, A.kb_rowptr.p, A.kb_col.p, A.kb_kij.p, A.kb_kji.p,
                                     A.pos.p, p, e0, rhs_form == PXH_RHS_POTENTIAL ? 1 : 0, mn, nv, d_rhs);
    if (A.model == PXH_WFQFMU) k_rhs_mu<<<(3 * n + 255) / 256, 256, 0, c->stream>>>(n, nv, e0, d_rhs);
    PXH_HIP(hipGetLastError());
}

__global__ void k_zero_qq_diag(int n, int nv, double *__restrict__ T) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) T[i + (size_t)nv * i] = 0.0;
}

// Dense T (n_var x n_var); with zero_qq_diag the q-q diagonal (eta) is left
// out so that the shared second stage adds it, as for the on-the-fly sweep.
void dense_t_atoms(pxh_ctx *c, double *d_T, bool zero_qq_diag) {
    Atoms &A = c->atoms;
    const int n = A.n, nv = A.n_var();
    k_dense_t<<<dim3((n + 127) / 128, n), 128, 0, c->stream>>>(n, A.nt, A.model == PXH_WFQFMU, A.kernel, A.pos.p,
                                                               A.pair.p, A.d_eta.p, nv, d_T);
    if (zero_qq_diag) k_zero_qq_diag<<<(n + 255) / 256, 256, 0, c->stream>>>(n, nv, d_T);
    PXH_HIP(hipGetLastError());
}

// y = A(omega) x with the stored T (the 'iterative' in-memory method,
// solve_w_iter_mem algorithm_module.f90:334-410): t = T x, then K-bar and
// the shifts.  T must have been built by prepare_stored_t.
void prepare_stored_t(pxh_ctx *c) {
    Atoms &A = c->atoms;
    const size_t nv = A.n_var();
    A.Tstore.alloc(nv * nv);
    dense_t_atoms(c, A.Tstore.p, true);
    PXH_HIP(hipStreamSynchronize(c->stream));
}

void apply_atoms_stored(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x,
                        double2 *y, int ld) {
    if (nsys <= 0) return;
    Atoms &A = c->atoms;
    const int nv = A.n_var();
    c->work_part.alloc((size_t)nsys * nv);
    c->work_t.alloc((size_t)nsys * nv);
    k_stored_t<<<dim3((nv + 255) / 256, nsys), 256, 0, c->stream>>>(nv, A.Tstore.p, nsys, d_sys_list, x, ld,
                                                                    c->work_part.p);
    PXH_HIP(hipGetLastError());
    finish_product(c, nsys, d_sys_list, d_par, x, y, ld, 1, false, c->work_part.p, c->work_t.p, false);
}

void constant_matrix_atoms(pxh_ctx *c, double *d_m) {
    Atoms &A = c->atoms;
    const int n = A.n, nv = A.n_var();
    DevBuf<double> T;
    T.alloc((size_t)nv * nv);
    dense_t_atoms(c, T.p, false);
    PXH_HIP(hipMemsetAsync(d_m, 0, sizeof(double) * (size_t)nv * nv, c->stream));
    const int gy = std::min(nv, 65535);
    if (!A.hetero)
        k_kbar_dense<<<dim3((nv + 63) / 64, (n + 63) / 64), 256, 0, c->stream>>>(n, A.kb_rowptr.p, A.kb_col.p,
                                                                                  A.kb_kij.p, nv, T.p, d_m);
    if (A.model == PXH_WFQFMU)
        k_copy_mu_rows<<<dim3((3 * n + 127) / 128, gy), 128, 0, c->stream>>>(n, nv, T.p, d_m);
    PXH_HIP(hipGetLastError());
    PXH_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace pxh
