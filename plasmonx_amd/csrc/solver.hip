// Per-frequency complex solves on gfx950: the reference's GMRES(m)
// (src/solvers/gmres_module.f90:30-228) for one system at a time in FP64,
// and a batched GMRES with FP64 iterative refinement in which many
// (frequency, field) systems share every operator application.
//
// Batched layout in HBM: slot-major.  Slot s owns
//   X[s], B[s], R[s], W[s]            n_var complex each,
//   V[s][0..m]                        the Krylov basis (FP64; complex FP32 in
//                                     the inner solves of the mixed path),
//   H[s] (m+1) x m, g[s] (m+1), cs[s], sn[s] (m)   Hessenberg and rotations.
// Dots and norms are reduced in a fixed order (one workgroup per dot), so
// a system's result does not depend on which other systems share the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "pxh_internal.h"

namespace pxh {

namespace {

constexpr int RB = 256;  // threads of the reduction workgroups

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconjmul(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Krylov basis element: FP64 (double2), or complex FP32 (float2) for the
// inner solves of the mixed path, whose operator is FP32-class anyway.
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2(v.x, v.y); }
template <typename VT>
__device__ __forceinline__ VT from_d2(double2 v);
template <>
__device__ __forceinline__ double2 from_d2<double2>(double2 v) {
    return v;
}
template <>
__device__ __forceinline__ float2 from_d2<float2>(double2 v) {
    return make_float2((float)v.x, (float)v.y);
}

__device__ __forceinline__ double2 block_sum2(double2 v) {
    __shared__ double sre[RB / 64], sim[RB / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_down(v.x, off, 64);
        v.y += __shfl_down(v.y, off, 64);
    }
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
        sre[w] = v.x;
        sim[w] = v.y;
    }
    __syncthreads();
    double2 r = make_double2(0, 0);
    if (threadIdx.x == 0)
        for (int k = 0; k < RB / 64; ++k) {
            r.x += sre[k];
            r.y += sim[k];
        }
    return r;
}

// h[s][i] = <V[s][i], w[s]> for i in [i0, jcount[s]) ; one workgroup per (i, slot).
__global__ __launch_bounds__(RB) void k_dots(int nv, const int *__restrict__ slots, const int *__restrict__ jcount,
                                             int i0, const double2 *__restrict__ V, size_t ldv_slot,
                                             const double2 *__restrict__ W, double2 *__restrict__ h, int ldh) {
    const int i = i0 + blockIdx.x;
    const int s = slots[blockIdx.y];
    if (i >= jcount[s]) return;
    const double2 *v = V + ldv_slot * s + (size_t)nv * i;
    const double2 *w = W + (size_t)nv * s;
    double2 acc = make_double2(0, 0);
    for (int r = threadIdx.x; r < nv; r += RB) {
        const double2 p = cconjmul(v[r], w[r]);
        acc.x += p.x;
        acc.y += p.y;
    }
    acc = block_sum2(acc);
    if (threadIdx.x == 0) h[(size_t)s * ldh + i] = acc;
}

// Batched dots in two stages: one wavefront per (slot, chunk of kDotRows rows)
// computes the chunk's partial sums for every basis vector, k_dots_sum adds
// the chunks in a fixed order (deterministic), W is read once per chunk.
constexpr int kDotRows = 512;
constexpr int kDotGroup = 8;  // basis vectors per pass over the chunk

// part[(s ldh + i) nch + ch] = sum_{r in chunk ch} conj(V[s][i][r]) W[s][r], i < jcount[s]
template <typename VT>
__global__ __launch_bounds__(256) void k_dots_part(int nv, const int *__restrict__ slots,
                                                   const int *__restrict__ jcount, const VT *__restrict__ V,
                                                   size_t ldv_slot, const double2 *__restrict__ W,
                                                   double2 *__restrict__ part, int ldh, int nch,
                                                   double *__restrict__ wn_part) {
    constexpr int U = kDotRows / 64;
    const int ch = blockIdx.x * 4 + threadIdx.x / 64;
    const int s = slots[blockIdx.y];
    const int lane = threadIdx.x % 64;
    if (ch >= nch) return;
    const int jc = jcount[s];
    const int r0 = ch * kDotRows + lane;
    double2 w[U];
    const double2 *ws = W + (size_t)nv * s;
#pragma unroll
    for (int u = 0; u < U; ++u) w[u] = r0 + 64 * u < nv ? ws[r0 + 64 * u] : make_double2(0, 0);
    if (wn_part && blockIdx.z == 0) {  // chunk partial of ||W[s]||^2
        double q = 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u) q += w[u].x * w[u].x + w[u].y * w[u].y;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off, 64);
        if (lane == 0) wn_part[(size_t)s * nch + ch] = q;
    }
    {
        const int i0 = blockIdx.z * kDotGroup;  // one group of basis vectors per workgroup layer
        if (i0 >= jc) return;
        double2 acc[kDotGroup];
#pragma unroll
        for (int g = 0; g < kDotGroup; ++g) {
            acc[g] = make_double2(0, 0);
            if (i0 + g < jc) {
                const VT *v = V + ldv_slot * s + (size_t)nv * (i0 + g);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (r0 + 64 * u < nv) {
                        const double2 p = cconjmul(to_d2(v[r0 + 64 * u]), w[u]);
                        acc[g].x += p.x;
                        acc[g].y += p.y;
                    }
                }
            }
        }
        // transposing reduction of the 8 sums over the wave: each exchange
        // halves the values a lane holds (10 shuffles instead of 48); lane l
        // ends with the sum of g = 4 b5 + 2 b4 + b3 (bits of l) when l % 8 == 0
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bool up = lane & 32;
            const double2 snd = up ? acc[q] : acc[q + 4];
            const double2 keep = up ? acc[q + 4] : acc[q];
            acc[q] = make_double2(keep.x + __shfl_xor(snd.x, 32, 64), keep.y + __shfl_xor(snd.y, 32, 64));
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const bool up = lane & 16;
            const double2 snd = up ? acc[q] : acc[q + 2];
            const double2 keep = up ? acc[q + 2] : acc[q];
            acc[q] = make_double2(keep.x + __shfl_xor(snd.x, 16, 64), keep.y + __shfl_xor(snd.y, 16, 64));
        }
        {
            const bool up = lane & 8;
            const double2 snd = up ? acc[0] : acc[1];
            const double2 keep = up ? acc[1] : acc[0];
            acc[0] = make_double2(keep.x + __shfl_xor(snd.x, 8, 64), keep.y + __shfl_xor(snd.y, 8, 64));
        }
#pragma unroll
        for (int off = 4; off > 0; off >>= 1) {
            acc[0].x += __shfl_xor(acc[0].x, off, 64);
            acc[0].y += __shfl_xor(acc[0].y, off, 64);
        }
        const int g = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
        if ((lane & 7) == 0 && i0 + g < jc) part[((size_t)s * ldh + i0 + g) * nch + ch] = acc[0];
    }
}

// h[s][i] = sum_ch part[s][i][ch]; one wavefront per (i, slot), fixed reduction tree
__global__ __launch_bounds__(256) void k_dots_sum(const int *__restrict__ slots, const int *__restrict__ jcount,
                                                  const double2 *__restrict__ part, int ldh, int nch,
                                                  double2 *__restrict__ h) {
    const int i = blockIdx.x * 4 + threadIdx.x / 64;
    const int s = slots[blockIdx.y];
    const int lane = threadIdx.x % 64;
    if (i >= jcount[s]) return;
    const double2 *p = part + ((size_t)s * ldh + i) * nch;
    double2 acc = make_double2(0, 0);
    for (int ch = lane; ch < nch; ch += 64) {
        acc.x += p[ch].x;
        acc.y += p[ch].y;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        acc.x += __shfl_xor(acc.x, off, 64);
        acc.y += __shfl_xor(acc.y, off, 64);
    }
    if (lane == 0) h[(size_t)s * ldh + i] = acc;
}

// out[s] = sum_b part[s][b] (sqrt when root); one wavefront per slot, only
// slots with gate[s] != 0 when gate is given
__global__ __launch_bounds__(64) void k_sum_slot(const int *__restrict__ slots, const double *__restrict__ part,
                                                 int nb, const int *__restrict__ gate, int root,
                                                 double *__restrict__ out) {
    const int s = slots[blockIdx.x];
    if (gate && gate[s] == 0) return;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += 64) acc += part[(size_t)s * nb + b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (threadIdx.x == 0) out[s] = root ? sqrt(acc) : acc;
}

// After the first Gram-Schmidt pass: ||w'|| from the update's partials, and
// the second pass only where the first removed more than (1 - kappa) of the
// norm (||w'|| < kappa ||w||, "twice is enough"); kappa >= 1: always.
__global__ __launch_bounds__(64) void k_reorth_decide(const int *__restrict__ slots, const int *__restrict__ jcount,
                                                      const double *__restrict__ npart, int nb,
                                                      const double *__restrict__ wn2, double kappa,
                                                      int *__restrict__ jc2, double *__restrict__ nrm) {
    const int s = slots[blockIdx.x];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += 64) acc += npart[(size_t)s * nb + b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (threadIdx.x == 0) {
        const bool again = kappa >= 1.0 || acc < kappa * kappa * wn2[s];
        jc2[s] = again ? jcount[s] : 0;
        nrm[s] = sqrt(acc);
    }
}

// W[s] -= sum_{i < jcount[s]} V[s][i] h[s][i]; with npart, also the
// per-workgroup partial sums of ||W[s]||^2 after the update.
template <typename VT>
__global__ __launch_bounds__(256) void k_gs_update(int nv, const int *__restrict__ slots,
                                                   const int *__restrict__ jcount, const VT *__restrict__ V,
                                                   size_t ldv_slot, const double2 *__restrict__ h, int ldh,
                                                   double2 *__restrict__ W, double *__restrict__ npart) {
    __shared__ double sq[4];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slots[blockIdx.y];
    const int jc = jcount[s];
    if (jc == 0) return;
    double q = 0.0;
    if (r < nv) {
        const VT *v = V + ldv_slot * s + r;
        const double2 *hs = h + (size_t)s * ldh;
        double2 acc = make_double2(0, 0);
        for (int i = 0; i < jc; ++i) {
            const double2 p = cmul(to_d2(v[(size_t)nv * i]), hs[i]);
            acc.x += p.x;
            acc.y += p.y;
        }
        double2 &wv = W[(size_t)nv * s + r];
        wv.x -= acc.x;
        wv.y -= acc.y;
        q = wv.x * wv.x + wv.y * wv.y;
    }
    if (!npart) return;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off, 64);
    if (threadIdx.x % 64 == 0) sq[threadIdx.x / 64] = q;
    __syncthreads();
    if (threadIdx.x == 0) npart[(size_t)s * gridDim.x + blockIdx.x] = (sq[0] + sq[1]) + (sq[2] + sq[3]);
}

// Single-vector MGS step: W[s] -= h V[s][i] with h = hbuf[s][i].
__global__ void k_axpy_col(int nv, const int *__restrict__ slots, int i, const double2 *__restrict__ V,
                           size_t ldv_slot, const double2 *__restrict__ h, int ldh, double2 *__restrict__ W) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slots[blockIdx.y];
    if (r >= nv) return;
    const double2 hv = h[(size_t)s * ldh + i];
    const double2 p = cmul(V[ldv_slot * s + (size_t)nv * i + r], hv);
    double2 &wv = W[(size_t)nv * s + r];
    wv.x -= p.x;
    wv.y -= p.y;
}

// nrm[s] = ||W[s]||_2 (sum of squares in a fixed order).
__global__ __launch_bounds__(RB) void k_norm(int nv, const int *__restrict__ slots, const double2 *__restrict__ W,
                                             size_t ldw, double *__restrict__ nrm) {
    const int s = slots[blockIdx.x];
    const double2 *w = W + ldw * s;
    double2 acc = make_double2(0, 0);
    for (int r = threadIdx.x; r < nv; r += RB) acc.x += w[r].x * w[r].x + w[r].y * w[r].y;
    acc = block_sum2(acc);
    if (threadIdx.x == 0) nrm[s] = sqrt(acc.x);
}

// V[s][j] = W[s] / nrm[s]   (nothing when nrm == 0, as gmres_module.f90:148-151)
template <typename VT>
__global__ void k_scale_into(int nv, const int *__restrict__ slots, const int *__restrict__ jcol,
                             const double2 *__restrict__ W, size_t ldw, const double *__restrict__ nrm,
                             VT *__restrict__ V, size_t ldv_slot) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slots[blockIdx.y];
    if (r >= nv) return;
    const double d = nrm[s];
    if (d == 0.0) return;
    const double2 w = W[ldw * s + r];
    V[ldv_slot * s + (size_t)nv * jcol[s] + r] = from_d2<VT>(make_double2(w.x / d, w.y / d));
}

// R[s] = B[s] - W[s]
__global__ void k_residual(int nv, const int *__restrict__ slots, const double2 *__restrict__ Bv,
                           const double2 *__restrict__ W, double2 *__restrict__ R) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slots[blockIdx.y];
    if (r >= nv) return;
    const size_t o = (size_t)nv * s + r;
    R[o] = make_double2(Bv[o].x - W[o].x, Bv[o].y - W[o].y);
}

// BLAS zrotg (reference algorithm) and zrot on one pair.
__device__ void zrotg_d(double2 &ca, double2 cb, double &c, double2 &s) {
    const double aa = hypot(ca.x, ca.y);
    if (aa == 0.0) {
        c = 0.0;
        s = make_double2(1.0, 0.0);
        ca = cb;
        return;
    }
    const double ab = hypot(cb.x, cb.y);
    const double scale = aa + ab;
    const double ra = aa / scale, rb = ab / scale;
    const double norm = scale * sqrt(ra * ra + rb * rb);
    const double2 alpha = make_double2(ca.x / aa, ca.y / aa);
    c = aa / norm;
    const double2 cbc = make_double2(cb.x, -cb.y);
    const double2 t = cmul(alpha, cbc);
    s = make_double2(t.x / norm, t.y / norm);
    ca = make_double2(alpha.x * norm, alpha.y * norm);
}
__device__ void zrot_d(double2 &x, double2 &y, double c, double2 s) {
    const double2 sy = cmul(s, y);
    const double2 sc = make_double2(s.x, -s.y);
    const double2 scx = cmul(sc, x);
    const double2 t = make_double2(c * x.x + sy.x, c * x.y + sy.y);
    y = make_double2(c * y.x - scx.x, c * y.y - scx.y);
    x = t;
}

// Givens update of column j of H[s] (gmres_module.f90:153-167), one
// workgroup per slot: the column and the previous rotations are staged in LDS
// by all lanes, lane 0 applies the rotations in order, all lanes write back.
// hcol holds the Gram-Schmidt coefficients h_0..h_j, hn the new norm.
__global__ __launch_bounds__(64) void k_givens(const int *__restrict__ slots, const int *__restrict__ jcol,
                                               const double2 *__restrict__ hcol, int ldh,
                                               const double *__restrict__ hn, double2 *__restrict__ H, int ldH,
                                               size_t ldH_slot, double2 *__restrict__ g, int ldg,
                                               double *__restrict__ cs, double2 *__restrict__ sn, int ldcs,
                                               double *__restrict__ resid) {
    extern __shared__ double2 sh_g[];  // col[ldH] | sn[ldcs] | cs[ldcs] (doubles)
    double2 *col = sh_g;
    double2 *snl = sh_g + ldH;
    double *csl = reinterpret_cast<double *>(sh_g + ldH + ldcs);
    const int s = slots[blockIdx.x];
    const int j = jcol[s];
    const int t = threadIdx.x;
    double2 *Hc = H + ldH_slot * s + (size_t)ldH * j;
    const double2 *hs = hcol + (size_t)s * ldh;
    double *css = cs + (size_t)s * ldcs;
    double2 *sns = sn + (size_t)s * ldcs;
    for (int i = t; i <= j; i += 64) col[i] = hs[i];
    for (int i = t; i < j; i += 64) {
        csl[i] = css[i];
        snl[i] = sns[i];
    }
    if (t == 0) col[j + 1] = make_double2(hn[s], 0.0);
    __syncthreads();
    if (t == 0) {
        for (int i = 0; i < j; ++i) zrot_d(col[i], col[i + 1], csl[i], snl[i]);
        double2 a = col[j];
        double c;
        double2 sv;
        zrotg_d(a, col[j + 1], c, sv);
        css[j] = c;
        sns[j] = sv;
        double2 *gs = g + (size_t)s * ldg;
        zrot_d(gs[j], gs[j + 1], c, sv);
        zrot_d(col[j], col[j + 1], c, sv);
        col[j + 1] = make_double2(0, 0);
        resid[s] = hypot(gs[j + 1].x, gs[j + 1].y);
    }
    __syncthreads();
    for (int i = t; i <= j + 1; i += 64) Hc[i] = col[i];
}

// Back substitution H(0:j,0:j) y = g(0:j) for each slot (one workgroup per slot).
__global__ __launch_bounds__(64) void k_trsv(const int *__restrict__ slots, const int *__restrict__ jlen,
                                             const double2 *__restrict__ H, int ldH, size_t ldH_slot,
                                             const double2 *__restrict__ g, int ldg, double2 *__restrict__ y,
                                             int ldy) {
    const int s = slots[blockIdx.x];
    const int n = jlen[s];
    const double2 *Hs = H + ldH_slot * s;
    double2 *ys = y + (size_t)s * ldy;
    const double2 *gs = g + (size_t)s * ldg;
    for (int i = threadIdx.x; i < n; i += 64) ys[i] = gs[i];
    __syncthreads();
    for (int i = n - 1; i >= 0; --i) {
        if (threadIdx.x == 0) {
            const double2 hii = Hs[i + (size_t)ldH * i];
            const double den = hii.x * hii.x + hii.y * hii.y;
            const double2 v = ys[i];
            ys[i] = make_double2((v.x * hii.x + v.y * hii.y) / den, (v.y * hii.x - v.x * hii.y) / den);
        }
        __syncthreads();
        const double2 yi = ys[i];
        for (int r = threadIdx.x; r < i; r += 64) {
            const double2 p = cmul(Hs[r + (size_t)ldH * i], yi);
            ys[r].x -= p.x;
            ys[r].y -= p.y;
        }
        __syncthreads();
    }
}

// X[s] (+)= V[s][0:jlen] y[s]
template <typename VT>
__global__ void k_update_x(int nv, const int *__restrict__ slots, const int *__restrict__ jlen,
                           const VT *__restrict__ V, size_t ldv_slot, const double2 *__restrict__ y, int ldy,
                           const double2 *__restrict__ dinv, double2 *__restrict__ X, int accumulate) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slots[blockIdx.y];
    if (r >= nv) return;
    const int n = jlen[s];
    const VT *v = V + ldv_slot * s + r;
    const double2 *ys = y + (size_t)s * ldy;
    double2 acc = make_double2(0, 0);
    for (int i = 0; i < n; ++i) {
        const double2 p = cmul(to_d2(v[(size_t)nv * i]), ys[i]);
        acc.x += p.x;
        acc.y += p.y;
    }
    if (dinv) acc = cmul(dinv[(size_t)nv * s + r], acc);  // x = D^-1 u (right preconditioning)
    double2 &xv = X[(size_t)nv * s + r];
    if (accumulate) {
        xv.x += acc.x;
        xv.y += acc.y;
    } else {
        xv = acc;
    }
}

// Operator input of each slot: stage[s] = V[s][j_s] (times D^-1 when preconditioned).
template <typename VT>
__global__ void k_stage(int nv, const int *__restrict__ slots, const int *__restrict__ jcol,
                        const VT *__restrict__ V, size_t ldv_slot, const double2 *__restrict__ dinv,
                        double2 *__restrict__ stage) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slots[blockIdx.y];
    if (r >= nv) return;
    double2 v = to_d2(V[ldv_slot * s + (size_t)nv * jcol[s] + r]);
    if (dinv) v = cmul(dinv[(size_t)nv * s + r], v);
    stage[(size_t)nv * s + r] = v;
}

// Start a Krylov cycle: V[s][0] = R[s] / beta[s], g[s] = (beta, 0, ...).
template <typename VT>
__global__ void k_start_cycle(int nv, const int *__restrict__ slots, const double2 *__restrict__ R,
                              const double *__restrict__ beta, VT *__restrict__ V, size_t ldv_slot,
                              double2 *__restrict__ g, int ldg) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slots[blockIdx.y];
    const double b = beta[s];
    if (r < ldg) g[(size_t)s * ldg + r] = make_double2(r == 0 ? b : 0.0, 0.0);
    if (r >= nv) return;
    const double2 v = R[(size_t)nv * s + r];
    V[ldv_slot * s + r] = from_d2<VT>(make_double2(v.x / b, v.y / b));
}

// Copy a right-hand side (column k of rhs, nv x 3) into slot s of B and R, zero X.
__global__ void k_load_rhs(int nv, int s, const double2 *__restrict__ rhs, int k, double2 *__restrict__ Bv,
                           double2 *__restrict__ R, double2 *__restrict__ X) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nv) return;
    const double2 v = rhs[(size_t)nv * k + r];
    const size_t o = (size_t)nv * s + r;
    Bv[o] = v;
    R[o] = v;
    X[o] = make_double2(0, 0);
}

// ----------------------------------------------------------------- driver
struct System {
    int ifreq, pol;
};

struct HostRing {
    char *base = nullptr;
    size_t chunk = 0;
    int count = 0, next = 0;
    HostRing() = default;
    HostRing(const HostRing &) = delete;
    HostRing &operator=(const HostRing &) = delete;
    ~HostRing() {
        if (base) (void)hipHostFree(base);
    }
    // grow-only: a ring that already holds n chunks of this size is reused
    // (the previous solve has synchronised its stream before returning)
    void init(size_t chunk_bytes, int n) {
        const size_t want = (chunk_bytes + 255) / 256 * 256;
        next = 0;
        if (base && want <= chunk && n <= count) return;
        if (base) (void)hipHostFree(base);
        base = nullptr;
        chunk = want;
        count = n;
        PXH_HIP(hipHostMalloc(reinterpret_cast<void **>(&base), chunk * (size_t)n, hipHostMallocDefault));
    }
    void *take() {
        void *p = base + chunk * (size_t)next;
        next = (next + 1) % count;
        return p;
    }
};

struct Batched {
    pxh_ctx *c;
    int nv, m, B;
    bool mixed, plain;
    DevBuf<double2> X, Bv, R, W, H, g, hbuf, hacc, y, sn, dpart;
    DevBuf<double> wpart, npart, wn2;  // norm partials of the Gram-Schmidt passes
    DevBuf<int> jc2_d;                 // columns of the second pass per slot (0 = skipped)
    DevBuf<double> cs, nrm, resid;
    DevBuf<int> slots_d, jcount_d, jp1_d, jl_d;
    DevBuf<double2> stage;
    DevBuf<SysParam> par;
    std::vector<int> h_list, h_jp1, h_jl, h_jc;
    std::vector<SysParam> h_par;
    std::vector<double2> h_lam;
    DevBuf<double2> lam;
    DevBuf<double2> dinv;  // right preconditioner 1/diag(A(omega_s)) per slot
    bool precond = false;
    std::vector<int> jcount;
    size_t ldv_slot = 0, ldH_slot = 0;
    int ldH = 0, ldg = 0, ldh = 0;
    HostRing ring;                 // pinned staging of the uploads
    double *pin_res = nullptr;     // pinned readback of per-slot reals (2 B)
    size_t pin_n = 0;
    ~Batched() {
        if (pin_res) (void)hipHostFree(pin_res);
    }
    // per-call state back to its defaults; the buffers stay (grow-only)
    void reset(pxh_ctx *ctx) {
        c = ctx;
        nv = m = B = 0;
        mixed = plain = precond = false;
        ldv_slot = ldH_slot = 0;
        ldH = ldg = ldh = 0;
    }
    void pin(size_t n) {
        if (pin_res && n <= pin_n) return;
        if (pin_res) (void)hipHostFree(pin_res);
        pin_res = nullptr;
        PXH_HIP(hipHostMalloc(reinterpret_cast<void **>(&pin_res), sizeof(double) * n, hipHostMallocDefault));
        pin_n = n;
    }
};

// Work buffers of the iterative solvers, kept with the context between calls
// (grow-only), so that a solve allocates only its Krylov basis and what
// scales with the number of systems: at a few ms per spectrum (BEM
// multi-shift) a round of hipMalloc / hipHostMalloc per call is a visible
// share of the time.
constexpr size_t kKeepSolveBytes = (size_t)256 << 20;

// A captured multi-shift Arnoldi step and the launch parameters it was
// captured with (sizes and device / pinned addresses): replayed only while
// they are all the same.
struct MsGraph {
    std::vector<uintptr_t> key;
    hipGraphExec_t exec = nullptr;
};

struct SolverWs {
    Batched b;
    DevBuf<double2> V, Wb, Hs, gq, snq, colq, sig;  // multi-shift
    DevBuf<double> csq, msres, mspart, mstres;
    std::vector<MsGraph> ms_graphs;  // one per number of active fields
    // per-call buffers of solve_spectrum kept between calls when together
    // under kKeepSolveBytes (each hipMalloc / hipFree pair of a 100 MB
    // solution block costs a C3-sized spectrum ~0.3 ms)
    DevBuf<double2> rhs_keep, sol_keep, part_keep;
    DevBuf<double4> bpos_keep;
    DevBuf<double> barea_keep;
    ~SolverWs() {
        for (MsGraph &g : ms_graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
    }
};

SolverWs &solver_ws(pxh_ctx *c) {
    if (!c->solver_ws) c->solver_ws = new SolverWs();
    return *static_cast<SolverWs *>(c->solver_ws);
}

// Small host-to-device uploads go through a ring of pinned chunks, so that
// the copies are asynchronous and need no synchronisation of their own: the
// loop synchronises once per step (to read the residual estimates), and a
// step uploads far fewer chunks than the ring holds.
template <typename T>
void upload(Batched &S, T *d_dst, const T *src, size_t n) {
    const size_t bytes = sizeof(T) * n;
    PXH_CHECK(bytes <= S.ring.chunk, PXH_EINVAL, "upload larger than a staging chunk");
    void *p = S.ring.take();
    std::memcpy(p, src, bytes);
    PXH_HIP(hipMemcpyAsync(d_dst, p, bytes, hipMemcpyHostToDevice, S.c->stream));
}

void upload_list(Batched &S, const std::vector<int> &list) { upload(S, S.slots_d.p, list.data(), list.size()); }

// hbuf[s][i] += hacc[s][i] for i <= j_s
__global__ void k_add_h(const int *__restrict__ slots, const int *__restrict__ jcount, double2 *__restrict__ h,
                        const double2 *__restrict__ h2, int ldh) {
    const int s = slots[blockIdx.x];
    for (int i = threadIdx.x; i < jcount[s]; i += blockDim.x) {
        h[(size_t)s * ldh + i].x += h2[(size_t)s * ldh + i].x;
        h[(size_t)s * ldh + i].y += h2[(size_t)s * ldh + i].y;
    }
}

// ------------------------------------------- delayed re-orthogonalisation
// Classical Gram-Schmidt with the second pass of each vector delayed to the
// next step (DCGS2: Swirydowicz, Langou, Ananthan, Yang, Thomas, Numer. Linear
// Algebra Appl. 2020; Bielich et al., Parallel Comput. 2022), so that a step
// reads the basis twice instead of four times.  At step j the basis holds
// the final q_0..q_{j-1} and the candidate u_j (one pass, normalised).  One
// pass of dots gives a = Q^H u_j, beta = u_j^H u_j, b = Q^H w, gamma = u_j^H w
// for w = A u_j; then
//   nu = sqrt(beta - |a|^2),  q_j = (u_j - Q a) / nu,
//   column j-1 of H becomes final: h_{i,j-1} += h_{j,j-1} a_i, h_{j,j-1} *= nu,
//   f = (b, (gamma - a^H b) / nu),  e = H a (final columns 0..j-1),
//   column j (pending): c = (f - e) / nu,
//   w <- (w - Q f_{0..j-1} - q_j f_j) / nu,  h_{j+1,j} = ||w||, u_{j+1} = w / ||w||,
// with A q_j = (w - A Q a) / nu and A Q a = Q_{j+1} e.  One pass of updates
// forms q_j and the new w together.

// Dots of basis columns 0..jc-1 of each slot against W and against column
// jc-1 itself (the candidate): partial sums per chunk of kDotRows rows, as
// k_dots_part, into pw and pu.
constexpr int kDot2Group = 4;  // basis vectors per pass of the two-vector dots (register budget)
// Dots of basis columns 0..jc-1 of each slot against W and against column
// jc-1 itself (the candidate): partial sums per chunk of kDotRows rows, as
// k_dots_part, into pw and pu; kDot2Group columns per workgroup layer.
template <typename VT>
__global__ __launch_bounds__(256) void k_dots2_part(int nv, const int *__restrict__ slots,
                                                    const int *__restrict__ jcount, const VT *__restrict__ V,
                                                    size_t ldv_slot, const double2 *__restrict__ W,
                                                    double2 *__restrict__ pw, double2 *__restrict__ pu, int ldh,
                                                    int nch) {
    constexpr int U = kDotRows / 64;
    const int ch = blockIdx.x * 4 + threadIdx.x / 64;
    const int s = slots[blockIdx.y];
    const int lane = threadIdx.x % 64;
    if (ch >= nch) return;
    const int jc = jcount[s];
    const int i0 = blockIdx.z * kDot2Group;
    if (i0 >= jc) return;
    const int r0 = ch * kDotRows + lane;
    double2 w[U], u[U];
    const double2 *ws = W + (size_t)nv * s;
    const VT *us = V + ldv_slot * s + (size_t)nv * (jc - 1);
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const bool in = r0 + 64 * k < nv;
        w[k] = in ? ws[r0 + 64 * k] : make_double2(0, 0);
        u[k] = in ? to_d2(us[r0 + 64 * k]) : make_double2(0, 0);
    }
    double2 aw[kDot2Group], au[kDot2Group];
#pragma unroll
    for (int g = 0; g < kDot2Group; ++g) {
        aw[g] = au[g] = make_double2(0, 0);
        if (i0 + g < jc) {
            const VT *v = V + ldv_slot * s + (size_t)nv * (i0 + g);
#pragma unroll
            for (int k = 0; k < U; ++k) {
                if (r0 + 64 * k < nv) {
                    const double2 vv = to_d2(v[r0 + 64 * k]);
                    const double2 p = cconjmul(vv, w[k]), q = cconjmul(vv, u[k]);
                    aw[g].x += p.x;
                    aw[g].y += p.y;
                    au[g].x += q.x;
                    au[g].y += q.y;
                }
            }
        }
    }
#pragma unroll
    for (int g = 0; g < kDot2Group; ++g) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            aw[g].x += __shfl_xor(aw[g].x, off, 64);
            aw[g].y += __shfl_xor(aw[g].y, off, 64);
            au[g].x += __shfl_xor(au[g].x, off, 64);
            au[g].y += __shfl_xor(au[g].y, off, 64);
        }
    }
    if (lane < kDot2Group && i0 + lane < jc) {
        double2 vw = aw[0], vu = au[0];
#pragma unroll
        for (int g = 1; g < kDot2Group; ++g)
            if (lane == g) {
                vw = aw[g];
                vu = au[g];
            }
        pw[((size_t)s * ldh + i0 + lane) * nch + ch] = vw;
        pu[((size_t)s * ldh + i0 + lane) * nch + ch] = vu;
    }
}

// Per slot (one wavefront), step j = jcount[s]: from b (hb, gamma at j) and a
// (ha, beta at j) -> nu, the finished column j-1 (into Hu, and gcol / ghn for
// the Givens step), the pending column j (pend) and the update coefficients
// (acoef: a, fcoef: f).  nu_out = 0 flags a candidate inside span(Q)
// (breakdown: the Krylov space is invariant).
__global__ __launch_bounds__(64) void k_dcgs_coef(const int *__restrict__ slots, const int *__restrict__ jcount,
                                                  const double2 *__restrict__ hb, const double2 *__restrict__ ha,
                                                  int ldh, double2 *__restrict__ pend, double2 *__restrict__ Hu,
                                                  int ldH, size_t ldH_slot, double2 *__restrict__ gcol,
                                                  double *__restrict__ ghn, double2 *__restrict__ acoef,
                                                  double2 *__restrict__ fcoef, double *__restrict__ nu_out) {
    const int s = slots[blockIdx.x];
    const int j = jcount[s];
    const int t = threadIdx.x;
    const double2 *bs = hb + (size_t)s * ldh;
    const double2 *as = ha + (size_t)s * ldh;
    double2 *ps = pend + (size_t)s * ldh;
    double2 *Hs = Hu + ldH_slot * s;
    // |a|^2 and a^H b over i < j
    double aa = 0.0, abr = 0.0, abi = 0.0;
    for (int i = t; i < j; i += 64) {
        const double2 ai = as[i], bi = bs[i];
        aa += ai.x * ai.x + ai.y * ai.y;
        abr += ai.x * bi.x + ai.y * bi.y;
        abi += ai.x * bi.y - ai.y * bi.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
        aa += __shfl_xor(aa, off, 64);
        abr += __shfl_xor(abr, off, 64);
        abi += __shfl_xor(abi, off, 64);
    }
    const double beta = as[j].x;
    const double nu2 = beta - aa;
    const bool bd = !(nu2 > 1e-24 * beta);
    const double nu = bd ? 1.0 : sqrt(nu2);
    if (t == 0) nu_out[s] = bd ? 0.0 : nu;
    // column j-1 becomes final
    if (j >= 1) {
        const double eta = ps[j].x;
        for (int i = t; i < j; i += 64) {
            const double2 v = make_double2(ps[i].x + eta * as[i].x, ps[i].y + eta * as[i].y);
            Hs[(size_t)ldH * (j - 1) + i] = v;
            gcol[(size_t)s * ldh + i] = v;
        }
        if (t == 0) {
            const double sub = bd ? 0.0 : eta * nu;
            Hs[(size_t)ldH * (j - 1) + j] = make_double2(sub, 0.0);
            ghn[s] = sub;
        }
    }
    __syncthreads();
    // e = H a over the final columns 0..j-1 (column k has rows 0..k+1);
    // f; pending column j: c = (f - e) / nu
    const double2 gam = bs[j];
    const double2 fj = make_double2((gam.x - abr) / nu, (gam.y - abi) / nu);
    for (int i = t; i <= j; i += 64) {
        double2 e = make_double2(0, 0);
        for (int k = (i > 0 ? i - 1 : 0); k < j; ++k) {
            const double2 h = Hs[(size_t)ldH * k + i];
            const double2 p = cmul(h, as[k]);
            e.x += p.x;
            e.y += p.y;
        }
        const double2 f = i < j ? bs[i] : fj;
        ps[i] = make_double2((f.x - e.x) / nu, (f.y - e.y) / nu);
        fcoef[(size_t)s * ldh + i] = f;
        if (i < j) acoef[(size_t)s * ldh + i] = as[i];
    }
}

// q_j = (u_j - Q a) / nu into the basis, W <- (W - Q f_{0..j-1} - q_j f_j) / nu,
// with per-workgroup partial sums of ||W||^2.
template <typename VT>
__global__ __launch_bounds__(256) void k_dcgs_update(int nv, const int *__restrict__ slots,
                                                     const int *__restrict__ jcount, VT *__restrict__ V,
                                                     size_t ldv_slot, const double2 *__restrict__ acoef,
                                                     const double2 *__restrict__ fcoef, int ldh,
                                                     const double *__restrict__ nu_in, double2 *__restrict__ W,
                                                     double *__restrict__ npart) {
    __shared__ double sq[4];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slots[blockIdx.y];
    const int j = jcount[s];
    const double nu = nu_in[s] > 0.0 ? nu_in[s] : 1.0;
    double q2 = 0.0;
    if (r < nv) {
        VT *vs = V + ldv_slot * s + r;
        const double2 *as = acoef + (size_t)s * ldh;
        const double2 *fs = fcoef + (size_t)s * ldh;
        double2 qa = make_double2(0, 0), qf = make_double2(0, 0);
        for (int i = 0; i < j; ++i) {
            const double2 v = to_d2(vs[(size_t)nv * i]);
            const double2 p = cmul(v, as[i]), q = cmul(v, fs[i]);
            qa.x += p.x;
            qa.y += p.y;
            qf.x += q.x;
            qf.y += q.y;
        }
        const double2 u = to_d2(vs[(size_t)nv * j]);
        const double2 qj = make_double2((u.x - qa.x) / nu, (u.y - qa.y) / nu);
        vs[(size_t)nv * j] = from_d2<VT>(qj);
        const double2 fj = fs[j];
        const double2 qfj = cmul(qj, fj);
        double2 &wv = W[(size_t)nv * s + r];
        wv = make_double2((wv.x - qf.x - qfj.x) / nu, (wv.y - qf.y - qfj.y) / nu);
        q2 = wv.x * wv.x + wv.y * wv.y;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) q2 += __shfl_xor(q2, off, 64);
    if (threadIdx.x % 64 == 0) sq[threadIdx.x / 64] = q2;
    __syncthreads();
    if (threadIdx.x == 0) npart[(size_t)s * gridDim.x + blockIdx.x] = (sq[0] + sq[1]) + (sq[2] + sq[3]);
}

// The same two passes for a complex FP32 basis with nv even: each work-item
// takes row pairs (one 16-byte load per basis vector instead of two 8-byte
// ones).  Measured at the C2 shape (62 732 rows, 48 slots, 2-41 vectors):
// the dots 2.6-2.9 -> 4.2-4.7 TB/s, the update 4.7-5.6 -> 4.9-6.3 TB/s.
constexpr int kPairRowsU = 2;                        // row pairs per work-item in the dots
constexpr int kDotPairRows = 2 * 256 * kPairRowsU;  // rows per workgroup (one chunk)
constexpr int kDotPairGroup = 8;                     // basis vectors per workgroup layer

// 32 values per lane summed over the wavefront by halving exchanges (31
// shuffles instead of 192); on return v[0] of lane L holds the sum of value
// L >> 1.
__device__ __forceinline__ void wave_transpose_sum32(double (&v)[32], int lane) {
    int n = 32;
#pragma unroll
    for (int o = 32; o >= 2; o >>= 1) {
        const int h = n / 2;
        const bool up = (lane & o) != 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (k < h) {
                const double send = up ? v[k] : v[k + h];
                const double keep = up ? v[k + h] : v[k];
                v[k] = keep + __shfl_xor(send, o, 64);
            }
        }
        n = h;
    }
    v[0] += __shfl_xor(v[0], 1, 64);
}

// As k_dots2_part: chunk partials of conj(q_i) . W and conj(q_i) . u_{jc-1}
// for kDotPairGroup basis vectors per workgroup layer, one chunk of
// kDotPairRows rows per workgroup.
__global__ __launch_bounds__(256) void k_dots2_pairs(int nv, const int *__restrict__ slots,
                                                     const int *__restrict__ jcount, const float2 *__restrict__ V,
                                                     size_t ldv_slot, const double2 *__restrict__ W,
                                                     double2 *__restrict__ pw, double2 *__restrict__ pu, int ldh,
                                                     int nch) {
    constexpr int G = kDotPairGroup, U2 = kPairRowsU;
    static_assert(4 * G == 32, "four values per basis vector, 32 per lane");
    __shared__ double red[4][32];
    const int ch = blockIdx.x;
    const int s = slots[blockIdx.y];
    const int jc = jcount[s];
    const int i0 = blockIdx.z * G;
    if (i0 >= jc) return;
    const int t = threadIdx.x, lane = t % 64, wv = t / 64;
    const int np = nv / 2;
    const int p0 = ch * 256 * U2 + t;
    double2 w[2 * U2], u[2 * U2];
    const double2 *ws = W + (size_t)nv * s;
    const float4 *us = reinterpret_cast<const float4 *>(V + ldv_slot * s + (size_t)nv * (jc - 1));
#pragma unroll
    for (int k = 0; k < U2; ++k) {
        const int p = p0 + 256 * k;
        const bool in = p < np;
        const float4 uu = in ? us[p] : make_float4(0, 0, 0, 0);
        u[2 * k] = make_double2(uu.x, uu.y);
        u[2 * k + 1] = make_double2(uu.z, uu.w);
        w[2 * k] = in ? ws[2 * p] : make_double2(0, 0);
        w[2 * k + 1] = in ? ws[2 * p + 1] : make_double2(0, 0);
    }
    double v[32];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        double2 aw = make_double2(0, 0), au = make_double2(0, 0);
        if (i0 + g < jc) {
            const float4 *vp = reinterpret_cast<const float4 *>(V + ldv_slot * s + (size_t)nv * (i0 + g));
#pragma unroll
            for (int k = 0; k < U2; ++k) {
                const int p = p0 + 256 * k;
                if (p < np) {
                    const float4 vv = vp[p];
                    const double2 a = make_double2(vv.x, vv.y), b = make_double2(vv.z, vv.w);
                    const double2 p1 = cconjmul(a, w[2 * k]), p2 = cconjmul(b, w[2 * k + 1]);
                    const double2 q1 = cconjmul(a, u[2 * k]), q2 = cconjmul(b, u[2 * k + 1]);
                    aw.x += p1.x + p2.x;
                    aw.y += p1.y + p2.y;
                    au.x += q1.x + q2.x;
                    au.y += q1.y + q2.y;
                }
            }
        }
        v[4 * g] = aw.x;
        v[4 * g + 1] = aw.y;
        v[4 * g + 2] = au.x;
        v[4 * g + 3] = au.y;
    }
    wave_transpose_sum32(v, lane);
    if ((lane & 1) == 0) red[wv][lane >> 1] = v[0];
    __syncthreads();
    if (t < 32) {
        const double x = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
        const int g = t / 4, c = t % 4;
        if (i0 + g < jc) {
            double *dst = reinterpret_cast<double *>((c < 2 ? pw : pu) + ((size_t)s * ldh + i0 + g) * nch + ch);
            dst[c & 1] = x;
        }
    }
}

// As k_dcgs_update, one row pair per work-item.
__global__ __launch_bounds__(256) void k_dcgs_update_pairs(int nv, const int *__restrict__ slots,
                                                           const int *__restrict__ jcount, float2 *__restrict__ V,
                                                           size_t ldv_slot, const double2 *__restrict__ acoef,
                                                           const double2 *__restrict__ fcoef, int ldh,
                                                           const double *__restrict__ nu_in, double2 *__restrict__ W,
                                                           double *__restrict__ npart) {
    __shared__ double sq[4];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slots[blockIdx.y];
    const int j = jcount[s];
    const double nu = nu_in[s] > 0.0 ? nu_in[s] : 1.0;
    const int np = nv / 2;
    double q2 = 0.0;
    if (p < np) {
        float4 *vs = reinterpret_cast<float4 *>(V + ldv_slot * s) + p;
        const size_t ld4 = (size_t)np;
        const double2 *as = acoef + (size_t)s * ldh;
        const double2 *fs = fcoef + (size_t)s * ldh;
        double2 qa0 = make_double2(0, 0), qf0 = qa0, qa1 = qa0, qf1 = qa0;
        for (int i = 0; i < j; ++i) {
            const float4 v4 = vs[ld4 * i];
            const double2 a = as[i], f = fs[i];
            const double2 v0 = make_double2(v4.x, v4.y), v1 = make_double2(v4.z, v4.w);
            const double2 pa0 = cmul(v0, a), pf0 = cmul(v0, f), pa1 = cmul(v1, a), pf1 = cmul(v1, f);
            qa0.x += pa0.x;
            qa0.y += pa0.y;
            qf0.x += pf0.x;
            qf0.y += pf0.y;
            qa1.x += pa1.x;
            qa1.y += pa1.y;
            qf1.x += pf1.x;
            qf1.y += pf1.y;
        }
        const float4 u4 = vs[ld4 * j];
        const double2 qj0 = make_double2((u4.x - qa0.x) / nu, (u4.y - qa0.y) / nu);
        const double2 qj1 = make_double2((u4.z - qa1.x) / nu, (u4.w - qa1.y) / nu);
        vs[ld4 * j] = make_float4((float)qj0.x, (float)qj0.y, (float)qj1.x, (float)qj1.y);
        const double2 fj = fs[j];
        const double2 qfj0 = cmul(qj0, fj), qfj1 = cmul(qj1, fj);
        double2 *wp = W + (size_t)nv * s + 2 * p;
        const double2 w0 = wp[0], w1 = wp[1];
        const double2 n0 = make_double2((w0.x - qf0.x - qfj0.x) / nu, (w0.y - qf0.y - qfj0.y) / nu);
        const double2 n1 = make_double2((w1.x - qf1.x - qfj1.x) / nu, (w1.y - qf1.y - qfj1.y) / nu);
        wp[0] = n0;
        wp[1] = n1;
        q2 = (n0.x * n0.x + n0.y * n0.y) + (n1.x * n1.x + n1.y * n1.y);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) q2 += __shfl_xor(q2, off, 64);
    if (threadIdx.x % 64 == 0) sq[threadIdx.x / 64] = q2;
    __syncthreads();
    if (threadIdx.x == 0) npart[(size_t)s * gridDim.x + blockIdx.x] = (sq[0] + sq[1]) + (sq[2] + sq[3]);
}

// pend[s][jcount[s] + 1] = ||w|| (the pending column's subdiagonal).
__global__ void k_dcgs_sub(const int *__restrict__ slots, const int *__restrict__ jcount,
                           const double *__restrict__ nrm, double2 *__restrict__ pend, int ldh) {
    const int s = slots[blockIdx.x];
    if (threadIdx.x == 0) pend[(size_t)s * ldh + jcount[s] + 1] = make_double2(nrm[s], 0.0);
}

// Closing a cycle after J = jcount[s] final vectors: the pending column J-1
// made final against the last candidate (W = ||w|| u_J still): with
// d = Q^H W (hd), column J-1 += d and its subdiagonal = sqrt(||W||^2 - |d|^2).
__global__ __launch_bounds__(64) void k_dcgs_close(const int *__restrict__ slots, const int *__restrict__ jcount,
                                                   const double2 *__restrict__ hd, const double2 *__restrict__ pend,
                                                   int ldh, double2 *__restrict__ gcol, double *__restrict__ ghn) {
    const int s = slots[blockIdx.x];
    const int J = jcount[s];
    const double2 *ps = pend + (size_t)s * ldh;
    const double2 *ds = hd + (size_t)s * ldh;
    double dd = 0.0;
    for (int i = threadIdx.x; i < J; i += 64) {
        const double2 d = ds[i];
        dd += d.x * d.x + d.y * d.y;
        gcol[(size_t)s * ldh + i] = make_double2(ps[i].x + d.x, ps[i].y + d.y);
    }
    for (int off = 32; off > 0; off >>= 1) dd += __shfl_xor(dd, off, 64);
    if (threadIdx.x == 0) {
        const double eta = ps[J].x;
        ghn[s] = sqrt(fmax(eta * eta - dd, 0.0));
    }
}

// ------------------------------------------------- projected starting iterate
// For a homogeneous atomistic target A(omega) = M + S(omega) with S diagonal
// (q rows -z_q(t_i), dipole rows Lambda_i(omega), the frequency-dependent
// shifts of the operator); for alloys M keeps a mild frequency dependence
// (the w_j/w_i ratios) and is taken at the neighbour's frequency.  Every converged system leaves x and M x (= A x -
// S x) behind; a newly admitted system at omega starts from the x0 in the
// span of up to kProjMax converged neighbours of the same field that
// minimises ||b - A(omega) x0||, with A(omega) x_g = M x_g + S(omega) x_g
// formed without a product.
constexpr int kProjMax = 16;
constexpr int kProjL = kProjMax * (kProjMax + 1) / 2 + kProjMax;  // upper triangle of G^H G, then G^H b
constexpr int kProjChunk = 64;                                     // rows staged in LDS at a time
constexpr int kProjRows = 2048;                                    // rows per workgroup

__device__ __forceinline__ double2 row_shift(int r, int n, int nt, const double4 *__restrict__ pos,
                                             const double *__restrict__ nbfrac, const SysParam &p) {
    if (r < n) {
        const int t = (int)pos[r].w;
        return make_double2(-p.zq[t].x, -p.zq[t].y);
    }
    const int i = (r - n) / 3;
    double2 l = make_double2(0, 0);
    for (int t = 0; t < nt; ++t) {
        const double f = nbfrac[t + (size_t)nt * i];
        l.x += f * p.inv_alpha[t].x;
        l.y += f * p.inv_alpha[t].y;
    }
    return l;
}

// mx[loc] = W[s] - S(omega_s) X[s] for a converged slot (W = A X, FP64).
__global__ void k_store_mx(int nv, int n, int nt, const double4 *__restrict__ pos, const double *__restrict__ nbfrac,
                           const SysParam *__restrict__ par, int s, const double2 *__restrict__ W,
                           const double2 *__restrict__ X, double2 *__restrict__ mx) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nv) return;
    const double2 sh = row_shift(r, n, nt, pos, nbfrac, par[s]);
    const double2 x = X[(size_t)nv * s + r];
    const double2 w = W[(size_t)nv * s + r];
    mx[r] = make_double2(w.x - (sh.x * x.x - sh.y * x.y), w.y - (sh.x * x.y + sh.y * x.x));
}

// Per admitted slot (info: k, then k neighbour offsets into sol / mx):
// partial sums over kProjRows rows of G^H G (upper triangle) and G^H b,
// G_i = mx_i + S(omega_s) x_i.  Rows are staged kProjChunk at a time in LDS
// as columns G_0..G_{k-1}, b; work-item e < k(k+1)/2 + k owns one entry.
// part[(a * nblk + blk) * kProjL + e].
__global__ __launch_bounds__(256) void k_proj_gram(int nv, int n, int nt, const double4 *__restrict__ pos,
                                                   const double *__restrict__ nbfrac,
                                                   const SysParam *__restrict__ par, const int *__restrict__ adm,
                                                   const long long *__restrict__ info,
                                                   const double2 *__restrict__ sol, const double2 *__restrict__ mx,
                                                   const double2 *__restrict__ Bv, double2 *__restrict__ part) {
    __shared__ double2 tile[kProjMax + 1][kProjChunk];
    __shared__ double2 shift[kProjChunk];
    __shared__ long long off[kProjMax];
    const int a = blockIdx.y;
    const int s = adm[a];
    const long long *in = info + (size_t)a * (kProjMax + 1);
    const int k = (int)in[0];
    if (threadIdx.x < kProjMax) off[threadIdx.x] = in[1 + threadIdx.x];
    // entry of this work-item: (i, j) with i <= j < k, or G_i^H b
    const int t = threadIdx.x;
    const int ntri = k * (k + 1) / 2;
    int ei = -1, ej = -1;
    if (t < ntri) {
        int i = 0, rem = t;
        while (rem >= k - i) {
            rem -= k - i;
            ++i;
        }
        ei = i;
        ej = i + rem;
    } else if (t < ntri + k) {
        ei = t - ntri;
        ej = k;  // the right-hand side column
    }
    double2 acc = make_double2(0, 0);
    const int r0 = blockIdx.x * kProjRows;
    const SysParam &p = par[s];
    for (int c0 = r0; c0 < min(nv, r0 + kProjRows); c0 += kProjChunk) {
        __syncthreads();
        if (t < kProjChunk) {
            const int r = c0 + t;
            shift[t] = r < nv ? row_shift(r, n, nt, pos, nbfrac, p) : make_double2(0, 0);
        }
        __syncthreads();
        for (int e = t; e < (k + 1) * kProjChunk; e += 256) {
            const int col = e / kProjChunk, rr = e % kProjChunk;
            const int r = c0 + rr;
            double2 v = make_double2(0, 0);
            if (r < nv) {
                if (col < k) {
                    const double2 x = sol[off[col] + r];
                    const double2 m = mx[off[col] + r];
                    const double2 sh = shift[rr];
                    v = make_double2(m.x + sh.x * x.x - sh.y * x.y, m.y + sh.x * x.y + sh.y * x.x);
                } else {
                    v = Bv[(size_t)nv * s + r];
                }
            }
            tile[col][rr] = v;
        }
        __syncthreads();
        if (ei >= 0) {
            for (int rr = 0; rr < kProjChunk; ++rr) {
                const double2 u = cconjmul(tile[ei][rr], tile[ej][rr]);
                acc.x += u.x;
                acc.y += u.y;
            }
        }
    }
    // entries are numbered for the slot's k; k_proj_solve reads them so
    if (t < kProjL) part[((size_t)a * gridDim.x + blockIdx.x) * kProjL + t] = acc;
}

// Coefficients of the projected start for each admitted slot: the partials
// summed in a fixed order, the normal equations scaled to a unit diagonal and
// solved by Cholesky; a column whose pivot falls below 1e-12 of its diagonal
// (dependent on the others) gets coefficient 0.  One wavefront per slot; the
// small matrix lives in LDS.
__global__ __launch_bounds__(64) void k_proj_solve(int nblk, const long long *__restrict__ info,
                                                   const double2 *__restrict__ part, double2 *__restrict__ coef) {
    __shared__ double2 acc[kProjL];
    __shared__ double2 G[kProjMax][kProjMax];
    __shared__ double2 c[kProjMax];
    __shared__ double d[kProjMax];
    __shared__ int keep[kProjMax];
    const int a = blockIdx.x;
    const int k = (int)info[(size_t)a * (kProjMax + 1)];
    const int ntri = k * (k + 1) / 2;
    for (int q = threadIdx.x; q < ntri + k; q += 64) {
        double2 v = make_double2(0, 0);
        for (int b = 0; b < nblk; ++b) {
            const double2 p = part[((size_t)a * nblk + b) * kProjL + q];
            v.x += p.x;
            v.y += p.y;
        }
        acc[q] = v;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    int e = 0;
    for (int i = 0; i < k; ++i)
        for (int j = i; j < k; ++j, ++e) {
            G[i][j] = acc[e];
            G[j][i] = make_double2(acc[e].x, -acc[e].y);
        }
    for (int i = 0; i < k; ++i, ++e) c[i] = acc[e];  // G^H b, then y, then the coefficients
    for (int i = 0; i < k; ++i) d[i] = G[i][i].x > 0.0 ? 1.0 / sqrt(G[i][i].x) : 0.0;
    for (int i = 0; i < k; ++i) {
        for (int j = 0; j < k; ++j) G[i][j] = make_double2(G[i][j].x * d[i] * d[j], G[i][j].y * d[i] * d[j]);
        c[i] = make_double2(c[i].x * d[i], c[i].y * d[i]);
    }
    // G = L L^H in place (lower triangle); a dropped column has L_jj = 0
    for (int j = 0; j < k; ++j) {
        double djj = G[j][j].x;
        for (int q = 0; q < j; ++q)
            if (keep[q]) djj -= G[j][q].x * G[j][q].x + G[j][q].y * G[j][q].y;
        keep[j] = d[j] > 0.0 && djj > 1e-12;
        const double ljj = keep[j] ? sqrt(djj) : 0.0;
        G[j][j] = make_double2(ljj, 0.0);
        for (int i = j + 1; i < k; ++i) {
            double2 v = G[i][j];
            for (int q = 0; q < j; ++q) {
                if (!keep[q]) continue;
                const double2 l1 = G[i][q], l2 = G[j][q];  // v -= L_iq conj(L_jq)
                v.x -= l1.x * l2.x + l1.y * l2.y;
                v.y -= l1.y * l2.x - l1.x * l2.y;
            }
            G[i][j] = keep[j] ? make_double2(v.x / ljj, v.y / ljj) : make_double2(0, 0);
        }
    }
    // L y = G^H b, L^H c = y
    for (int i = 0; i < k; ++i) {
        double2 v = c[i];
        for (int q = 0; q < i; ++q) {
            if (!keep[q]) continue;
            const double2 p = cmul(G[i][q], c[q]);
            v.x -= p.x;
            v.y -= p.y;
        }
        c[i] = keep[i] ? make_double2(v.x / G[i][i].x, v.y / G[i][i].x) : make_double2(0, 0);
    }
    for (int i = k - 1; i >= 0; --i) {
        double2 v = c[i];
        for (int q = i + 1; q < k; ++q) {
            if (!keep[i] || !keep[q]) continue;
            const double2 p = cconjmul(G[q][i], c[q]);
            v.x -= p.x;
            v.y -= p.y;
        }
        c[i] = keep[i] ? make_double2(v.x / G[i][i].x, v.y / G[i][i].x) : make_double2(0, 0);
    }
    for (int i = 0; i < kProjMax; ++i)
        coef[(size_t)a * kProjMax + i] = i < k ? make_double2(c[i].x * d[i], c[i].y * d[i]) : make_double2(0, 0);
}

// X[s] = sum_i coef_i x_i for each admitted slot.
__global__ void k_proj_x0(int nv, int n, int nt, const double4 *__restrict__ pos, const double *__restrict__ nbfrac,
                          const SysParam *__restrict__ par, const int *__restrict__ adm,
                          const long long *__restrict__ info, const double2 *__restrict__ coef,
                          const double2 *__restrict__ sol, const double2 *__restrict__ mx,
                          const double2 *__restrict__ Bv, int with_res, double2 *__restrict__ X,
                          double2 *__restrict__ R) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (r >= nv) return;
    const int s = adm[a];
    const long long *in = info + (size_t)a * (kProjMax + 1);
    const int k = (int)in[0];
    double2 v = make_double2(0, 0), m = make_double2(0, 0);
    for (int i = 0; i < k; ++i) {
        const double2 ci = coef[(size_t)a * kProjMax + i];
        const double2 p = cmul(ci, sol[in[1 + i] + r]);
        v.x += p.x;
        v.y += p.y;
        if (with_res) {
            const double2 q = cmul(ci, mx[in[1 + i] + r]);
            m.x += q.x;
            m.y += q.y;
        }
    }
    X[(size_t)nv * s + r] = v;
    if (with_res) {
        // A(omega) x0 = sum_i c_i (M x_i + S(omega) x_i) = M x0 + S(omega) x0
        const double2 sh = row_shift(r, n, nt, pos, nbfrac, par[s]);
        const double2 sx = cmul(sh, v);
        const double2 b = Bv[(size_t)nv * s + r];
        R[(size_t)nv * s + r] = make_double2(b.x - (m.x + sx.x), b.y - (m.y + sx.y));
    }
}

// ------------------------------------------------------------ multi-shift
// Shifted systems (A0 + sigma_s I) x_s = b, x_s(0) = 0, share one Krylov
// space K_m(A0, b): a single Arnoldi process per right-hand side serves every
// shift.  The Hessenberg matrix of shift s is H + sigma_s [I; 0]; each shift
// keeps its own Givens rotations and rotated right-hand side g, whose
// |g_{j+1}| is that shift's GMRES residual norm.  The state of (shift s,
// right-hand side f) sits at q = f nsh + s, ldg entries each.

// Column j of the Arnoldi matrix of each listed right-hand side:
// Hs[f][j][0..j] = h, Hs[f][j][j+1] = ||w||.
__global__ void k_ms_store_col(const int *__restrict__ slots, const int *__restrict__ jcol,
                               const double2 *__restrict__ hbuf, int ldh, const double *__restrict__ nrm,
                               double2 *__restrict__ Hs, int ldH, size_t ldH_f) {
    const int f = slots[blockIdx.x];
    const int j = jcol[f];
    double2 *Hc = Hs + ldH_f * f + (size_t)ldH * j;
    for (int i = threadIdx.x; i <= j; i += blockDim.x) Hc[i] = hbuf[(size_t)f * ldh + i];
    if (threadIdx.x == 0) Hc[j + 1] = make_double2(nrm[f], 0.0);
}

// End of an Arnoldi step: jcount[s] += 1 and jp1[s] = jcount[s] + 1 for the
// listed slots.  The step's counters live on the device, so that one step
// captured as a HIP graph can be replayed for the next ones.
__global__ void k_ms_advance(int na, const int *__restrict__ slots, int *__restrict__ jcount,
                             int *__restrict__ jp1) {
    for (int k = threadIdx.x; k < na; k += blockDim.x) {
        const int s = slots[k];
        const int j = jcount[s] + 1;
        jcount[s] = j;
        jp1[s] = j + 1;
    }
}

// g[q] = (beta_f, 0, ...) for every shift of the listed right-hand sides.
__global__ void k_ms_init(int nsh, const int *__restrict__ slots, const double *__restrict__ beta,
                          double2 *__restrict__ g, int ldg) {
    const int f = slots[blockIdx.y];
    double2 *gq = g + ((size_t)f * nsh + blockIdx.x) * ldg;
    for (int i = threadIdx.x; i < ldg; i += blockDim.x) gq[i] = make_double2(i == 0 ? beta[f] : 0.0, 0.0);
}

// Step j for every shift (grid nsh x listed right-hand sides): the stored
// rotations 0..j-1 applied to column j of H + sigma_s I, then the rotation
// that zeroes its subdiagonal, applied to g as well; resid[q] = |g_{j+1}|.
// Lane 0 runs the rotation chain over chunks of kMsChunk entries that all
// lanes stage in LDS; the rotated column itself is not kept.
constexpr int kMsChunk = 512;
__global__ __launch_bounds__(64) void k_ms_givens(int nsh, const int *__restrict__ slots,
                                                  const int *__restrict__ jcol, const double2 *__restrict__ Hs,
                                                  int ldH, size_t ldH_f, const double2 *__restrict__ sigma,
                                                  double *__restrict__ cs, double2 *__restrict__ sn,
                                                  double2 *__restrict__ g, int ldg, double *__restrict__ resid) {
    __shared__ double2 col[kMsChunk + 1], snl[kMsChunk];
    __shared__ double csl[kMsChunk];
    const int s = blockIdx.x;
    const int f = slots[blockIdx.y];
    const int j = jcol[f];
    const size_t q = (size_t)f * nsh + s;
    const double2 *Hc = Hs + ldH_f * f + (size_t)ldH * j;
    const double2 sg = sigma[s];
    double2 carry = make_double2(0, 0);
    // chunk [b0, b0 + kMsChunk) of the rotations; col holds entries b0 .. b0 + kMsChunk
    for (int b0 = 0; b0 <= j; b0 += kMsChunk) {
        const int nr = min(kMsChunk, j - b0);  // rotations in this chunk
        __syncthreads();
        for (int i = threadIdx.x; i <= nr; i += 64) {
            double2 h = Hc[b0 + i];
            if (b0 + i == j) {
                h.x += sg.x;
                h.y += sg.y;
            }
            col[i] = h;
        }
        for (int i = threadIdx.x; i < nr; i += 64) {
            csl[i] = cs[q * ldg + b0 + i];
            snl[i] = sn[q * ldg + b0 + i];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (b0 == 0) carry = col[0];
            for (int i = 0; i < nr; ++i) {
                double2 nx = col[i + 1];
                zrot_d(carry, nx, csl[i], snl[i]);
                carry = nx;
            }
        }
    }
    if (threadIdx.x != 0) return;
    double c;
    double2 sv;
    zrotg_d(carry, Hc[j + 1], c, sv);
    cs[q * ldg + j] = c;
    sn[q * ldg + j] = sv;
    double2 *gq = g + q * ldg;
    zrot_d(gq[j], gq[j + 1], c, sv);
    resid[q] = hypot(gq[j + 1].x, gq[j + 1].y);
}

// g[q](0:J) <- R^-1 g[q](0:J) for every shift, R the rotated (H + sigma_s
// I)(0:J, 0:J).  Column-oriented back substitution from the last column; each
// column of R is rebuilt from H and the stored rotations when it is needed
// (rotation chain in lane 0, staged by chunks, into the per-shift scratch
// colq), so no shift stores its triangle.
__global__ __launch_bounds__(64) void k_ms_solve(int nsh, const int *__restrict__ slots,
                                                 const int *__restrict__ jlen, const double2 *__restrict__ Hs,
                                                 int ldH, size_t ldH_f, const double2 *__restrict__ sigma,
                                                 const double *__restrict__ cs, const double2 *__restrict__ sn,
                                                 double2 *__restrict__ g, int ldg, double2 *__restrict__ colq) {
    __shared__ double2 col[kMsChunk + 1], snl[kMsChunk];
    __shared__ double csl[kMsChunk];
    __shared__ double2 yj_s;
    const int s = blockIdx.x;
    const int f = slots[blockIdx.y];
    const int J = jlen[f];
    const size_t q = (size_t)f * nsh + s;
    double2 *gq = g + q * ldg;
    double2 *cq = colq + q * ldg;
    const double2 sg = sigma[s];
    for (int jj = J - 1; jj >= 0; --jj) {
        const double2 *Hc = Hs + ldH_f * f + (size_t)ldH * jj;
        double2 carry = make_double2(0, 0);
        // rotations 0..jj on column jj (entries 0..jj+1)
        for (int b0 = 0; b0 <= jj; b0 += kMsChunk) {
            const int nr = min(kMsChunk, jj + 1 - b0);
            __syncthreads();
            for (int i = threadIdx.x; i <= nr; i += 64) {
                double2 h = Hc[b0 + i];
                if (b0 + i == jj) {
                    h.x += sg.x;
                    h.y += sg.y;
                }
                col[i] = h;
            }
            for (int i = threadIdx.x; i < nr; i += 64) {
                csl[i] = cs[q * ldg + b0 + i];
                snl[i] = sn[q * ldg + b0 + i];
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                if (b0 == 0) carry = col[0];
                for (int i = 0; i < nr; ++i) {
                    double2 nx = col[i + 1];
                    zrot_d(carry, nx, csl[i], snl[i]);
                    cq[b0 + i] = carry;
                    carry = nx;
                }
            }
        }
        if (threadIdx.x == 0) {
            const double2 r = cq[jj];
            const double den = r.x * r.x + r.y * r.y;
            const double2 v = gq[jj];
            const double2 yv = make_double2((v.x * r.x + v.y * r.y) / den, (v.y * r.x - v.x * r.y) / den);
            gq[jj] = yv;
            yj_s = yv;
        }
        __syncthreads();
        const double2 yj = yj_s;
        for (int i = threadIdx.x; i < jj; i += 64) {
            const double2 p = cmul(cq[i], yj);
            gq[i].x -= p.x;
            gq[i].y -= p.y;
        }
    }
}

// Column jcol[s] of a per-slot basis from the slot vectors: B[s][jcol[s]] = W[s].
__global__ void k_save_col(int nv, const int *__restrict__ slots, const int *__restrict__ jcol,
                           const double2 *__restrict__ W, double2 *__restrict__ Bs, size_t ldb_slot) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slots[blockIdx.y];
    if (r >= nv) return;
    Bs[ldb_slot * s + (size_t)nv * jcol[s] + r] = W[(size_t)nv * s + r];
}

// Solutions and FP64 true residuals of every shift of one right-hand side f:
// x_s = V_f(:, 0:J) y_s, written to sol + stride s + nv pol, and
// b - A(omega_s) x_s = b - (W_f y_s + sigma_s x_s) with W_f(:, j) = A0 v_j,
// the FP64 products of the Arnoldi process (A0 x_s = sum_j y_sj A0 v_j).
// part[(blk * nsh + s)] = sum over the workgroup's rows of |r|^2.
constexpr int kMsShifts = 8;
__global__ __launch_bounds__(256) void k_ms_finish(int nv, int nsh, int f, int pol, const int *__restrict__ jlen,
                                                   const double2 *__restrict__ V, const double2 *__restrict__ Wb,
                                                   size_t ldv_slot, const double2 *__restrict__ y, int ldy,
                                                   const double2 *__restrict__ sigma, const double2 *__restrict__ b,
                                                   double2 *__restrict__ sol, size_t stride,
                                                   double *__restrict__ part) {
    __shared__ double red[4][kMsShifts];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s0 = blockIdx.y * kMsShifts;
    const int J = jlen[f];
    const int ns = min(kMsShifts, nsh - s0);
    const double2 *yq = y + ((size_t)f * nsh + s0) * ldy;
    double2 ax[kMsShifts], aw[kMsShifts];
#pragma unroll
    for (int k = 0; k < kMsShifts; ++k) ax[k] = aw[k] = make_double2(0, 0);
    if (r < nv) {
        const double2 *v = V + ldv_slot * f + r;
        const double2 *w = Wb + ldv_slot * f + r;
        for (int i = 0; i < J; ++i) {
            const double2 vi = v[(size_t)nv * i], wi = w[(size_t)nv * i];
#pragma unroll
            for (int k = 0; k < kMsShifts; ++k) {
                if (k >= ns) break;
                const double2 yk = yq[(size_t)k * ldy + i];
                const double2 p = cmul(vi, yk), q = cmul(wi, yk);
                ax[k].x += p.x;
                ax[k].y += p.y;
                aw[k].x += q.x;
                aw[k].y += q.y;
            }
        }
    }
    const double2 bv = r < nv ? b[r] : make_double2(0, 0);
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
#pragma unroll
    for (int k = 0; k < kMsShifts; ++k) {
        double q2 = 0.0;
        if (k < ns && r < nv) {
            sol[stride * (s0 + k) + (size_t)nv * pol + r] = ax[k];
            const double2 sg = sigma[s0 + k];
            const double2 sx = cmul(sg, ax[k]);
            const double rx = bv.x - (aw[k].x + sx.x), ry = bv.y - (aw[k].y + sx.y);
            q2 = rx * rx + ry * ry;
        }
        for (int off = 32; off > 0; off >>= 1) q2 += __shfl_xor(q2, off, 64);
        if (lane == 0) red[wv][k] = q2;
    }
    __syncthreads();
    if ((int)threadIdx.x < ns)
        part[(size_t)blockIdx.x * nsh + s0 + threadIdx.x] =
            (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// tres[q] = sqrt(sum over workgroups of part), fixed order; one thread per shift.
__global__ void k_ms_tres(int nsh, int nblk, int f, const double *__restrict__ part, double *__restrict__ tres) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsh) return;
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) acc += part[(size_t)b * nsh + s];
    tres[(size_t)f * nsh + s] = sqrt(acc);
}

}  // namespace

// Operator application for slots in `list` (device copy in S.slots_d).
static void apply_op(Batched &S, int nsl, bool fp64, const double2 *x, double2 *y) {
    pxh_ctx *c = S.c;
    if (c->target == PXH_BEM) {
        apply_bem(c, nsl, S.slots_d.p, S.lam.p, x, y, S.nv, !fp64);
    } else if (c->use_stored) {
        apply_atoms_stored(c, nsl, S.slots_d.p, S.par.p, x, y, S.nv);
    } else if (fp64) {
        apply_atoms_fp64(c, nsl, S.slots_d.p, S.par.p, x, y, S.nv);
    } else {
        apply_atoms_fp32(c, nsl, S.slots_d.p, S.par.p, x, y, S.nv);
    }
}

// lambda(omega) = 2 pi (eps_s + eps) / (eps_s - eps)  (bem_module.f90:898-906)
double2 bem_lambda(const pxh_ctx *c, int i_freq) {
    const double *e = c->eps.data() + 2 * (size_t)i_freq;
    const double es = c->mesh.eps_solvent;
    const double nr = es + e[0], ni = e[1], dr = es - e[0], di = -e[1];
    const double den = dr * dr + di * di;
    return make_double2(2.0 * M_PI * (nr * dr + ni * di) / den, 2.0 * M_PI * (ni * dr - nr * di) / den);
}

// Solve all (ifreq, pol) systems.  rhs_all: device, [n_freq][nv x 3].
// VT: element type of the Krylov basis.
// warm: the systems start from the solutions already in d_sol (else x = 0).
template <typename VT>
static int run_batched_t(pxh_ctx *c, int nv, const std::vector<System> &systems, const double2 *d_rhs_all,
                         int rhs_fstride, int rhs_stride, const double *freq, double tol, int kdim, int precision, bool plain,
                         bool warm, double2 *d_sol /* [n_freq][nv x 3] */, std::vector<int> &iters_out,
                         std::vector<double> &tres_out, std::vector<double> &t_admit, std::vector<double> &t_done) {
    Batched &S = solver_ws(c).b;
    S.reset(c);
    DevBuf<VT> V;
    const auto t_start = std::chrono::steady_clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count(); };
    S.nv = nv;
    S.mixed = precision == PXH_MIXED;
    S.plain = plain;
    S.B = plain ? 1 : std::max(1, std::min<int>(c->opt.batch, (int)systems.size()));
    S.m = std::max(1, std::min(kdim, nv));
    const int m = S.m, B = S.B;
    hipStream_t st = c->stream;
    S.ldv_slot = (size_t)nv * (m + 1);
    S.ldH = m + 2;
    S.ldH_slot = (size_t)S.ldH * (m + 1);
    S.ldg = m + 2;
    S.ldh = m + 2;
    S.X.alloc((size_t)nv * B);
    S.Bv.alloc((size_t)nv * B);
    S.R.alloc((size_t)nv * B);
    S.W.alloc((size_t)nv * B);
    V.alloc(S.ldv_slot * B);
    const int nch = (nv + kDotRows - 1) / kDotRows;
    S.dpart.alloc((size_t)B * S.ldh * nch);
    S.wpart.alloc((size_t)B * nch);
    S.npart.alloc((size_t)B * ((nv + 255) / 256));
    S.wn2.alloc(B);
    S.jc2_d.alloc(B);
    S.H.alloc(S.ldH_slot * B);
    S.g.alloc((size_t)S.ldg * B);
    S.hbuf.alloc((size_t)S.ldh * B);
    S.hacc.alloc((size_t)S.ldh * B);
    S.y.alloc((size_t)S.ldh * B);
    S.cs.alloc((size_t)S.ldg * B);
    S.sn.alloc((size_t)S.ldg * B);
    S.nrm.alloc(B);
    S.resid.alloc(B);
    S.jcount_d.alloc(B);
    S.jp1_d.alloc(B);
    S.jl_d.alloc(B);
    S.stage.alloc((size_t)nv * B);
    // diagonal right preconditioning (x = D^-1 u) for the batched path; the
    // plain path is the reference's unpreconditioned GMRES
    // (measured: it cuts the omegaFQFmu and BEM DPCM iteration counts, and
    // makes weakly damped omegaFQ spectra stagnate, so omegaFQ runs without)
    S.precond = !plain && c->opt.precond &&
                (c->target == PXH_BEM ? c->mesh.variant == PXH_DPCM : c->atoms.model == PXH_WFQFMU);
    if (S.precond) S.dinv.alloc((size_t)nv * B);
    S.h_jp1.assign(B, 0);
    S.h_jl.assign(B, 0);
    S.h_jc.assign(B, 0);
    if (c->target == PXH_BEM) {
        S.lam.alloc(B);
        S.h_lam.assign(B, make_double2(0, 0));
    } else {
        S.par.alloc(B);
        S.h_par.assign(B, SysParam{});
    }
    S.jcount.assign(B, 0);
    S.slots_d.alloc(std::max<size_t>(1, (size_t)B));
    S.ring.init(std::max({sizeof(SysParam), sizeof(int) * (size_t)B, sizeof(long long) * (kProjMax + 1) * (size_t)B}),
                4 * B + 64);
    S.pin(2 * (size_t)B);

    // projected starting iterates (atomistic targets, cold start): the
    // systems are admitted coarse to fine over the frequency grid, so that
    // most start between converged neighbours.  For alloys the q rows also
    // carry the frequency-dependent ratios w_j/w_i; M x is then taken at the
    // neighbour's frequency, so the projection minimises a nearby residual
    // (the start is still checked by its FP64 true residual)
    const bool proj = !plain && !warm && c->opt.x0_projection > 0 && c->target != PXH_BEM;
    const int kproj = std::min(kProjMax, std::max(1, c->opt.x0_projection));
    int n_freq_all = 0;
    for (const System &sy : systems) n_freq_all = std::max(n_freq_all, sy.ifreq + 1);
    std::vector<int> perm(systems.size());
    for (size_t k = 0; k < perm.size(); ++k) perm[k] = (int)k;
    DevBuf<double2> mx, proj_part, proj_coef;
    DevBuf<long long> proj_info;
    DevBuf<int> proj_adm;
    std::vector<char> have;  // [ifreq * 3 + pol]: x and M x stored
    const int proj_nblk = (nv + kProjRows - 1) / kProjRows;
    if (proj) {
        auto level = [](int f) {
            int l = 0;
            while (l < 30 && (f >> l) % 2 == 0) ++l;
            return f == 0 ? 31 : l;
        };
        std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) {
            return level(systems[a].ifreq) > level(systems[b].ifreq);
        });
        mx.alloc((size_t)rhs_stride * n_freq_all);
        proj_part.alloc((size_t)B * proj_nblk * kProjL);
        proj_coef.alloc((size_t)B * kProjMax);
        proj_info.alloc((size_t)B * (kProjMax + 1));
        proj_adm.alloc(B);
        have.assign((size_t)3 * n_freq_all, 0);
    }
    // delayed re-orthogonalisation (DCGS2) on the batched path
    const bool dcgs = !plain && c->opt.dcgs2;
    DevBuf<double2> dc_part_u, dc_a, dc_pend, dc_Hu, dc_acoef, dc_fcoef, dc_gcol;
    DevBuf<double> dc_nu, dc_ghn;
    DevBuf<int> dc_jm1, dc_slots2;
    if (dcgs) {
        dc_part_u.alloc((size_t)B * S.ldh * nch);
        dc_a.alloc((size_t)B * S.ldh);
        dc_pend.alloc((size_t)B * S.ldh);
        dc_Hu.alloc(S.ldH_slot * B);
        dc_acoef.alloc((size_t)B * S.ldh);
        dc_fcoef.alloc((size_t)B * S.ldh);
        dc_gcol.alloc((size_t)B * S.ldh);
        dc_nu.alloc(B);
        dc_ghn.alloc(B);
        dc_jm1.alloc(B);
        dc_slots2.alloc(B);
    }
    std::vector<int> h_jm1(B, 0);
    std::vector<int> adm_proj;                 // slots admitted this step with a projected start
    std::vector<long long> h_info;

    enum Phase { FREE, INNER, NEED_RES };
    std::vector<int> phase(B, FREE), sys_of(B, -1), outer(B, 0), res_wait(B, 0);
    const int kResBatch = std::max(1, c->opt.res_batch), kResWait = std::max(0, c->opt.res_wait);
    std::vector<double> beta(B, 0.0), target(B, 0.0);
    std::vector<int> total_it(B, 0);
    size_t next = 0, done = 0;
    iters_out.assign(systems.size(), 0);
    tres_out.assign(systems.size(), 0.0);
    t_admit.assign(systems.size(), 0.0);
    t_done.assign(systems.size(), 0.0);
    const double inner_rel = S.mixed ? c->opt.inner_tol : 0.0;
    auto grid_rows = [&](int nsl) { return dim3((nv + 255) / 256, nsl); };
    std::vector<double> hbeta(B);
    int status = 0;

    auto admit = [&](int s) {
        const int id = perm[next++];
        const System &sy = systems[id];
        t_admit[id] = now_s();
        sys_of[s] = id;
        if (c->target == PXH_BEM) {
            S.h_lam[s] = bem_lambda(c, sy.ifreq);
            upload(S, S.lam.p + s, &S.h_lam[s], 1);
        } else {
            S.h_par[s] = make_sysparam(c, freq[sy.ifreq], sy.ifreq);
            upload(S, S.par.p + s, &S.h_par[s], 1);
        }
        k_load_rhs<<<(nv + 255) / 256, 256, 0, st>>>(nv, s, d_rhs_all + (size_t)rhs_fstride * sy.ifreq, sy.pol,
                                                    S.Bv.p, S.R.p, S.X.p);
        if (S.precond) {
            if (c->target == PXH_BEM)
                precond_diag_bem(c, S.lam.p + s, S.dinv.p + (size_t)nv * s);
            else
                precond_diag_atoms(c, S.par.p + s, S.dinv.p + (size_t)nv * s);
        }
        if (warm)
            PXH_HIP(hipMemcpyAsync(S.X.p + (size_t)nv * s, d_sol + (size_t)rhs_stride * sy.ifreq + (size_t)nv * sy.pol,
                                   sizeof(double2) * nv, hipMemcpyDeviceToDevice, st));
        outer[s] = 0;
        total_it[s] = 0;
        phase[s] = NEED_RES;  // with x = 0 the residual is b: computed below without a product
        if (proj) {
            // the kproj nearest converged frequencies of this field
            std::vector<std::pair<int, int>> cand;
            for (int g = 0; g < n_freq_all; ++g)
                if (have[(size_t)g * 3 + sy.pol]) cand.push_back({std::abs(g - sy.ifreq), g});
            if (!cand.empty()) {
                std::sort(cand.begin(), cand.end());
                const int k = std::min<int>(kproj, (int)cand.size());
                h_info.push_back(k);
                for (int i = 0; i < kProjMax; ++i)
                    h_info.push_back(i < k ? (long long)rhs_stride * cand[i].second + (long long)nv * sy.pol : 0);
                adm_proj.push_back(s);
            }
        }
    };

    std::vector<int> res_first;  // slots whose residual R is already known (x = 0, or a projected start)
    while (done < systems.size()) {
        // fill free slots
        res_first.clear();
        adm_proj.clear();
        h_info.clear();
        for (int s = 0; s < B && next < systems.size(); ++s)
            if (phase[s] == FREE) {
                admit(s);
                if (!warm && (adm_proj.empty() || adm_proj.back() != s)) res_first.push_back(s);
            }
        if (!adm_proj.empty()) {
            const int na = (int)adm_proj.size();
            {
                void *pi = S.ring.take();
                PXH_CHECK(sizeof(long long) * h_info.size() <= S.ring.chunk, PXH_EINVAL, "projection list too long");
                std::memcpy(pi, h_info.data(), sizeof(long long) * h_info.size());
                PXH_HIP(hipMemcpyAsync(proj_info.p, pi, sizeof(long long) * h_info.size(), hipMemcpyHostToDevice, st));
            }
            upload(S, proj_adm.p, adm_proj.data(), adm_proj.size());
            const Atoms &A = c->atoms;
            k_proj_gram<<<dim3(proj_nblk, na), 256, 0, st>>>(nv, A.n, A.nt, A.pos.p, A.d_nbfrac.p, S.par.p,
                                                                   proj_adm.p, proj_info.p, d_sol, mx.p, S.Bv.p,
                                                                   proj_part.p);
            k_proj_solve<<<na, 64, 0, st>>>(proj_nblk, proj_info.p, proj_part.p, proj_coef.p);
            // homogeneous targets: M is frequency independent, so the
            // residual of the start follows from the neighbours' FP64
            // products without one of its own; alloys take a product
            const int with_res = A.hetero ? 0 : 1;
            k_proj_x0<<<grid_rows(na), 256, 0, st>>>(nv, A.n, A.nt, A.pos.p, A.d_nbfrac.p, S.par.p, proj_adm.p,
                                                     proj_info.p, proj_coef.p, d_sol, mx.p, S.Bv.p, with_res,
                                                     S.X.p, S.R.p);
            if (with_res)
                for (int sl : adm_proj) res_first.push_back(sl);
            PXH_HIP(hipGetLastError());
        }
        // ----- residual checks (FP64).  The slots whose x needs a product
        // wait (idle, at most res_wait steps) until res_batch of them can
        // share one FP64 product, or no slot is left in an inner cycle; an
        // idle slot costs the low-precision products nothing
        std::vector<int> need, prod;
        int n_inner = 0;
        bool overdue = false;
        for (int s = 0; s < B; ++s) {
            if (phase[s] == INNER) ++n_inner;
            if (phase[s] != NEED_RES) continue;
            if (std::find(res_first.begin(), res_first.end(), s) != res_first.end()) {
                need.push_back(s);
            } else {
                prod.push_back(s);
                overdue = overdue || res_wait[s] >= kResWait;
            }
        }
        if (!prod.empty() && !plain && (int)prod.size() < kResBatch && n_inner > 0 && !overdue) {
            for (int s : prod) res_wait[s] += 1;
            prod.clear();
        }
        for (int s : prod) {
            res_wait[s] = 0;
            need.push_back(s);
        }
        std::sort(need.begin(), need.end());
        if (!need.empty()) {
            if (!prod.empty()) {
                upload_list(S, prod);
                apply_op(S, (int)prod.size(), true, S.X.p, S.W.p);
                k_residual<<<grid_rows((int)prod.size()), 256, 0, st>>>(nv, S.slots_d.p, S.Bv.p, S.W.p, S.R.p);
                c->stats.v[2] += 1;
                c->stats.v[0] += prod.size();
            }
            upload_list(S, need);
            k_norm<<<(int)need.size(), RB, 0, st>>>(nv, S.slots_d.p, S.R.p, (size_t)nv, S.nrm.p);
            PXH_HIP(hipMemcpyAsync(S.pin_res, S.nrm.p, sizeof(double) * B, hipMemcpyDeviceToHost, st));
            PXH_HIP(hipStreamSynchronize(st));
            std::memcpy(hbeta.data(), S.pin_res, sizeof(double) * B);
            std::vector<int> restart;
            for (int s : need) {
                const bool first = std::find(res_first.begin(), res_first.end(), s) != res_first.end();
                const double rn = hbeta[s];
                const int id = sys_of[s];
                if (rn == 0.0 || (!first && rn < tol)) {
                    // converged (or zero RHS: the solution is zero, gmres_module.f90:84-85)
                    tres_out[id] = rn;
                    iters_out[id] = total_it[s];
                    t_done[id] = now_s();
                    const System &sy = systems[id];
                    PXH_HIP(hipMemcpyAsync(d_sol + (size_t)rhs_stride * sy.ifreq + (size_t)nv * sy.pol,
                                           S.X.p + (size_t)nv * s, sizeof(double2) * nv, hipMemcpyDeviceToDevice,
                                           st));
                    if (proj && !first) {
                        // W holds A(omega) x from this step's FP64 product
                        const Atoms &A = c->atoms;
                        k_store_mx<<<(nv + 255) / 256, 256, 0, st>>>(
                            nv, A.n, A.nt, A.pos.p, A.d_nbfrac.p, S.par.p, s, S.W.p, S.X.p,
                            mx.p + (size_t)rhs_stride * sy.ifreq + (size_t)nv * sy.pol);
                        have[(size_t)sy.ifreq * 3 + sy.pol] = 1;
                    }
                    phase[s] = FREE;
                    ++done;
                    continue;
                }
                if (outer[s] >= c->opt.max_outer) {
                    tres_out[id] = rn;
                    iters_out[id] = total_it[s];
                    t_done[id] = now_s();
                    status = PXH_ENOCONV;
                    phase[s] = FREE;
                    ++done;
                    continue;
                }
                beta[s] = rn;
                target[s] = plain ? tol : std::max(0.5 * tol, inner_rel * rn);
                restart.push_back(s);
            }
            if (!restart.empty()) {
                upload_list(S, restart);
                k_start_cycle<VT><<<dim3((std::max(nv, S.ldg) + 255) / 256, (int)restart.size()), 256, 0, st>>>(
                    nv, S.slots_d.p, S.R.p, S.nrm.p, V.p, S.ldv_slot, S.g.p, S.ldg);
                for (int s : restart) {
                    phase[s] = INNER;
                    S.jcount[s] = 0;
                    outer[s] += 1;
                    c->stats.v[7] += 1;
                }
            }
        }
        // ----- one Arnoldi step for every slot in INNER
        std::vector<int> act;
        for (int s = 0; s < B; ++s)
            if (phase[s] == INNER) act.push_back(s);
        if (act.empty()) continue;
        const int na = (int)act.size();
        int jmax = 0;
        for (int s : act) jmax = std::max(jmax, S.jcount[s]);
        upload_list(S, act);
        // W = A V_j: V_j of each slot is staged at stage + nv * slot (the operator's layout)
        upload(S, S.jcount_d.p, S.jcount.data(), (size_t)B);
        k_stage<VT><<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, S.jcount_d.p, V.p, S.ldv_slot,
                                                  S.precond ? S.dinv.p : nullptr, S.stage.p);
        apply_op(S, na, !S.mixed, S.stage.p, S.W.p);
        if (S.mixed)
            c->stats.v[1] += 1;
        else
            c->stats.v[2] += 1;
        c->stats.v[0] += na;
        // jcount_d holds j; the GS uses j+1 columns
        for (int s : act) S.h_jp1[s] = S.jcount[s] + 1;
        upload(S, S.jp1_d.p, S.h_jp1.data(), (size_t)B);
        int *jp1_dp = S.jp1_d.p;
        if (dcgs) {
            const int nb = (nv + 255) / 256;
            // complex FP32 basis with an even row count: row-pair kernels
            // (fewer chunks, so the partial buffers sized for nch suffice)
            constexpr bool f32_basis = std::is_same<VT, float2>::value;
            const bool pairs = f32_basis && nv % 2 == 0;
            const int nch_u = pairs ? (nv + kDotPairRows - 1) / kDotPairRows : nch;
            const int nb_u = pairs ? (nv / 2 + 255) / 256 : nb;
            const dim3 gs((jmax + 1 + 3) / 4, na);
            // one pass of dots over q_0..q_{j-1}, u_j against w and u_j
            if constexpr (f32_basis) {
                if (pairs)
                    k_dots2_pairs<<<dim3(nch_u, na, (jmax + 1 + kDotPairGroup - 1) / kDotPairGroup), 256, 0, st>>>(
                        nv, S.slots_d.p, jp1_dp, V.p, S.ldv_slot, S.W.p, S.dpart.p, dc_part_u.p, S.ldh, nch_u);
            }
            if (!pairs) {
                const dim3 gd((nch + 3) / 4, na, (jmax + 1 + kDot2Group - 1) / kDot2Group);
                k_dots2_part<VT><<<gd, 256, 0, st>>>(nv, S.slots_d.p, jp1_dp, V.p, S.ldv_slot, S.W.p, S.dpart.p,
                                                     dc_part_u.p, S.ldh, nch);
            }
            k_dots_sum<<<gs, 256, 0, st>>>(S.slots_d.p, jp1_dp, S.dpart.p, S.ldh, nch_u, S.hbuf.p);
            k_dots_sum<<<gs, 256, 0, st>>>(S.slots_d.p, jp1_dp, dc_part_u.p, S.ldh, nch_u, dc_a.p);
            k_dcgs_coef<<<na, 64, 0, st>>>(S.slots_d.p, S.jcount_d.p, S.hbuf.p, dc_a.p, S.ldh, dc_pend.p, dc_Hu.p,
                                           S.ldH, S.ldH_slot, dc_gcol.p, dc_ghn.p, dc_acoef.p, dc_fcoef.p,
                                           dc_nu.p);
            // one pass of updates: q_j into the basis and the new w
            if constexpr (f32_basis) {
                if (pairs)
                    k_dcgs_update_pairs<<<dim3(nb_u, na), 256, 0, st>>>(nv, S.slots_d.p, S.jcount_d.p, V.p,
                                                                        S.ldv_slot, dc_acoef.p, dc_fcoef.p, S.ldh,
                                                                        dc_nu.p, S.W.p, S.npart.p);
            }
            if (!pairs)
                k_dcgs_update<VT><<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, S.jcount_d.p, V.p, S.ldv_slot,
                                                                dc_acoef.p, dc_fcoef.p, S.ldh, dc_nu.p, S.W.p,
                                                                S.npart.p);
            k_sum_slot<<<na, 64, 0, st>>>(S.slots_d.p, S.npart.p, nb_u, nullptr, 1, S.nrm.p);
            k_dcgs_sub<<<na, 64, 0, st>>>(S.slots_d.p, S.jcount_d.p, S.nrm.p, dc_pend.p, S.ldh);
            k_scale_into<VT><<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, jp1_dp, S.W.p, (size_t)nv, S.nrm.p,
                                                           V.p, S.ldv_slot);
            // Givens step on the column that became final (j - 1)
            std::vector<int> gl;
            for (int s : act)
                if (S.jcount[s] >= 1) {
                    gl.push_back(s);
                    h_jm1[s] = S.jcount[s] - 1;
                }
            const size_t lds_g = sizeof(double2) * (S.ldH + S.ldg) + sizeof(double) * S.ldg;
            if (!gl.empty()) {
                upload(S, dc_slots2.p, gl.data(), gl.size());
                upload(S, dc_jm1.p, h_jm1.data(), (size_t)B);
                k_givens<<<(int)gl.size(), 64, lds_g, st>>>(dc_slots2.p, dc_jm1.p, dc_gcol.p, S.ldh, dc_ghn.p, S.H.p,
                                                           S.ldH, S.ldH_slot, S.g.p, S.ldg, S.cs.p, S.sn.p, S.ldg,
                                                           S.resid.p);
            }
            PXH_HIP(hipGetLastError());
            PXH_HIP(hipMemcpyAsync(S.pin_res, S.resid.p, sizeof(double) * B, hipMemcpyDeviceToHost, st));
            PXH_HIP(hipMemcpyAsync(S.pin_res + B, dc_nu.p, sizeof(double) * B, hipMemcpyDeviceToHost, st));
            PXH_HIP(hipStreamSynchronize(st));
            std::vector<int> fin, close;
            for (int s : act) {
                const int j = S.jcount[s];
                const double res = S.pin_res[s], nu = S.pin_res[B + s];
                total_it[s] += 1;
                if (nu == 0.0) {
                    // the candidate lies in span(Q): q_0..q_{j-1} solve the cycle exactly
                    S.h_jl[s] = j;
                    fin.push_back(s);
                    continue;
                }
                S.jcount[s] = j + 1;
                if ((j >= 1 && res < target[s]) || S.jcount[s] >= m) {
                    S.h_jl[s] = S.jcount[s];
                    fin.push_back(s);
                    close.push_back(s);
                }
            }
            if (!close.empty()) {
                // the pending column of each closing slot made final against
                // its last candidate, then its Givens step
                const int nc = (int)close.size();
                int jcl = 0;
                for (int s : close) {
                    jcl = std::max(jcl, S.jcount[s]);
                    h_jm1[s] = S.jcount[s] - 1;
                }
                upload_list(S, close);
                upload(S, S.jcount_d.p, S.jcount.data(), (size_t)B);
                upload(S, dc_jm1.p, h_jm1.data(), (size_t)B);
                const dim3 gd2((nch + 3) / 4, nc, (jcl + kDotGroup - 1) / kDotGroup);
                const dim3 gs2((jcl + 3) / 4, nc);
                k_dots_part<VT><<<gd2, 256, 0, st>>>(nv, S.slots_d.p, S.jcount_d.p, V.p, S.ldv_slot, S.W.p,
                                                     S.dpart.p, S.ldh, nch, nullptr);
                k_dots_sum<<<gs2, 256, 0, st>>>(S.slots_d.p, S.jcount_d.p, S.dpart.p, S.ldh, nch, S.hacc.p);
                k_dcgs_close<<<nc, 64, 0, st>>>(S.slots_d.p, S.jcount_d.p, S.hacc.p, dc_pend.p, S.ldh, dc_gcol.p,
                                                dc_ghn.p);
                k_givens<<<nc, 64, lds_g, st>>>(S.slots_d.p, dc_jm1.p, dc_gcol.p, S.ldh, dc_ghn.p, S.H.p, S.ldH,
                                               S.ldH_slot, S.g.p, S.ldg, S.cs.p, S.sn.p, S.ldg, S.resid.p);
            }
            if (!fin.empty()) {
                upload_list(S, fin);
                upload(S, S.jl_d.p, S.h_jl.data(), (size_t)B);
                k_trsv<<<(int)fin.size(), 64, 0, st>>>(S.slots_d.p, S.jl_d.p, S.H.p, S.ldH, S.ldH_slot, S.g.p,
                                                       S.ldg, S.y.p, S.ldh);
                k_update_x<VT><<<grid_rows((int)fin.size()), 256, 0, st>>>(
                    nv, S.slots_d.p, S.jl_d.p, V.p, S.ldv_slot, S.y.p, S.ldh, S.precond ? S.dinv.p : nullptr,
                    S.X.p, 1);
                PXH_HIP(hipGetLastError());
                for (int s : fin) phase[s] = NEED_RES;
            }
            continue;
        }
        const bool mgs = plain && (jmax + 1 < 25);
        if constexpr (std::is_same<VT, double2>::value) {
            if (mgs) {
                // modified Gram-Schmidt (gmres_module.f90:123-127); plain path has one slot
                for (int i = 0; i <= jmax; ++i) {
                    k_dots<<<dim3(1, na), RB, 0, st>>>(nv, S.slots_d.p, jp1_dp, i, V.p, S.ldv_slot, S.W.p,
                                                       S.hbuf.p, S.ldh);
                    k_axpy_col<<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, i, V.p, S.ldv_slot, S.hbuf.p,
                                                             S.ldh, S.W.p);
                }
            }
        }
        if (!mgs) {
            // classical Gram-Schmidt (:128-142); the second pass runs for every
            // slot on the plain path (the reference's CGS2) and where the first
            // pass lost more than 1 - 1/sqrt(2) of the norm on the batched path
            const dim3 gd((nch + 3) / 4, na, (jmax + 1 + kDotGroup - 1) / kDotGroup);
            const dim3 gs((jmax + 1 + 3) / 4, na);
            const int nb = (nv + 255) / 256;
            k_dots_part<VT><<<gd, 256, 0, st>>>(nv, S.slots_d.p, jp1_dp, V.p, S.ldv_slot, S.W.p, S.dpart.p,
                                                S.ldh, nch, S.wpart.p);
            k_dots_sum<<<gs, 256, 0, st>>>(S.slots_d.p, jp1_dp, S.dpart.p, S.ldh, nch, S.hbuf.p);
            k_sum_slot<<<na, 64, 0, st>>>(S.slots_d.p, S.wpart.p, nch, nullptr, 0, S.wn2.p);
            k_gs_update<VT><<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, jp1_dp, V.p, S.ldv_slot, S.hbuf.p,
                                                          S.ldh, S.W.p, S.npart.p);
            k_reorth_decide<<<na, 64, 0, st>>>(S.slots_d.p, jp1_dp, S.npart.p, nb, S.wn2.p,
                                               plain ? 2.0 : c->opt.reorth_kappa, S.jc2_d.p, S.nrm.p);
            k_dots_part<VT><<<gd, 256, 0, st>>>(nv, S.slots_d.p, S.jc2_d.p, V.p, S.ldv_slot, S.W.p, S.dpart.p,
                                                S.ldh, nch, nullptr);
            k_dots_sum<<<gs, 256, 0, st>>>(S.slots_d.p, S.jc2_d.p, S.dpart.p, S.ldh, nch, S.hacc.p);
            k_gs_update<VT><<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, S.jc2_d.p, V.p, S.ldv_slot,
                                                          S.hacc.p, S.ldh, S.W.p, S.npart.p);
            k_sum_slot<<<na, 64, 0, st>>>(S.slots_d.p, S.npart.p, nb, S.jc2_d.p, 1, S.nrm.p);
            k_add_h<<<na, 64, 0, st>>>(S.slots_d.p, S.jc2_d.p, S.hbuf.p, S.hacc.p, S.ldh);
        } else {
            k_norm<<<na, RB, 0, st>>>(nv, S.slots_d.p, S.W.p, (size_t)nv, S.nrm.p);
        }
        k_scale_into<VT><<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, jp1_dp, S.W.p, (size_t)nv, S.nrm.p, V.p,
                                                       S.ldv_slot);
        k_givens<<<na, 64, sizeof(double2) * (S.ldH + S.ldg) + sizeof(double) * S.ldg, st>>>(
            S.slots_d.p, S.jcount_d.p, S.hbuf.p, S.ldh, S.nrm.p, S.H.p,
                                                S.ldH, S.ldH_slot, S.g.p, S.ldg, S.cs.p, S.sn.p, S.ldg, S.resid.p);
        PXH_HIP(hipGetLastError());
        std::vector<double> hres(B);
        PXH_HIP(hipMemcpyAsync(S.pin_res, S.resid.p, sizeof(double) * B, hipMemcpyDeviceToHost, st));
        PXH_HIP(hipStreamSynchronize(st));
        std::memcpy(hres.data(), S.pin_res, sizeof(double) * B);
        std::vector<int> fin;
        for (int s : act) {
            S.jcount[s] += 1;
            total_it[s] += 1;
            if (hres[s] < target[s] || S.jcount[s] >= m) fin.push_back(s);
        }
        if (!fin.empty()) {
            upload_list(S, fin);
            for (int s : fin) S.h_jl[s] = S.jcount[s];
            upload(S, S.jl_d.p, S.h_jl.data(), (size_t)B);
            int *jl_dp = S.jl_d.p;
            k_trsv<<<(int)fin.size(), 64, 0, st>>>(S.slots_d.p, jl_dp, S.H.p, S.ldH, S.ldH_slot, S.g.p, S.ldg,
                                                   S.y.p, S.ldh);
            k_update_x<VT><<<grid_rows((int)fin.size()), 256, 0, st>>>(nv, S.slots_d.p, jl_dp, V.p, S.ldv_slot,
                                                                  S.y.p, S.ldh, S.precond ? S.dinv.p : nullptr,
                                                                  S.X.p, 1);
            PXH_HIP(hipGetLastError());
            for (int s : fin) phase[s] = NEED_RES;
        }
    }
    PXH_HIP(hipStreamSynchronize(st));
    return status;
}

static int run_batched(pxh_ctx *c, int nv, const std::vector<System> &systems, const double2 *d_rhs_all,
                       int rhs_fstride, int rhs_stride, const double *freq, double tol, int kdim, int precision, bool plain, bool warm,
                       double2 *d_sol, std::vector<int> &iters_out, std::vector<double> &tres_out,
                       std::vector<double> &t_admit, std::vector<double> &t_done) {
    if (precision == PXH_MIXED && !plain)
        return run_batched_t<float2>(c, nv, systems, d_rhs_all, rhs_fstride, rhs_stride, freq, tol, kdim, precision, plain, warm,
                                     d_sol, iters_out, tres_out, t_admit, t_done);
    return run_batched_t<double2>(c, nv, systems, d_rhs_all, rhs_fstride, rhs_stride, freq, tol, kdim, precision, plain, warm,
                                  d_sol, iters_out, tres_out, t_admit, t_done);
}

// Frequency families with a scalar shift: A(omega_f) = A0 + sigma_f I and a
// frequency-independent right-hand side.  BEM DPCM: A0 = D, sigma = lambda
// (bem_module.f90:898-906); homogeneous omegaFQ: A0 = K-bar T, sigma = -z_q
// (wfq_module.f90:234-246).  The omegaFQFmu operator shifts its charge and
// dipole blocks by different amounts and is no such family.
//
// BEM IEF, right-preconditioned by S^-1: the operator -(D + lambda I) S with
// u = S q is -(D + lambda I), A0 = -D and sigma = -lambda (operator form
// bem_op = 1).  The charges q = S^-1 u are not formed: alpha needs only
// z_a^T u with z_a = S^-1 c_a (S symmetric), and the residual of the IEF
// system at q is the residual of the shifted system at u.  Used when the
// caller asks for polarizabilities only (polar_only).
static bool shift_family(const pxh_ctx *c, int n_freq, const double *freq, bool polar_only,
                         std::vector<double2> &sigma) {
    sigma.assign(n_freq, make_double2(0, 0));
    if (c->target == PXH_BEM) {
        const bool ief = c->mesh.variant == PXH_IEFPCM;
        if (ief && !polar_only) return false;
        for (int f = 0; f < n_freq; ++f) {
            const double2 l = bem_lambda(c, f);
            sigma[f] = ief ? make_double2(-l.x, -l.y) : l;
        }
        return true;
    }
    const Atoms &A = c->atoms;
    if (A.model != PXH_WFQ || A.hetero) return false;
    for (int f = 0; f < n_freq; ++f) {
        const SysParam p = make_sysparam(c, freq[f], f);
        for (int t = 1; t < A.nt; ++t)
            if (p.zq[t].x != p.zq[0].x || p.zq[t].y != p.zq[0].y) return false;
        sigma[f] = make_double2(-p.zq[0].x, -p.zq[0].y);
    }
    return true;
}

namespace {
}  // namespace

// Multi-shift GMRES over a shift family (shift_family): for each field
// direction one FP64 Arnoldi process on A0 (CGS2, x0 = 0) serves all
// frequencies, and runs until every frequency's GMRES residual is below tol/2
// or the basis holds krylov_dim vectors.  Every solution is then checked
// against its own FP64 true residual b - A(omega) x; the systems above tol
// (and those not converged in the basis) are listed in `unsolved`, their
// current iterate left in d_sol as a starting point.
static int run_multishift(pxh_ctx *c, int nv, const std::vector<System> &systems, int n_freq,
                          const std::vector<double2> &sigma, const double *freq, const double2 *d_rhs_all,
                          int rhs_fstride, int rhs_stride, double tol, int kdim, double2 *d_sol, std::vector<int> &iters_out,
                          std::vector<double> &tres_out, std::vector<double> &t_done, std::vector<int> &unsolved) {
    PXH_CHECK(rhs_fstride == 0, PXH_EINVAL, "multi-shift GMRES needs one right-hand side for all frequencies");
    hipStream_t st = c->stream;
    const auto t_start = std::chrono::steady_clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count(); };
    const bool bem = c->target == PXH_BEM;
    std::vector<int> pols;
    for (const System &sy : systems)
        if (std::find(pols.begin(), pols.end(), sy.pol) == pols.end()) pols.push_back(sy.pol);
    const int F = (int)pols.size();
    const int nsh = n_freq;
    const int m = std::max(1, std::min(kdim, nv));
    SolverWs &ws = solver_ws(c);
    Batched &S = ws.b;
    S.reset(c);
    S.nv = nv;
    S.m = m;
    S.B = std::max(F, std::max(1, std::min<int>(c->opt.batch, (int)systems.size())));
    const int B = S.B;
    S.ldv_slot = (size_t)nv * (m + 1);
    S.ldH = m + 2;
    S.ldg = m + 2;
    S.ldh = m + 2;
    const int ldH = S.ldH, ldg = S.ldg, ldh = S.ldh;
    const size_t ldH_f = (size_t)ldH * (m + 1);
    const int nch = (nv + kDotRows - 1) / kDotRows;
    const int nb = (nv + 255) / 256;
    S.X.alloc((size_t)nv * B);
    S.Bv.alloc((size_t)nv * B);
    S.R.alloc((size_t)nv * B);
    S.W.alloc((size_t)nv * B);
    S.stage.alloc((size_t)nv * B);
    S.dpart.alloc((size_t)F * ldh * nch);
    S.wpart.alloc((size_t)F * nch);
    S.npart.alloc((size_t)F * nb);
    S.wn2.alloc(F);
    S.jc2_d.alloc(F);
    S.g.alloc((size_t)ldg * F);
    S.hbuf.alloc((size_t)ldh * F);
    S.hacc.alloc((size_t)ldh * F);
    S.nrm.alloc(B);
    S.jcount_d.alloc(B);
    S.jp1_d.alloc(B);
    S.jl_d.alloc(B);
    S.slots_d.alloc(B);
    S.ring.init(std::max(sizeof(SysParam), sizeof(int) * (size_t)B) * (size_t)B, 4 * B + 64);
    if (bem) {
        S.lam.alloc(B);
        S.h_lam.assign(B, make_double2(0, 0));
    } else {
        S.par.alloc(B);
        S.h_par.assign(B, SysParam{});
    }
    DevBuf<double2> &V = ws.V, &Wb = ws.Wb, &Hs = ws.Hs, &gq = ws.gq, &snq = ws.snq, &colq = ws.colq,
                    &sig = ws.sig;
    DevBuf<double> &csq = ws.csq, &msres = ws.msres, &mspart = ws.mspart, &mstres = ws.mstres;
    V.alloc(S.ldv_slot * F);
    Wb.alloc(S.ldv_slot * F);  // A0 v_j, kept for the true residuals
    const int fin_blk = (nv + 255) / 256;
    mspart.alloc((size_t)fin_blk * nsh);
    mstres.alloc((size_t)nsh * F);
    PXH_HIP(hipMemsetAsync(mstres.p, 0, sizeof(double) * nsh * F, st));
    Hs.alloc(ldH_f * F);
    gq.alloc((size_t)nsh * F * ldg);
    snq.alloc((size_t)nsh * F * ldg);
    csq.alloc((size_t)nsh * F * ldg);
    colq.alloc((size_t)nsh * F * ldg);
    msres.alloc((size_t)nsh * F);
    sig.upload(sigma.data(), nsh, st);
    S.pin((size_t)nsh * F + B);
    struct {
        double *p;
    } pin{S.pin_res};
    auto grid_rows = [&](int nsl) { return dim3((nv + 255) / 256, nsl); };

    // system index of (frequency, field slot)
    std::vector<int> sys_at((size_t)nsh * F, -1);
    for (size_t k = 0; k < systems.size(); ++k) {
        const int f = (int)(std::find(pols.begin(), pols.end(), systems[k].pol) - pols.begin());
        sys_at[(size_t)f * nsh + systems[k].ifreq] = (int)k;
    }
    iters_out.assign(systems.size(), 0);
    tres_out.assign(systems.size(), 0.0);
    t_done.assign(systems.size(), 0.0);
    unsolved.clear();

    // base operator A0 on the field slots, start vectors b / ||b||
    std::vector<int> all(F);
    for (int f = 0; f < F; ++f) {
        all[f] = f;
        if (bem) {
            upload(S, S.lam.p + f, &S.h_lam[f], 1);
        } else {
            S.h_par[f] = make_sysparam(c, freq[0], 0);
            for (int t = 0; t < c->atoms.nt; ++t) S.h_par[f].zq[t] = make_double2(0, 0);
            upload(S, S.par.p + f, &S.h_par[f], 1);
        }
        k_load_rhs<<<(nv + 255) / 256, 256, 0, st>>>(nv, f, d_rhs_all, pols[f], S.Bv.p, S.R.p, S.X.p);
    }
    upload_list(S, all);
    k_norm<<<F, RB, 0, st>>>(nv, S.slots_d.p, S.R.p, (size_t)nv, S.nrm.p);
    PXH_HIP(hipMemcpyAsync(pin.p, S.nrm.p, sizeof(double) * F, hipMemcpyDeviceToHost, st));
    PXH_HIP(hipStreamSynchronize(st));
    std::vector<int> act;
    for (int f = 0; f < F; ++f)
        if (pin.p[f] > 0.0) act.push_back(f);  // a zero right-hand side has the solution 0
    std::vector<int> conv_it((size_t)nsh * F, 0);
    S.jcount.assign(B, 0);
    S.h_jp1.assign(B, 0);
    S.h_jl.assign(B, 0);
    if (!act.empty()) {
        upload_list(S, act);
        k_start_cycle<double2><<<dim3((std::max(nv, ldg) + 255) / 256, (int)act.size()), 256, 0, st>>>(
            nv, S.slots_d.p, S.R.p, S.nrm.p, V.p, S.ldv_slot, S.g.p, ldg);
        k_ms_init<<<dim3(nsh, (int)act.size()), 64, 0, st>>>(nsh, S.slots_d.p, S.nrm.p, gq.p, ldg);
        for (size_t k = 0; k < act.size(); ++k) c->stats.v[7] += 1;
    }
    // One Arnoldi step of the active fields: every launch reads the step
    // counters (jcount, jp1) from the device and advances them at the end, and
    // the per-shift residuals and norms go to pinned memory, so the same
    // sequence can be captured once as a HIP graph (BEM operator: kernels and
    // read-backs only, the work buffers already sized by an eager step) and
    // replayed while the active set is unchanged.  jmax bounds the Gram-Schmidt
    // grids: the current maximum when launched directly, m when captured.
    auto step = [&](int na, int jmax) {
        k_stage<double2><<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, S.jcount_d.p, V.p, S.ldv_slot, nullptr,
                                                       S.stage.p);
        apply_op(S, na, true, S.stage.p, S.W.p);
        k_save_col<<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, S.jcount_d.p, S.W.p, Wb.p, S.ldv_slot);
        // classical Gram-Schmidt, two passes
        const dim3 gd((nch + 3) / 4, na, (jmax + 1 + kDotGroup - 1) / kDotGroup);
        const dim3 gs((jmax + 1 + 3) / 4, na);
        k_dots_part<double2><<<gd, 256, 0, st>>>(nv, S.slots_d.p, S.jp1_d.p, V.p, S.ldv_slot, S.W.p, S.dpart.p, ldh,
                                                 nch, S.wpart.p);
        k_dots_sum<<<gs, 256, 0, st>>>(S.slots_d.p, S.jp1_d.p, S.dpart.p, ldh, nch, S.hbuf.p);
        k_sum_slot<<<na, 64, 0, st>>>(S.slots_d.p, S.wpart.p, nch, nullptr, 0, S.wn2.p);
        k_gs_update<double2><<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, S.jp1_d.p, V.p, S.ldv_slot, S.hbuf.p,
                                                           ldh, S.W.p, S.npart.p);
        k_reorth_decide<<<na, 64, 0, st>>>(S.slots_d.p, S.jp1_d.p, S.npart.p, nb, S.wn2.p, 2.0, S.jc2_d.p, S.nrm.p);
        k_dots_part<double2><<<gd, 256, 0, st>>>(nv, S.slots_d.p, S.jc2_d.p, V.p, S.ldv_slot, S.W.p, S.dpart.p, ldh,
                                                 nch, nullptr);
        k_dots_sum<<<gs, 256, 0, st>>>(S.slots_d.p, S.jc2_d.p, S.dpart.p, ldh, nch, S.hacc.p);
        k_gs_update<double2><<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, S.jc2_d.p, V.p, S.ldv_slot, S.hacc.p,
                                                           ldh, S.W.p, S.npart.p);
        k_sum_slot<<<na, 64, 0, st>>>(S.slots_d.p, S.npart.p, nb, S.jc2_d.p, 1, S.nrm.p);
        k_add_h<<<na, 64, 0, st>>>(S.slots_d.p, S.jc2_d.p, S.hbuf.p, S.hacc.p, ldh);
        k_scale_into<double2><<<grid_rows(na), 256, 0, st>>>(nv, S.slots_d.p, S.jp1_d.p, S.W.p, (size_t)nv, S.nrm.p,
                                                            V.p, S.ldv_slot);
        k_ms_store_col<<<na, 64, 0, st>>>(S.slots_d.p, S.jcount_d.p, S.hbuf.p, ldh, S.nrm.p, Hs.p, ldH, ldH_f);
        k_ms_givens<<<dim3(nsh, na), 64, 0, st>>>(nsh, S.slots_d.p, S.jcount_d.p, Hs.p, ldH, ldH_f, sig.p,
                                                           csq.p, snq.p, gq.p, ldg, msres.p);
        k_ms_advance<<<1, 64, 0, st>>>(na, S.slots_d.p, S.jcount_d.p, S.jp1_d.p);
        PXH_HIP(hipGetLastError());
        PXH_HIP(hipMemcpyAsync(pin.p, msres.p, sizeof(double) * nsh * F, hipMemcpyDeviceToHost, st));
        PXH_HIP(hipMemcpyAsync(pin.p + (size_t)nsh * F, S.nrm.p, sizeof(double) * F, hipMemcpyDeviceToHost, st));
    };
    // Everything the captured step takes as launch parameters (the step above
    // and apply_bem): a cached graph is replayed only when all of it matches.
    const Mesh &Msh = c->mesh;
    auto graph_key = [&](int na) {
        std::vector<uintptr_t> k{(uintptr_t)na, (uintptr_t)nv, (uintptr_t)nch, (uintptr_t)nb, (uintptr_t)m,
                                 (uintptr_t)S.ldv_slot, (uintptr_t)ldh, (uintptr_t)ldH, (uintptr_t)ldH_f,
                                 (uintptr_t)nsh, (uintptr_t)ldg, (uintptr_t)F, (uintptr_t)Msh.n,
                                 (uintptr_t)Msh.variant, (uintptr_t)c->bem_op};
        for (const void *p : {(const void *)S.slots_d.p, (const void *)S.jcount_d.p, (const void *)S.jp1_d.p,
                              (const void *)V.p, (const void *)S.stage.p, (const void *)S.lam.p, (const void *)S.W.p,
                              (const void *)Msh.D.p, (const void *)Msh.SA.p, (const void *)c->work_t.p,
                              (const void *)c->work_pack.p, (const void *)c->work_gemm.p, (const void *)Wb.p,
                              (const void *)S.dpart.p, (const void *)S.wpart.p, (const void *)S.hbuf.p,
                              (const void *)S.wn2.p, (const void *)S.npart.p, (const void *)S.jc2_d.p,
                              (const void *)S.nrm.p, (const void *)S.hacc.p, (const void *)Hs.p, (const void *)sig.p,
                              (const void *)csq.p, (const void *)snq.p, (const void *)gq.p, (const void *)msres.p,
                              (const void *)pin.p})
            k.push_back((uintptr_t)p);
        return k;
    };
    hipGraphExec_t gexec = nullptr;  // the captured step for the current number of active fields
    int eager_steps = 0;             // steps launched directly since the active set last changed
    bool list_dirty = true;
    while (!act.empty()) {
        const int na = (int)act.size();
        if (list_dirty) {  // the active set changed: its list and counters to the device
            upload_list(S, act);
            for (int f = 0; f < B; ++f) S.h_jp1[f] = S.jcount[f] + 1;
            upload(S, S.jcount_d.p, S.jcount.data(), (size_t)B);
            upload(S, S.jp1_d.p, S.h_jp1.data(), (size_t)B);
            list_dirty = false;
            eager_steps = 0;
            gexec = nullptr;
            if (bem && c->opt.graphs) {
                if ((int)ws.ms_graphs.size() <= na) ws.ms_graphs.resize(na + 1);
                MsGraph &g = ws.ms_graphs[na];
                if (g.exec && g.key == graph_key(na)) gexec = g.exec;
            }
        }
        if (!gexec && bem && c->opt.graphs && eager_steps >= 1) {
            // capture one step (relaxed mode: nothing in it allocates, the
            // eager step before it has sized every buffer)
            MsGraph &g = ws.ms_graphs[na];
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
            g.key.clear();
            PXH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            c->capturing = true;
            hipGraph_t graph = nullptr;
            try {
                step(na, m);
            } catch (...) {
                c->capturing = false;
                (void)hipStreamEndCapture(st, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            c->capturing = false;
            PXH_HIP(hipStreamEndCapture(st, &graph));
            const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) g.exec = nullptr;
            PXH_CHECK(ie == hipSuccess, PXH_EHIP, "hipGraphInstantiate of the multi-shift step failed");
            g.key = graph_key(na);
            gexec = g.exec;
        }
        if (gexec) {
            PXH_HIP(hipGraphLaunch(gexec, st));
            if (bem) c->stats.v[6] += (double)c->mesh.n * c->mesh.n * na;
        } else {
            int jmax = 0;
            for (int f : act) jmax = std::max(jmax, S.jcount[f]);
            step(na, jmax);
            ++eager_steps;
        }
        c->stats.v[2] += 1;
        c->stats.v[0] += na;
        PXH_HIP(hipStreamSynchronize(st));
        std::vector<int> fin, keep;
        for (int f : act) {
            S.jcount[f] += 1;
            const int j = S.jcount[f];
            bool all_conv = true;
            for (int s = 0; s < nsh; ++s) {
                const size_t q = (size_t)f * nsh + s;
                if (conv_it[q] == 0 && pin.p[q] < 0.5 * tol) conv_it[q] = j;
                all_conv = all_conv && conv_it[q] != 0;
            }
            // a zero norm is an invariant subspace: every shift is solved
            const bool breakdown = pin.p[(size_t)nsh * F + f] == 0.0;
            if (all_conv || breakdown || j >= m)
                fin.push_back(f);
            else
                keep.push_back(f);
        }
        for (int f : fin) {
            S.h_jl[f] = S.jcount[f];
            upload(S, S.jl_d.p, S.h_jl.data(), (size_t)B);
            std::vector<int> one(1, f);
            upload_list(S, one);
            k_ms_solve<<<dim3(nsh, 1), 64, 0, st>>>(nsh, S.slots_d.p, S.jl_d.p, Hs.p, ldH, ldH_f, sig.p, csq.p,
                                                    snq.p, gq.p, ldg, colq.p);
            k_ms_finish<<<dim3(fin_blk, (nsh + kMsShifts - 1) / kMsShifts), 256, 0, st>>>(
                nv, nsh, f, pols[f], S.jl_d.p, V.p, Wb.p, S.ldv_slot, gq.p, ldg, sig.p, S.Bv.p + (size_t)nv * f,
                d_sol, (size_t)rhs_stride, mspart.p);
            k_ms_tres<<<(nsh + 255) / 256, 256, 0, st>>>(nsh, fin_blk, f, mspart.p, mstres.p);
            PXH_HIP(hipGetLastError());
            const double t = now_s();
            for (int s = 0; s < nsh; ++s) {
                const size_t q = (size_t)f * nsh + s;
                const int k = sys_at[q];
                if (k < 0) continue;
                iters_out[k] = conv_it[q] ? conv_it[q] : S.jcount[f];
                t_done[k] = t;
            }
        }
        if (!fin.empty()) list_dirty = true;  // the finishes used the slot list for one field each
        act = keep;
    }
    // FP64 true residuals (k_ms_finish), one readback
    {
        std::vector<double> tr((size_t)nsh * F);
        PXH_HIP(hipMemcpyAsync(tr.data(), mstres.p, sizeof(double) * tr.size(), hipMemcpyDeviceToHost, st));
        PXH_HIP(hipStreamSynchronize(st));
        for (int f = 0; f < F; ++f)
            for (int s = 0; s < nsh; ++s) {
                const int k = sys_at[(size_t)f * nsh + s];
                if (k < 0) continue;
                tres_out[k] = tr[(size_t)f * nsh + s];
                if (!(tres_out[k] < tol)) unsolved.push_back(k);
            }
    }
    return 0;
}


// polar[a + 3 b] = sum_i var_b(i) x_a(i) / E0 (+ dipoles); BEM: area-weighted for DPCM.
__global__ void k_polar_partial(int n, int nv, int fmu, int bem_dpcm, const double4 *__restrict__ pos,
                                const double *__restrict__ area, const double2 *__restrict__ var_all,
                                size_t var_stride, double2 *__restrict__ out_all /* [freq][blocks][9] */) {
    __shared__ double2 sh[9][RB];
    const double2 *var = var_all + var_stride * blockIdx.y;
    double2 *out = out_all + (size_t)gridDim.x * 9 * blockIdx.y;
    const int tid = threadIdx.x;
    double2 acc[9];
    for (int k = 0; k < 9; ++k) acc[k] = make_double2(0, 0);
    for (int i = blockIdx.x * RB + tid; i < n; i += gridDim.x * RB) {
        const double4 p = pos[i];
        const double w = bem_dpcm ? area[i] : 1.0;
        for (int b = 0; b < 3; ++b) {
            const double2 v = var[(size_t)nv * b + i];
            acc[0 + 3 * b].x += w * v.x * p.x;
            acc[0 + 3 * b].y += w * v.y * p.x;
            acc[1 + 3 * b].x += w * v.x * p.y;
            acc[1 + 3 * b].y += w * v.y * p.y;
            acc[2 + 3 * b].x += w * v.x * p.z;
            acc[2 + 3 * b].y += w * v.y * p.z;
            if (fmu)
                for (int a = 0; a < 3; ++a) {
                    const double2 mu = var[(size_t)nv * b + n + 3 * i + a];
                    acc[a + 3 * b].x += mu.x;
                    acc[a + 3 * b].y += mu.y;
                }
        }
    }
    for (int k = 0; k < 9; ++k) sh[k][tid] = acc[k];
    __syncthreads();
    for (int off = RB / 2; off > 0; off >>= 1) {
        if (tid < off)
            for (int k = 0; k < 9; ++k) {
                sh[k][tid].x += sh[k][tid + off].x;
                sh[k][tid].y += sh[k][tid + off].y;
            }
        __syncthreads();
    }
    if (tid == 0)
        for (int k = 0; k < 9; ++k) out[(size_t)blockIdx.x * 9 + k] = sh[k][0];
}

// sum[f][k] = the k_polar_partial block sums of frequency f added in block
// order (on the device, so that only 9 values per frequency are read back)
__global__ void k_polar_sum(int nb, int n_freq, const double2 *__restrict__ part, double2 *__restrict__ sum) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 9 * n_freq) return;
    const int f = t / 9, k = t % 9;
    double re = 0.0, im = 0.0;
    for (int b = 0; b < nb; ++b) {
        const double2 v = part[((size_t)nb * f + b) * 9 + k];
        re += v.x;
        im += v.y;
    }
    sum[t] = make_double2(re, im);
}

}  // namespace pxh

namespace pxh {

void results_row(double freq, const double *pol /* 18: complex 3x3 */, double *r);

// Host: the 20 result columns (wfq_module.f90:561-581, target_module.f90:1989-2054).
void results_row(double freq, const double *pol, double *r) {
    const double light = 137.035227, pi = M_PI;
    auto re = [&](int a) { return pol[2 * a]; };
    auto im = [&](int a) { return pol[2 * a + 1]; };
    r[0] = (re(0) + re(4) + re(8)) / 3.0;
    r[1] = (im(0) + im(4) + im(8)) / 3.0;
    r[2] = re(0);
    r[3] = re(4);
    r[4] = re(8);
    r[5] = im(0);
    r[6] = im(4);
    r[7] = im(8);
    const double m_iso = r[0] * r[0] + r[1] * r[1], m_x = r[2] * r[2] + r[5] * r[5], m_y = r[3] * r[3] + r[6] * r[6],
                 m_z = r[4] * r[4] + r[7] * r[7];
    r[8] = 4.0 * pi * freq * r[1] / light;
    r[9] = 4.0 * pi * freq * r[5] / light;
    r[10] = 4.0 * pi * freq * r[6] / light;
    r[11] = 4.0 * pi * freq * r[7] / light;
    const double c4 = 3.0 * light * light * light * light, f4 = 8.0 * pi * freq * freq * freq * freq;
    r[12] = f4 * m_iso / c4;
    r[13] = f4 * m_x / c4;
    r[14] = f4 * m_y / c4;
    r[15] = f4 * m_z / c4;
    for (int k = 0; k < 4; ++k) r[16 + k] = r[8 + k] + r[12 + k];
}

int solve_dense_inversion(pxh_ctx *c, int n_freq, const double *freq, const double2 *d_rhs_all, int rhs_fstride,
                          int rhs_stride,
                          double2 *d_sol, int pol_mask);

namespace {
__global__ void k_pack_z(int n, const double2 *__restrict__ z, double4 *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_double4(z[i].x, z[(size_t)n + i].x, z[2 * (size_t)n + i].x, 0.0);
}
}  // namespace

// z_a = S^-1 c_a for the three Cartesian components of the tessera centroids
// (FP64 batched GMRES on the operator form S, to 1e-13 relative), packed as
// one double4 per tessera.  S is real symmetric positive definite, so z is
// real.  It depends on the mesh alone and is kept with it.
static void ief_polar_weights(pxh_ctx *c) {
    if (c->mesh.ief_z_valid) return;
    DevBuf<double4> &z4 = c->mesh.ief_z;
    const int n = c->mesh.n;
    hipStream_t st = c->stream;
    std::vector<double2> hc((size_t)3 * n);
    double cmax = 0.0;
    for (int k = 0; k < 3; ++k) {
        double q = 0.0;
        for (int i = 0; i < n; ++i) {
            const double v = c->mesh.h_coord[3 * (size_t)i + k];
            hc[(size_t)n * k + i] = make_double2(v, 0.0);
            q += v * v;
        }
        cmax = std::max(cmax, std::sqrt(q));
    }
    DevBuf<double2> rhs, z;
    rhs.upload(hc.data(), hc.size(), st);
    z.alloc(hc.size());
    PXH_HIP(hipMemsetAsync(z.p, 0, sizeof(double2) * hc.size(), st));
    std::vector<System> sys;
    for (int k = 0; k < 3; ++k) sys.push_back({0, k});
    std::vector<int> it;
    std::vector<double> tr, ta, tf;
    const double f0 = 0.0;
    const int saved = c->bem_op;
    c->bem_op = 2;
    const int stt = run_batched(c, n, sys, rhs.p, 0, 3 * n, &f0, 1e-13 * std::max(cmax, 1e-300), std::min(n, 300),
                                PXH_FP64, false, false, z.p, it, tr, ta, tf);
    c->bem_op = saved;
    PXH_CHECK(stt == PXH_OK, PXH_ENOCONV, "IEF: S z = c did not converge");
    z4.alloc(n);
    k_pack_z<<<(n + 255) / 256, 256, 0, st>>>(n, z.p, z4.p);
    PXH_HIP(hipGetLastError());
    PXH_HIP(hipStreamSynchronize(st));
    c->mesh.ief_z_valid = true;
}

int solve_spectrum(pxh_ctx *c, int n_freq, const double *freq, double e0, int rhs_form, int pol_mask, int method,
                   double tol, int kdim, int precision, double *variables, double *polar, double *results,
                   int32_t *iters, double *true_res) {
    const auto t0 = std::chrono::steady_clock::now();
    for (double &v : c->stats.v) v = 0.0;
    hipStream_t st = c->stream;
    const bool bem = c->target == PXH_BEM;
    const int nv = bem ? c->mesh.n : c->atoms.n_var();
    PXH_CHECK(nv > 0, PXH_EINVAL, "no target set (pxh_set_atoms / pxh_set_mesh)");
    if (bem) PXH_CHECK(c->eps_nf >= n_freq, PXH_EINVAL, "permittivity table shorter than the frequency list");
    if (!bem && c->atoms.model == PXH_WFQFMU)
        PXH_CHECK(c->alpha_nf >= n_freq, PXH_EINVAL, "alpha(omega) table shorter than the frequency list");
    const size_t stride = (size_t)nv * 3;
    // right-hand sides: one per frequency for alloys (H-bar(omega) enters
    // them), else one shared by all frequencies (frequency stride 0)
    const bool rhs_per_freq = !bem && c->atoms.hetero;
    const size_t rhs_fstride = rhs_per_freq ? stride : 0;
    const size_t rhs_n = rhs_per_freq ? stride * n_freq : stride, sol_n = stride * n_freq;
    SolverWs &ws = solver_ws(c);
    const bool keep = (rhs_n + sol_n) * sizeof(double2) <= kKeepSolveBytes;
    DevBuf<double2> rhs_local, sol_local;
    DevBuf<double2> &rhs_all = keep ? ws.rhs_keep : rhs_local;
    DevBuf<double2> &sol = keep ? ws.sol_keep : sol_local;
    rhs_all.alloc(rhs_n);
    sol.alloc(sol_n);
    // starting iterates: the caller's variables with the "initial_guess"
    // option on the iterative methods (else x0 = 0, as the reference)
    const bool x0_given = c->opt.initial_guess && variables && method != PXH_INVERSION;
    if (x0_given)
        PXH_HIP(hipMemcpyAsync(sol.p, variables, sizeof(double2) * stride * n_freq, hipMemcpyHostToDevice, st));
    else
        PXH_HIP(hipMemsetAsync(sol.p, 0, sizeof(double2) * stride * n_freq, st));
    if (bem)
        dynamic_rhs_bem(c, e0, rhs_all.p);
    else
        for (int f = 0; f < (rhs_per_freq ? n_freq : 1); ++f)
            dynamic_rhs_atoms(c, freq[f], e0, rhs_form, rhs_all.p + stride * f);
    std::vector<int> it_sys;
    std::vector<double> tr_sys, t_adm, t_fin;
    bool ief_u = false;      // IEF solved for u = S q (polarizabilities only): alpha uses z_a = S^-1 c_a
    std::vector<System> systems;
    for (int f = 0; f < n_freq; ++f)
        for (int k = 0; k < 3; ++k)
            if (pol_mask & (1 << k)) systems.push_back({f, k});
    int status = 0;
    if (method == PXH_INVERSION) {
        status = solve_dense_inversion(c, n_freq, freq, rhs_all.p, (int)rhs_fstride, (int)stride, sol.p, pol_mask);
        it_sys.assign(systems.size(), 0);
        tr_sys.assign(systems.size(), 0.0);
    } else {
        // 'iterative' keeps the frequency-independent T (atoms) or D (BEM) in
        // HBM; 'iterative on the fly' recomputes T in every product.
        c->use_stored = method == PXH_ITERATIVE && !bem;
        if (c->use_stored) prepare_stored_t(c);
        // omegaFQ (charges only): the FP64 pair kernel issues at the FP32 rate
        // here, and the restarts of the refinement cost more iterations than
        // the low-precision products save (1745-atom Na sphere, 61 w: 326
        // iterations per system in FP64, 804 refined with FP64 products, 891
        // with FP32 products), so a mixed request runs the batched FP64 solve
        if (!bem && c->atoms.model == PXH_WFQ && precision == PXH_MIXED) precision = PXH_FP64;
        const bool plain = precision == PXH_FP64 && c->opt.batch <= 1;
        const int prec = c->use_stored ? PXH_FP64 : precision;
        std::vector<double2> sigma;
        // the IEF operator form is reset whatever happens below
        struct BemOpGuard {
            pxh_ctx *c;
            ~BemOpGuard() { c->bem_op = 0; }
        } guard{c};
        try {
            if (!plain && !x0_given && c->opt.multishift &&
                shift_family(c, n_freq, freq, variables == nullptr, sigma)) {
                if (bem && c->mesh.variant == PXH_IEFPCM) {
                    c->bem_op = 1;  // unknown u = S q
                    ief_u = true;
                }
                // one Krylov basis for all frequencies; the few systems it
                // leaves above tol continue in the batched solver from there
                std::vector<int> rest;
                status = run_multishift(c, nv, systems, n_freq, sigma, freq, rhs_all.p, (int)rhs_fstride, (int)stride,
                                        tol, kdim,
                                        sol.p, it_sys, tr_sys, t_fin, rest);
                t_adm.assign(systems.size(), 0.0);
                if (!rest.empty()) {
                    std::vector<System> sub;
                    for (int k : rest) sub.push_back(systems[k]);
                    std::vector<int> it2;
                    std::vector<double> tr2, ta2, tf2;
                    const double t_ms = *std::max_element(t_fin.begin(), t_fin.end());
                    status = run_batched(c, nv, sub, rhs_all.p, (int)rhs_fstride, (int)stride, freq, tol, kdim, prec,
                                         false, true,
                                         sol.p, it2, tr2, ta2, tf2);
                    for (size_t i = 0; i < rest.size(); ++i) {
                        it_sys[rest[i]] += it2[i];
                        tr_sys[rest[i]] = tr2[i];
                        t_fin[rest[i]] = t_ms + tf2[i];
                    }
                }
            } else {
                status = run_batched(c, nv, systems, rhs_all.p, (int)rhs_fstride, (int)stride, freq, tol, kdim, prec,
                                     plain, x0_given,
                                     sol.p, it_sys, tr_sys, t_adm, t_fin);
            }
            if (ief_u) ief_polar_weights(c);
        } catch (...) {
            c->use_stored = false;
            c->atoms.Tstore.release();
            throw;
        }
        c->use_stored = false;
        c->atoms.Tstore.release();
    }
    // service time of each frequency point: admission of its first field to
    // convergence of its last (iterative paths), p50 / p99 in ms
    if (!t_fin.empty()) {
        std::vector<double> a(n_freq, 1e300), d(n_freq, 0.0), lat;
        for (size_t k = 0; k < systems.size(); ++k) {
            a[systems[k].ifreq] = std::min(a[systems[k].ifreq], t_adm[k]);
            d[systems[k].ifreq] = std::max(d[systems[k].ifreq], t_fin[k]);
        }
        for (int f = 0; f < n_freq; ++f)
            if (a[f] < 1e299) lat.push_back(1e3 * (d[f] - a[f]));
        std::sort(lat.begin(), lat.end());
        if (!lat.empty()) {
            auto pct = [&](double q) { return lat[std::min(lat.size() - 1, (size_t)std::floor(q * (lat.size() - 1) + 0.5))]; };
            c->stats.v[14] = pct(0.50);
            c->stats.v[15] = pct(0.99);
        }
    }
    // polarizabilities and results
    const int nb = 64;
    DevBuf<double2> &part = ws.part_keep;
    part.alloc((size_t)nb * 9 * n_freq + (size_t)9 * n_freq);
    const double4 *posp = nullptr;
    const double *area = nullptr;
    DevBuf<double4> &bpos = ws.bpos_keep;
    DevBuf<double> &barea = ws.barea_keep;
    if (bem) {
        std::vector<double4> hp(nv);
        for (int i = 0; i < nv; ++i)
            hp[i] = make_double4(c->mesh.h_coord[3 * i], c->mesh.h_coord[3 * i + 1], c->mesh.h_coord[3 * i + 2], 0.0);
        bpos.upload(hp.data(), hp.size(), st);
        barea.upload(c->mesh.h_area.data(), nv, st);
        posp = ief_u ? c->mesh.ief_z.p : bpos.p;
        area = barea.p;
    } else {
        posp = c->atoms.pos.p;
    }
    const int n_sites = bem ? nv : c->atoms.n;
    const int fmu = (!bem && c->atoms.model == PXH_WFQFMU) ? 1 : 0;
    const int dpcm = (bem && c->mesh.variant == PXH_DPCM) ? 1 : 0;
    k_polar_partial<<<dim3(nb, n_freq), RB, 0, st>>>(n_sites, nv, fmu, dpcm, posp, area, sol.p, stride, part.p);
    double2 *psum = part.p + (size_t)nb * 9 * n_freq;
    k_polar_sum<<<(9 * n_freq + 255) / 256, 256, 0, st>>>(nb, n_freq, part.p, psum);
    PXH_HIP(hipGetLastError());
    std::vector<double2> hp((size_t)9 * n_freq);
    PXH_HIP(hipMemcpyAsync(hp.data(), psum, sizeof(double2) * hp.size(), hipMemcpyDeviceToHost, st));
    if (variables)
        PXH_HIP(hipMemcpyAsync(variables, sol.p, sizeof(double2) * stride * n_freq, hipMemcpyDeviceToHost, st));
    PXH_HIP(hipStreamSynchronize(st));
    timer_drain(c);
    for (int f = 0; f < n_freq; ++f) {
        double pol[18];
        for (int k = 0; k < 9; ++k) {
            pol[2 * k] = hp[(size_t)9 * f + k].x / e0;
            pol[2 * k + 1] = hp[(size_t)9 * f + k].y / e0;
        }
        if (polar) std::memcpy(polar + 18 * (size_t)f, pol, sizeof(pol));
        if (results) {
            double r[20];
            results_row(freq[f], pol, r);
            for (int k = 0; k < 20; ++k) results[f + (size_t)n_freq * k] = r[k];
        }
    }
    if (iters || true_res) {
        for (int f = 0; f < 3 * n_freq; ++f) {
            if (iters) iters[f] = 0;
            if (true_res) true_res[f] = 0.0;
        }
        for (size_t k = 0; k < systems.size(); ++k) {
            const int o = systems[k].pol + 3 * systems[k].ifreq;
            if (iters && k < it_sys.size()) iters[o] = it_sys[k];
            if (true_res && k < tr_sys.size()) true_res[o] = tr_sys[k];
        }
    }
    // the product scratch is per call (a context holds only its target between calls)
    c->work_part.release();
    c->work_t.release();
    c->work_xh.release();
    c->work_xl.release();
    c->stats.v[5] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return status;
}

size_t estimate_solve_bytes(const pxh_ctx *c, int n_freq, int pol_mask, int method, int kdim, int precision) {
    const bool bem = c->target == PXH_BEM;
    const size_t nv = bem ? (size_t)c->mesh.n : (size_t)c->atoms.n_var();
    const size_t nsys = (size_t)n_freq * (size_t)__builtin_popcount(pol_mask & 7);
    size_t b = 2 * nv * 3 * (size_t)n_freq * sizeof(double2);  // right-hand sides and solutions
    if (method == PXH_INVERSION) {
        const size_t k = std::min<size_t>(std::min<size_t>(n_freq, kMaxFreqChunk),
                                          std::max<size_t>(1, kDenseChunkBytes / (nv * nv * sizeof(double2))));
        b += k * nv * nv * 2 * sizeof(double2) + 2 * nv * nv * sizeof(double);
    } else {
        // omegaFQ (charges only): the FP64 pair kernel issues at the FP32 rate
        // here, and the restarts of the refinement cost more iterations than
        // the low-precision products save (1745-atom Na sphere, 61 w: 326
        // iterations per system in FP64, 804 refined with FP64 products, 891
        // with FP32 products), so a mixed request runs the batched FP64 solve
        if (!bem && c->atoms.model == PXH_WFQ && precision == PXH_MIXED) precision = PXH_FP64;
        const bool plain = precision == PXH_FP64 && c->opt.batch <= 1;
        const size_t B = plain ? 1 : std::max<size_t>(1, std::min<size_t>(c->opt.batch, nsys));
        const size_t m = std::min<size_t>(std::max(kdim, 1), nv);
        const bool vf = precision == PXH_MIXED && !plain;
        b += B * (m + 1) * nv * (vf ? sizeof(float2) : sizeof(double2));  // Krylov basis
        b += 8 * B * nv * sizeof(double2);                                  // slot vectors
        b += 9 * B * nv * sizeof(double2);                                  // product partials and t
        if (!bem && method == PXH_ITERATIVE) b += nv * nv * sizeof(double);  // stored T
    }
    return b + b / 10;
}

void release_solver_ws(pxh_ctx *c) {
    delete static_cast<SolverWs *>(c->solver_ws);
    c->solver_ws = nullptr;
}

}  // namespace pxh
