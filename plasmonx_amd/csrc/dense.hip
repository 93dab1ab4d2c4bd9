// Dense per-frequency solves on gfx950: assembly of A(omega) for every
// frequency, a blocked complex LU with partial pivoting batched over the
// frequencies (LAPACK ZGETRF/ZGETRS semantics, called by the reference at
// src/solvers/algorithm_module.f90:510-541), the BEM charge constraint
// (:266-319), and GMRES on the stored matrix (solve_w_iter_mem, :334-410).
//
// Layout: one n x n column-major complex (double2) matrix per frequency,
// frequencies contiguous; every kernel takes blockIdx.z as the frequency.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <cmath>
#include <type_traits>
#include <vector>

#include "pxh_internal.h"

namespace pxh {

namespace {

constexpr int NB = 32;  // panel width

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    const double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
// Pivot magnitude |re| + |im|, as LAPACK's IZAMAX (via DCABS1) ranks the
// candidates of ZGETRF.
__device__ __forceinline__ double cabs1(double2 a) { return fabs(a.x) + fabs(a.y); }

// The panel's row interchanges as one permutation of the rows they touch,
// for the other columns: window row k + i takes the old row src_w[i]; rows
// below the window that the interchanges reach, dst_o[j] (j < m_o), take
// the old window row src_o[j].
struct SwapSet {
    int src_w[NB];
    int dst_o[NB];
    int src_o[NB];
    int m_o;
};

// Net permutation of the interchanges pv[0..w) (pivot rows, global indices)
// of the window rows [w0, w0 + w), w <= 32, by one wave: lane i < w holds
// window row w0 + i, lanes from 32 the rows below the window in order of
// first touch; each lane carries the old row now at its position.  Writes
// src_w[0..w), dst_o / src_o[0..m_o) and m_o as SwapSet describes.
__device__ void net_permutation(int lane, int w0, int w, const int *pv, int *src_w, int *dst_o, int *src_o,
                                int *m_o) {
    int row = lane < w ? w0 + lane : -1;
    int cont = row;
    int extra = 0;
    for (int i = 0; i < w; ++i) {
        const int p = pv[i];
        if (p == w0 + i) continue;
        int slot;
        if (p < w0 + w) {
            slot = p - w0;
        } else {
            const unsigned long long hit = __ballot(lane >= 32 && row == p);
            if (hit) {
                slot = __ffsll((long long)hit) - 1;
            } else {
                slot = 32 + extra++;
                if (lane == slot) row = cont = p;
            }
        }
        const int ci = __shfl(cont, i, 64), cp = __shfl(cont, slot, 64);
        if (lane == i) cont = cp;
        if (lane == slot) cont = ci;
    }
    if (lane < w) src_w[lane] = cont;
    if (lane >= 32 && lane < 32 + extra) {
        dst_o[lane - 32] = row;
        src_o[lane - 32] = cont;
    }
    if (lane == 0) *m_o = extra;
}

// a[ld dst[j]] = a[ld src[j]] for j < m, where no destination row is a source
// row (a net permutation's rows below its window take old window rows): the
// loads go in batches of 8 ahead of their stores, instead of one load-store
// round trip per row.
__device__ __forceinline__ void move_rows(double2 *a, size_t ld, const int *src, const int *dst, int m) {
#pragma unroll
    for (int j0 = 0; j0 < NB; j0 += 8) {
        if (j0 >= m) break;
        double2 t0 = a[ld * src[j0]], t1 = a[ld * src[min(j0 + 1, m - 1)]], t2 = a[ld * src[min(j0 + 2, m - 1)]],
                t3 = a[ld * src[min(j0 + 3, m - 1)]], t4 = a[ld * src[min(j0 + 4, m - 1)]], t5 = a[ld * src[min(j0 + 5, m - 1)]],
                t6 = a[ld * src[min(j0 + 6, m - 1)]], t7 = a[ld * src[min(j0 + 7, m - 1)]];
        a[ld * dst[j0]] = t0;
        if (j0 + 1 < m) a[ld * dst[j0 + 1]] = t1;
        if (j0 + 2 < m) a[ld * dst[j0 + 2]] = t2;
        if (j0 + 3 < m) a[ld * dst[j0 + 3]] = t3;
        if (j0 + 4 < m) a[ld * dst[j0 + 4]] = t4;
        if (j0 + 5 < m) a[ld * dst[j0 + 5]] = t5;
        if (j0 + 6 < m) a[ld * dst[j0 + 6]] = t6;
        if (j0 + 7 < m) a[ld * dst[j0 + 7]] = t7;
    }
}

// Layout of the factorisation.  The matrices A(w) are assembled column-major
// (element (i, j) at i + n j).  The kernels below factorise B = A^T, i.e. the
// same memory read row-major (element (r, c) of B at c + n r), with row
// interchanges on B: P B = L U.  A row of B is then contiguous, so the
// interchanges, which move whole rows, move contiguous segments.  In A's
// terms the memory ends up holding U^T in the lower triangle (with the
// diagonal) and L^T (unit diagonal) strictly above it, and
//   A x = b  <=>  U^T (L^T (P x)) = b,
// solved by k_lu_solve and the scatter x[perm[i]] = (P x)[i].
//
// Panel factorisation of columns [k, k+nb) of B over rows [k, n), one
// workgroup per frequency.  Pivots (global row indices) go to piv[f][k..k+nb)
// and their net permutation to swaps[f].  Row swaps are applied inside the panel only.  One
// pass over the rows per column: scale column c, update the remaining panel
// columns with row c (staged in LDS), and find the pivot candidates of column
// c + 1 on the way.
constexpr int kPanelThreads = 512;
__global__ __launch_bounds__(kPanelThreads) void k_panel(int n, int k, int nb, double2 *__restrict__ A,
                                                         size_t lda_f, int *__restrict__ piv,
                                                         SwapSet *__restrict__ swaps, int *__restrict__ info) {
    const int f = blockIdx.z;
    double2 *Af = A + lda_f * f;
    __shared__ double sval[kPanelThreads];
    __shared__ int sidx[kPanelThreads];
    __shared__ double2 urow[NB];
    __shared__ int spiv[NB];
    const int tid = threadIdx.x;
    // pivot candidates of column k (first maximum, as IZAMAX)
    double best = -1.0;
    int bi = k;
    for (int r = k + tid; r < n; r += kPanelThreads) {
        const double v = cabs1(Af[k + (size_t)n * r]);
        if (v > best) {
            best = v;
            bi = r;
        }
    }
    for (int c = k; c < k + nb; ++c) {
        sval[tid] = best;
        sidx[tid] = bi;
        __syncthreads();
        for (int off = kPanelThreads / 2; off > 0; off >>= 1) {
            if (tid < off) {
                if (sval[tid + off] > sval[tid] || (sval[tid + off] == sval[tid] && sidx[tid + off] < sidx[tid])) {
                    sval[tid] = sval[tid + off];
                    sidx[tid] = sidx[tid + off];
                }
            }
            __syncthreads();
        }
        const int p = sidx[0];
        const double pv = sval[0];
        __syncthreads();
        if (tid == 0) {
            piv[(size_t)n * f + c] = p;
            spiv[c - k] = pv == 0.0 ? c : p;
            if (pv == 0.0 && info[f] == 0) info[f] = c + 1;
        }
        best = -1.0;
        bi = c + 1;
        if (pv == 0.0) {
            // singular column: nothing to eliminate; candidates of the next
            if (c + 1 < k + nb)
                for (int r = c + 1 + tid; r < n; r += kPanelThreads) {
                    const double v = cabs1(Af[(c + 1) + (size_t)n * r]);
                    if (v > best) {
                        best = v;
                        bi = r;
                    }
                }
            continue;
        }
        // swap rows c and p inside the panel, row c of the remaining columns to LDS
        if (p != c)
            for (int cc = k + tid; cc < k + nb; cc += kPanelThreads) {
                const double2 t = Af[cc + (size_t)n * c];
                Af[cc + (size_t)n * c] = Af[cc + (size_t)n * p];
                Af[cc + (size_t)n * p] = t;
            }
        __syncthreads();
        for (int cc = c + 1 + tid; cc < k + nb; cc += kPanelThreads) urow[cc - k] = Af[cc + (size_t)n * c];
        __syncthreads();
        const double2 d = Af[c + (size_t)n * c];
        for (int r = c + 1 + tid; r < n; r += kPanelThreads) {
            const double2 l = cdiv(Af[c + (size_t)n * r], d);
            Af[c + (size_t)n * r] = l;
            for (int cc = c + 1; cc < k + nb; ++cc) {
                const double2 pr = cmul(l, urow[cc - k]);
                double2 a = Af[cc + (size_t)n * r];
                a.x -= pr.x;
                a.y -= pr.y;
                Af[cc + (size_t)n * r] = a;
                if (cc == c + 1) {
                    const double v = cabs1(a);
                    if (v > best) {
                        best = v;
                        bi = r;
                    }
                }
            }
        }
        __syncthreads();
    }
    // net permutation of the panel's interchanges for the other columns
    if (tid < 64) {
        SwapSet &S = swaps[f];
        net_permutation(tid, k, nb, spiv, S.src_w, S.dst_o, S.src_o, &S.m_o);
    }
}

// The same panel factorisation with the panel held in registers: work-item t
// owns rows k + t + kPanelThreads i (i < R), and the panel is processed in
// sub-panels of SUB = min(8, 32 / R) columns (R SUB complex values per
// work-item; the column loop is unrolled, so SUB stays small).
// Per sub-panel: load it, bring it up to date with the sub-panels before it
// (their U rows by forward substitution in LDS, the rows below by the rank
// update with their L), factorise its columns one by one in registers (each
// wave publishes its best candidate row in LDS together with row c, so one
// barrier per column settles the pivot and the exchange), write it back, and
// apply its interchanges to the other columns of the panel as one net
// permutation.  Same pivots as k_panel; column c below the diagonal is scaled
// by the reciprocal of the pivot, as LAPACK's ZGETF2 does.
//
// PK: the panel has been copied to a column-major buffer (k_panel_pack:
// column c - k of matrix f at P + (f NB + c - k) ldp, rows from k), so that a
// column read is one contiguous run instead of one element per row of the
// row-major matrix; k_panel_unpack copies it back.
template <int R, bool PK>
__global__ __launch_bounds__(kPanelThreads) void k_panel_reg(int n, int k, int nb, double2 *__restrict__ A,
                                                             size_t lda_f, int *__restrict__ piv,
                                                             SwapSet *__restrict__ swaps, int *__restrict__ info,
                                                             double2 *__restrict__ P, int ldp) {
    constexpr int SUB = NB / R < 8 ? NB / R : 8;
    constexpr int W = kPanelThreads / 64;
    constexpr int kUpdUnroll = R <= 8 ? 4 : 1;  // L loads in flight per row in the rank update
    const int f = blockIdx.z;
    double2 *Af = A + lda_f * f;
    double2 *Pf = PK ? P + (size_t)f * NB * ldp : nullptr;
    // element (row r, column col) of the panel, rows >= k, k <= col < k + nb
    auto el = [&](int r, int col) -> double2 & {
        if constexpr (PK)
            return Pf[(size_t)(col - k) * ldp + (r - k)];
        else
            return Af[col + (size_t)n * r];
    };
    const int tid = threadIdx.x, lane = tid % 64, wid = tid / 64;
    __shared__ double wv[2][W];
    __shared__ int wi[2][W];
    __shared__ double2 cand[2][W][SUB], crow[2][SUB];  // by column parity
    __shared__ int spiv[NB];
    __shared__ double2 Lw[NB][NB + 1];  // unit lower block of the finished sub-panels, [row][col]
    __shared__ double2 Uw[NB][SUB];     // their U rows in the current sub-panel's columns
    __shared__ int ps_src_w[SUB], ps_dst[NB], ps_src[NB], ps_m;
    double2 v[R][SUB];
    for (int s0 = k; s0 < k + nb; s0 += SUB) {
        const int sb = min(SUB, k + nb - s0);
        const int done = s0 - k;
        // load the sub-panel (rows k..n)
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = k + tid + kPanelThreads * i;
#pragma unroll
            for (int j = 0; j < SUB; ++j)
                v[i][j] = (r < n && j < sb) ? el(r, s0 + j) : make_double2(0, 0);
        }
        if (done > 0) {
            // U rows [k, s0): forward substitution with the unit lower block
            for (int e = tid; e < done * done; e += kPanelThreads) {
                const int rr = e % done, cc = e / done;
                Lw[rr][cc] = el(k + rr, k + cc);
            }
            if (tid < done)
#pragma unroll
                for (int j = 0; j < SUB; ++j) Uw[tid][j] = v[0][j];
            __syncthreads();
            // wave w solves column w, lane rr holding row rr
            if (wid < sb) {
                double2 val = lane < done ? Uw[lane][wid] : make_double2(0, 0);
                for (int q = 0; q + 1 < done; ++q) {
                    const double2 xq = make_double2(__shfl(val.x, q, 64), __shfl(val.y, q, 64));
                    if (lane > q && lane < done) {
                        const double2 pr = cmul(Lw[lane][q], xq);
                        val.x -= pr.x;
                        val.y -= pr.y;
                    }
                }
                if (lane < done) Uw[lane][wid] = val;
            }
            __syncthreads();
            // rows s0..n: rank update with the finished sub-panels' L
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int r = k + tid + kPanelThreads * i;
                if (r >= s0 && r < n) {
#pragma unroll kUpdUnroll
                    for (int q = 0; q < done; ++q) {
                        const double2 l = el(r, k + q);
#pragma unroll
                        for (int j = 0; j < SUB; ++j) {
                            const double2 u = Uw[q][j];
                            v[i][j].x -= l.x * u.x - l.y * u.y;
                            v[i][j].y -= l.x * u.y + l.y * u.x;
                        }
                    }
                }
            }
            if (tid < done)
#pragma unroll
                for (int j = 0; j < SUB; ++j) v[0][j] = Uw[tid][j];
        }
        // factorise the sub-panel's columns
#pragma unroll
        for (int j = 0; j < SUB; ++j) {
            if (j < sb) {
                const int c = s0 + j, b = j & 1;
                double best = -1.0;
                int bi = c;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int r = k + tid + kPanelThreads * i;
                    if (r >= c && r < n) {
                        const double a = cabs1(v[i][j]);
                        if (a > best) {
                            best = a;
                            bi = r;
                        }
                    }
                }
                // first maximum of the wave; its owner publishes the row
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const double ob = __shfl_xor(best, off, 64);
                    const int oi = __shfl_xor(bi, off, 64);
                    if (ob > best || (ob == best && oi < bi)) {
                        best = ob;
                        bi = oi;
                    }
                }
                if (best >= 0.0) {
                    const int tw = (bi - k) % kPanelThreads, iw = (bi - k) / kPanelThreads;
#pragma unroll
                    for (int i = 0; i < R; ++i)
                        if (tid == tw && i == iw)
#pragma unroll
                            for (int jj = 0; jj < SUB; ++jj) cand[b][wid][jj] = v[i][jj];
                }
                if (lane == 0) {
                    wv[b][wid] = best;
                    wi[b][wid] = bi;
                }
                const int tc = (c - k) % kPanelThreads, ic = (c - k) / kPanelThreads;
#pragma unroll
                for (int i = 0; i < R; ++i)
                    if (tid == tc && i == ic)
#pragma unroll
                        for (int jj = 0; jj < SUB; ++jj) crow[b][jj] = v[i][jj];
                __syncthreads();
                // first maximum over the workgroup
                best = wv[b][0];
                bi = wi[b][0];
                int ws = 0;
#pragma unroll
                for (int w = 1; w < W; ++w)
                    if (wv[b][w] > best || (wv[b][w] == best && wi[b][w] < bi)) {
                        best = wv[b][w];
                        bi = wi[b][w];
                        ws = w;
                    }
                const int p = bi;
                if (tid == 0) {
                    piv[(size_t)n * f + c] = p;
                    spiv[c - k] = best == 0.0 ? c : p;
                    if (best == 0.0 && info[f] == 0) info[f] = c + 1;
                }
                if (best != 0.0) {
                    // exchange rows c and p
                    const double2 *prow = cand[b][ws];
                    const int tp = (p - k) % kPanelThreads, ip = (p - k) / kPanelThreads;
#pragma unroll
                    for (int i = 0; i < R; ++i) {
                        if (tid == tc && i == ic)
#pragma unroll
                            for (int jj = 0; jj < SUB; ++jj) v[i][jj] = prow[jj];
                        if (p != c && tid == tp && i == ip)
#pragma unroll
                            for (int jj = 0; jj < SUB; ++jj) v[i][jj] = crow[b][jj];
                    }
                    // scale column c below the diagonal, update the columns right of it
                    const double2 d = prow[j];
                    const double dd = d.x * d.x + d.y * d.y;
                    const double rdd = 1.0 / dd;
                    const double2 dinv = make_double2(d.x * rdd, -d.y * rdd);
#pragma unroll
                    for (int i = 0; i < R; ++i) {
                        const int r = k + tid + kPanelThreads * i;
                        if (r > c && r < n) {
                            const double2 l = cmul(v[i][j], dinv);
                            v[i][j] = l;
#pragma unroll
                            for (int jj = j + 1; jj < SUB; ++jj) {
                                const double2 u = prow[jj];
                                v[i][jj].x -= l.x * u.x - l.y * u.y;
                                v[i][jj].y -= l.x * u.y + l.y * u.x;
                            }
                        }
                    }
                }
            }
        }
        // write the sub-panel back
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = k + tid + kPanelThreads * i;
            if (r < n)
#pragma unroll
                for (int j = 0; j < SUB; ++j)
                    if (j < sb) el(r, s0 + j) = v[i][j];
        }
        __syncthreads();
        // its interchanges on the other columns of the panel
        if (nb > sb) {
            if (tid < 64) net_permutation(tid, s0, sb, spiv + done, ps_src_w, ps_dst, ps_src, &ps_m);
            __syncthreads();
            if (tid < nb - sb) {
                const int col = tid < done ? k + tid : k + tid + sb;
                // a[ld * row] = element (row, col)
                double2 *a = PK ? Pf + (size_t)(col - k) * ldp - k : Af + col;
                const size_t ld = PK ? 1 : (size_t)n;
                double2 *x = v[0];  // the sub-panel is written back; its registers are free
#pragma unroll
                for (int j = 0; j < SUB; ++j) x[j] = j < sb ? a[ld * ps_src_w[j]] : make_double2(0, 0);
                move_rows(a, ld, ps_src, ps_dst, ps_m);
#pragma unroll
                for (int j = 0; j < SUB; ++j)
                    if (j < sb) a[ld * (s0 + j)] = x[j];
            }
            __syncthreads();
        }
    }
    // net permutation of the panel's interchanges for the other columns
    if (tid < 64) {
        SwapSet &S = swaps[f];
        net_permutation(tid, k, nb, spiv, S.src_w, S.dst_o, S.src_o, &S.m_o);
    }
}

// Copy of the panel (rows [k, n), columns [k, k + nb)) between the row-major
// matrix and the column-major panel buffer, 64 rows per workgroup through an
// LDS tile: each row's nb columns are one contiguous read (or write), each
// column's 64 rows one contiguous write (or read).
template <bool TO_P>
__global__ __launch_bounds__(256) void k_panel_copy(int n, int k, int nb, double2 *__restrict__ A, size_t lda_f,
                                                    double2 *__restrict__ P, int ldp) {
    __shared__ double2 tile[NB][65];
    const int f = blockIdx.z, r0 = k + 64 * blockIdx.x;
    double2 *Af = A + lda_f * f;
    double2 *Pf = P + (size_t)f * NB * ldp;
    if (TO_P) {
        for (int e = threadIdx.x; e < 64 * NB; e += 256) {
            const int rr = e / NB, cc = e % NB, r = r0 + rr;
            if (r < n && cc < nb) tile[cc][rr] = Af[(k + cc) + (size_t)n * r];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * NB; e += 256) {
            const int cc = e / 64, rr = e % 64, r = r0 + rr;
            if (r < n && cc < nb) Pf[(size_t)cc * ldp + (r - k)] = tile[cc][rr];
        }
    } else {
        for (int e = threadIdx.x; e < 64 * NB; e += 256) {
            const int cc = e / 64, rr = e % 64, r = r0 + rr;
            if (r < n && cc < nb) tile[cc][rr] = Pf[(size_t)cc * ldp + (r - k)];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * NB; e += 256) {
            const int rr = e / NB, cc = e % NB, r = r0 + rr;
            if (r < n && cc < nb) Af[(k + cc) + (size_t)n * r] = tile[cc][rr];
        }
    }
}

// The panel's row interchanges applied to the columns outside [k, k+nb),
// and U12 = L11^-1 A12 (unit lower L11 of the panel) for the columns right of
// it; one work-item per column, the window of the column in registers (loops
// over the constant panel width, fully unrolled).  All old values a column
// needs are read before any of its rows is written.  Columns by mode:
//   kSwapAll    every column outside the panel;
//   kSwapRight  the columns right of the panel, whose window rows first take
//               the update by the previous panel [k - NB, k) (its L rows
//               read at their places before this panel's interchanges, its
//               U rows above the window), as the second panel of a pair;
//   kSwapLeft   the columns left of the panel (interchanges only), after
//               kSwapRight.
enum { kSwapAll = 0, kSwapRight = 1, kSwapLeft = 2 };
__global__ __launch_bounds__(128) void k_swap_trsm(int n, int k, int nb, double2 *__restrict__ A, size_t lda_f,
                                                   const SwapSet *__restrict__ swaps, int mode) {
    const int f = blockIdx.z;
    double2 *Af = A + lda_f * f;
    __shared__ double2 L[NB][NB];
    __shared__ double2 Lp[NB][NB + 1];  // kSwapRight: previous panel's L in this panel's window rows
    __shared__ SwapSet S;
    for (int e = threadIdx.x; e < (int)(sizeof(SwapSet) / sizeof(int)); e += blockDim.x)
        reinterpret_cast<int *>(&S)[e] = reinterpret_cast<const int *>(swaps + f)[e];
    __syncthreads();
    const bool trsm = mode != kSwapLeft && k + nb < n;  // then nb == NB
    if (trsm)
        for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
            const int r = e % NB, c = e / NB;
            L[r][c] = Af[(k + c) + (size_t)n * (k + r)];
        }
    if (mode == kSwapRight)
        for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
            const int r = e % NB, q = e / NB;
            Lp[r][q] = Af[(k - NB + q) + (size_t)n * S.src_w[r]];
        }
    __syncthreads();
    const int cidx = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = mode == kSwapAll ? n - nb : mode == kSwapRight ? n - k - nb : k;
    if (cidx >= ncol) return;
    const int col = mode == kSwapRight ? k + nb + cidx : (mode == kSwapLeft || cidx < k) ? cidx : cidx + nb;
    double2 *a = Af + col;  // column col of B, rows n apart
    double2 x[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) x[i] = i < nb ? a[(size_t)n * S.src_w[i]] : make_double2(0, 0);
    move_rows(a, n, S.src_o, S.dst_o, S.m_o);
    if (mode == kSwapRight)
        for (int q = 0; q < NB; ++q) {
            const double2 u = a[(size_t)n * (k - NB + q)];
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const double2 p = cmul(Lp[i][q], u);
                x[i].x -= p.x;
                x[i].y -= p.y;
            }
        }
    if (trsm && col > k) {
#pragma unroll
        for (int r = 0; r < NB; ++r)
#pragma unroll
            for (int q = r + 1; q < NB; ++q) {
                const double2 p = cmul(L[q][r], x[r]);
                x[q].x -= p.x;
                x[q].y -= p.y;
            }
    }
#pragma unroll
    for (int i = 0; i < NB; ++i)
        if (i < nb) a[(size_t)n * (k + i)] = x[i];
}

// Solve U^T (L^T y) = b in place (column-major view: U^T is the lower
// triangle with the diagonal, L^T the unit upper triangle) for the
// right-hand sides selected by mask; one workgroup per (rhs, frequency).
// Blocked by kSolveBlock columns: the diagonal block is staged in LDS, wave 0
// solves it with the values in its lanes (shuffles), then all work-items
// update the rows outside the block (one row each, the block's columns read
// together).
constexpr int kSolveBlock = 32, kSolveThreads = 1024;

// x[r] -= sum_j A[r][c0 + j] xb[j] for the rows r handled by this work-item
__device__ inline void lu_solve_update(int n, const double2 *__restrict__ Af, int c0, int cb, int r0, int r1,
                                       const double2 *xb, double2 *__restrict__ x) {
    for (int r = r0 + (int)threadIdx.x; r < r1; r += kSolveThreads) {
        const double2 *a = Af + r + (size_t)n * c0;
        double2 acc = make_double2(0, 0);
#pragma unroll 8
        for (int j = 0; j < cb; ++j) {
            const double2 p = cmul(a[(size_t)n * j], xb[j]);
            acc.x += p.x;
            acc.y += p.y;
        }
        x[r].x -= acc.x;
        x[r].y -= acc.y;
    }
}

__global__ __launch_bounds__(kSolveThreads) void k_lu_solve(int n, const double2 *__restrict__ A, size_t lda_f,
                                                            double2 *__restrict__ X, size_t ldx_f, int ldx,
                                                            int mask) {
    const int f = blockIdx.z, rcol = blockIdx.x;
    if (!((mask >> rcol) & 1)) return;
    const double2 *Af = A + lda_f * f;
    double2 *x = X + ldx_f * f + (size_t)ldx * rcol;
    const int tid = threadIdx.x, lane = tid % 64;
    __shared__ double2 xb[kSolveBlock];
    __shared__ double2 Ab[kSolveBlock][kSolveBlock + 1];  // [col][row]
    auto stage = [&](int c0, int cb) {
        for (int i = tid; i < cb * cb; i += kSolveThreads) {
            const int r = i % cb, col = i / cb;
            Ab[col][r] = Af[(c0 + r) + (size_t)n * (c0 + col)];
        }
        __syncthreads();
    };
    // lower with its diagonal, blocks top to bottom
    for (int c0 = 0; c0 < n; c0 += kSolveBlock) {
        const int cb = min(kSolveBlock, n - c0);
        stage(c0, cb);
        if (tid < 64) {
            double2 v = lane < cb ? x[c0 + lane] : make_double2(0, 0);
            for (int j = 0; j < cb; ++j) {
                if (lane == j) v = cdiv(v, Ab[j][j]);
                const double2 xj = make_double2(__shfl(v.x, j, 64), __shfl(v.y, j, 64));
                if (lane > j && lane < cb) {
                    const double2 p = cmul(Ab[j][lane], xj);
                    v.x -= p.x;
                    v.y -= p.y;
                }
            }
            if (lane < cb) {
                x[c0 + lane] = v;
                xb[lane] = v;
            }
        }
        __syncthreads();
        lu_solve_update(n, Af, c0, cb, c0 + cb, n, xb, x);
        __syncthreads();
    }
    // unit upper, blocks bottom to top
    for (int c1 = n; c1 > 0; c1 -= kSolveBlock) {
        const int c0 = max(0, c1 - kSolveBlock), cb = c1 - c0;
        stage(c0, cb);
        if (tid < 64) {
            double2 v = lane < cb ? x[c0 + lane] : make_double2(0, 0);
            for (int j = cb - 1; j >= 0; --j) {
                const double2 xj = make_double2(__shfl(v.x, j, 64), __shfl(v.y, j, 64));
                if (lane < j) {
                    const double2 p = cmul(Ab[j][lane], xj);
                    v.x -= p.x;
                    v.y -= p.y;
                }
            }
            if (lane < cb) {
                x[c0 + lane] = v;
                xb[lane] = v;
            }
        }
        __syncthreads();
        lu_solve_update(n, Af, c0, cb, 0, c0, xb, x);
        __syncthreads();
    }
}

// A_f = M + diag shift (atoms, homogeneous): q rows -z_q, mu rows Lambda.
__global__ void k_assemble_atoms(int n_at, int nv, int nt, int fmu, const double *__restrict__ M,
                                 const double4 *__restrict__ pos, const double *__restrict__ nbfrac,
                                 const SysParam *__restrict__ par, double2 *__restrict__ A, size_t lda_f) {
    const int f = blockIdx.z;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    if (r >= nv) return;
    double2 v = make_double2(M[r + (size_t)nv * col], 0.0);
    if (r == col) {
        if (r < n_at) {
            const double2 z = par[f].zq[(int)pos[r].w];
            v.x -= z.x;
            v.y -= z.y;
        } else if (fmu) {
            const int i = (r - n_at) / 3;
            double2 lam = make_double2(0, 0);
            for (int t = 0; t < nt; ++t) {
                lam.x += nbfrac[t + (size_t)nt * i] * par[f].inv_alpha[t].x;
                lam.y += nbfrac[t + (size_t)nt * i] * par[f].inv_alpha[t].y;
            }
            v = lam;  // wfqfmu_module.f90:525-528 sets the mu diagonal
        }
    }
    A[lda_f * f + r + (size_t)nv * col] = v;
}

// Alloys: q rows of A_f = 1/2 (K + H(omega_f)) T(q rows), via the K-bar CSR.
__global__ void k_assemble_hetero_q(int n_at, int nv, int nt, const int *__restrict__ rowptr,
                                    const int *__restrict__ col, const double *__restrict__ kij,
                                    const double *__restrict__ kji, const double4 *__restrict__ pos,
                                    const double *__restrict__ T, const SysParam *__restrict__ par,
                                    double2 *__restrict__ A, size_t lda_f) {
    const int f = blockIdx.z;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (c >= nv) return;
    const int ti = (int)pos[i].w;
    const double tic = T[i + (size_t)nv * c];
    double2 acc = make_double2(0, 0);
    for (int z = rowptr[i]; z < rowptr[i + 1]; ++z) {
        const int j = col[z];
        const double2 wr = par[f].wratio[ti + nt * (int)pos[j].w];
        const double d = T[j + (size_t)nv * c] - tic;
        acc.x += 0.5 * (kij[z] + wr.x * kji[z]) * d;
        acc.y += 0.5 * wr.y * kji[z] * d;
    }
    double2 v = acc;
    if (c == i) {
        const double2 z = par[f].zq[ti];
        v.x -= z.x;
        v.y -= z.y;
    }
    A[lda_f * f + i + (size_t)nv * c] = v;
}

// BEM: DPCM A = D + lambda I; IEF A = -(D + lambda I) S A^-1.
__global__ void k_assemble_bem_dpcm(int n, const double *__restrict__ D, const double2 *__restrict__ lam,
                                    double2 *__restrict__ A, size_t lda_f) {
    const int f = blockIdx.z;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (r >= n) return;
    double2 v = make_double2(D[r + (size_t)n * c], 0.0);
    if (r == c) {
        v.x += lam[f].x;
        v.y += lam[f].y;
    }
    A[lda_f * f + r + (size_t)n * c] = v;
}

__global__ __launch_bounds__(256) void k_assemble_bem_ief(int n, const double *__restrict__ D,
                                                          const double *__restrict__ SA,
                                                          const double2 *__restrict__ lam, double2 *__restrict__ A,
                                                          size_t lda_f) {
    // (D S)(r, c) computed with a 64 x 64 tile per workgroup; then - (DS + lambda S)
    const int f = blockIdx.z;
    const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
    __shared__ double sD[64][33], sS[32][65];
    const int tid = threadIdx.x, tr = tid % 16, tc = tid / 16;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < n; k0 += 32) {
        for (int e = tid; e < 64 * 32; e += 256) {
            const int r = e % 64, q = e / 64;
            sD[r][q] = (r0 + r < n && k0 + q < n) ? D[(r0 + r) + (size_t)n * (k0 + q)] : 0.0;
        }
        for (int e = tid; e < 32 * 64; e += 256) {
            const int q = e % 32, cc = e / 32;
            sS[q][cc] = (k0 + q < n && c0 + cc < n) ? SA[(k0 + q) + (size_t)n * (c0 + cc)] : 0.0;
        }
        __syncthreads();
        for (int q = 0; q < 32; ++q)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) acc[a][b] += sD[tr + 16 * a][q] * sS[q][tc + 16 * b];
        __syncthreads();
    }
    const double2 l = lam[f];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            const int r = r0 + tr + 16 * a, c = c0 + tc + 16 * b;
            if (r < n && c < n) {
                const double s = SA[r + (size_t)n * c];
                A[lda_f * f + r + (size_t)n * c] = make_double2(-(acc[a][b] + l.x * s), -(l.y * s));
            }
        }
}

// The row interchanges of a factorisation of B as one permutation: perm[f][i]
// is the row of B that P B holds at i (the interchanges applied in order to
// the identity), so x = P^T z puts z[i] at x[perm[i]].  The pivots are staged through LDS in chunks so the serial pass
// of work-item 0 touches LDS only; the permutation itself lives in LDS when
// it fits.
constexpr int kPermLds = 16384, kPivChunk = 2048;
__global__ __launch_bounds__(256) void k_perm_from_piv(int n, const int *__restrict__ piv, int *__restrict__ perm) {
    __shared__ int sp[kPermLds];
    __shared__ int pc[kPivChunk];
    const int f = blockIdx.x;
    const int *pf = piv + (size_t)n * f;
    int *out = perm + (size_t)n * f;
    const bool in_lds = n <= kPermLds;
    int *q = in_lds ? sp : out;
    for (int i = threadIdx.x; i < n; i += 256) q[i] = i;
    __syncthreads();
    for (int c0 = 0; c0 < n; c0 += kPivChunk) {
        const int cn = min(kPivChunk, n - c0);
        for (int i = threadIdx.x; i < cn; i += 256) pc[i] = pf[c0 + i];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int c = 0; c < cn; ++c) {
                const int p = pc[c];
                const int t = q[c0 + c];
                q[c0 + c] = q[p];
                q[p] = t;
            }
        __syncthreads();
    }
    if (in_lds)
        for (int i = threadIdx.x; i < n; i += 256) out[i] = sp[i];
}

// Z[f][k] = b_k for the selected right-hand sides (zero otherwise).
__global__ void k_copy_rhs(int n, int mask, const double2 *__restrict__ rhs, size_t ld_rhs_f,
                           double2 *__restrict__ Z, size_t ldz_f) {
    const int f = blockIdx.z;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (r >= n) return;
    const bool on = (mask >> k) & 1;
    Z[ldz_f * f + (size_t)n * k + r] = on ? rhs[ld_rhs_f * f + (size_t)n * k + r] : make_double2(0, 0);
}

// x = P^T z: X[f][k][perm_f[i]] = Z[f][k][i].
__global__ void k_scatter_rows(int n, const int *__restrict__ perm, const double2 *__restrict__ Z, size_t ldz_f,
                               double2 *__restrict__ X, size_t ldx_f) {
    const int f = blockIdx.z;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (r >= n) return;
    X[ldx_f * f + (size_t)n * k + perm[(size_t)n * f + r]] = Z[ldz_f * f + (size_t)n * k + r];
}

__global__ void k_fill_one(int n, double2 *__restrict__ X, size_t ldx_f) {
    const int f = blockIdx.z;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) X[ldx_f * f + r] = make_double2(1.0, 0.0);
}

// Charge constraint: x_k -= x1 * sum(x_k) / sum(x1)   (algorithm_module.f90:306-311),
// for the field directions k of pol_mask (the others hold no solution)
__global__ __launch_bounds__(256) void k_charge_constraint(int n, int pol_mask, double2 *__restrict__ X,
                                                           size_t ldx_f, const double2 *__restrict__ X1,
                                                           size_t ldx1_f) {
    const int f = blockIdx.z, k = blockIdx.x;
    if (!((pol_mask >> k) & 1)) return;
    double2 *x = X + ldx_f * f + (size_t)n * k;
    const double2 *x1 = X1 + ldx1_f * f;
    __shared__ double s[4][256];
    double a = 0, b = 0, c = 0, d = 0;
    for (int r = threadIdx.x; r < n; r += 256) {
        a += x[r].x;
        b += x[r].y;
        c += x1[r].x;
        d += x1[r].y;
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    s[2][threadIdx.x] = c;
    s[3][threadIdx.x] = d;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int q = 0; q < 4; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + off];
        __syncthreads();
    }
    const double2 sx = make_double2(s[0][0], s[1][0]), s1 = make_double2(s[2][0], s[3][0]);
    const double2 ratio = cdiv(sx, s1);
    for (int r = threadIdx.x; r < n; r += 256) {
        const double2 p = cmul(x1[r], ratio);
        x[r].x -= p.x;
        x[r].y -= p.y;
    }
}

// Stored-matrix operator: y_s = A_{f(s)} x_s.
__global__ void k_dense_matvec(int n, const double2 *__restrict__ A, size_t lda_f, const int *__restrict__ slots,
                               const int *__restrict__ slot_freq, const double2 *__restrict__ x, int ld,
                               double2 *__restrict__ y) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = slots[blockIdx.y];
    if (r >= n) return;
    const double2 *Af = A + lda_f * slot_freq[s];
    const double2 *xs = x + (size_t)ld * s;
    double2 acc = make_double2(0, 0);
    for (int c = 0; c < n; ++c) {
        const double2 p = cmul(Af[r + (size_t)n * c], xs[c]);
        acc.x += p.x;
        acc.y += p.y;
    }
    y[(size_t)ld * s + r] = acc;
}

}  // namespace

// The context's rocBLAS handle (created on first use, bound to the context's
// stream; atomics off, so that every result is reproducible bit for bit).
rocblas_handle blas_handle(pxh_ctx *c) {
    if (!c->blas) {
        rocblas_handle h = nullptr;
        PXH_CHECK(rocblas_create_handle(&h) == rocblas_status_success, PXH_EHIP, "rocblas_create_handle failed");
        c->blas = h;
        PXH_CHECK(rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed) == rocblas_status_success, PXH_EHIP,
                  "rocblas_set_atomics_mode failed");
    }
    rocblas_handle h = static_cast<rocblas_handle>(c->blas);
    PXH_CHECK(rocblas_set_stream(h, c->stream) == rocblas_status_success, PXH_EHIP, "rocblas_set_stream failed");
    return h;
}

namespace {

// Right-looking blocked LU of nf matrices with partial pivoting.  Panels go in
// pairs where a trailing part follows: the second panel's columns take the
// first panel's update alone, the second panel's window rows of the columns
// right of it take it while their U rows are solved, and the rest of the
// trailing matrix takes both panels in one update of depth 2 NB, so it is
// read and written once per 2 NB columns.
void lu_factor(pxh_ctx *c, int n, int nf, double2 *A, int *piv, SwapSet *swaps, int *info) {
    hipStream_t st = c->stream;
    const size_t ldf = (size_t)n * n;
    // panels of more than 4 kPanelThreads rows are factorised in the
    // column-major panel buffer (contiguous column reads)
    const bool pack = n > 4 * kPanelThreads;
    double2 *P = nullptr;
    if (pack) {
        c->work_panel.alloc((size_t)nf * NB * n);
        P = c->work_panel.p;
    }
    auto panel = [&](int k, int nb) {
        const int rows = n - k;
        auto reg = [&](auto r_tag, auto pk_tag) {
            constexpr int R = decltype(r_tag)::value;
            constexpr bool PK = decltype(pk_tag)::value;
            k_panel_reg<R, PK><<<dim3(1, 1, nf), kPanelThreads, 0, st>>>(n, k, nb, A, ldf, piv, swaps, info, P, n);
        };
        using T = std::true_type;
        using F = std::false_type;
        if (rows <= kPanelThreads)
            reg(std::integral_constant<int, 1>{}, F{});
        else if (rows <= 2 * kPanelThreads)
            reg(std::integral_constant<int, 2>{}, F{});
        else if (rows <= 4 * kPanelThreads)
            reg(std::integral_constant<int, 4>{}, F{});
        else if (rows <= 24 * kPanelThreads) {
            const dim3 cg((rows + 63) / 64, 1, nf);
            k_panel_copy<true><<<cg, 256, 0, st>>>(n, k, nb, A, ldf, P, n);
            if (rows <= 8 * kPanelThreads)
                reg(std::integral_constant<int, 8>{}, T{});
            else if (rows <= 12 * kPanelThreads)
                reg(std::integral_constant<int, 12>{}, T{});
            else if (rows <= 16 * kPanelThreads)
                reg(std::integral_constant<int, 16>{}, T{});
            else if (rows <= 20 * kPanelThreads)
                reg(std::integral_constant<int, 20>{}, T{});
            else
                reg(std::integral_constant<int, 24>{}, T{});
            k_panel_copy<false><<<cg, 256, 0, st>>>(n, k, nb, A, ldf, P, n);
        } else {
            k_panel<<<dim3(1, 1, nf), kPanelThreads, 0, st>>>(n, k, nb, A, ldf, piv, swaps, info);
        }
    };
    // The trailing updates are plain complex GEMMs, done by rocBLAS
    // (zgemm_strided_batched over the chunk's matrices).  In B = A^T
    // row-major, rows [r0, n) and columns [r0, c_end) take
    // B[r][c] -= sum_d B[r][k + d] B[k + d][c]; the same memory read
    // column-major is X = B^T, and the update is
    // X(c0:c_end, r0:n) -= X(c0:c_end, k:k+depth) X(k:k+depth, r0:n).
    // Each launch is timed on the context's stream, with its algorithmic
    // flops: 8 real flop per complex multiply-add.
    rocblas_handle h = blas_handle(c);
    auto update = [&](int k, int depth, int r0, int c0, int c_end) {
        const int rows = n - r0, cols = c_end - c0;
        if (rows <= 0 || cols <= 0) return;
        const rocblas_double_complex minus_one(-1.0, 0.0), one(1.0, 0.0);
        const rocblas_stride sf = (rocblas_stride)ldf;
        const auto *X = reinterpret_cast<rocblas_double_complex *>(A);
        auto *Xw = reinterpret_cast<rocblas_double_complex *>(A);
        const int t = timer_begin(c);
        const rocblas_status rs = rocblas_zgemm_strided_batched(
            h, rocblas_operation_none, rocblas_operation_none, cols, rows, depth, &minus_one,
            X + c0 + (size_t)n * k, n, sf, X + k + (size_t)n * r0, n, sf, &one, Xw + c0 + (size_t)n * r0, n, sf, nf);
        timer_end(c, t, 19);
        PXH_CHECK(rs == rocblas_status_success, PXH_EHIP, "rocblas_zgemm_strided_batched failed");
        c->stats.v[20] += 1;
        c->stats.v[21] += 8.0 * rows * cols * depth * nf;
    };
    // trailing update of the rows and columns from k + depth
    auto trailing = [&](int k, int depth) { update(k, depth, k + depth, k + depth, n); };
    int k = 0;
    while (k < n) {
        const int nb = std::min(NB, n - k);
        panel(k, nb);
        if (n > nb)
            k_swap_trsm<<<dim3((n - nb + 127) / 128, 1, nf), 128, 0, st>>>(n, k, nb, A, ldf, swaps, kSwapAll);
        const int rest = n - k - nb;
        if (rest > NB) {
            // second panel of the pair, with a trailing part after it
            const int k2 = k + NB, rest2 = rest - NB;
            update(k, NB, k + NB, k + NB, k2 + NB);
            panel(k2, NB);
            k_swap_trsm<<<dim3((rest2 + 127) / 128, 1, nf), 128, 0, st>>>(n, k2, NB, A, ldf, swaps, kSwapRight);
            k_swap_trsm<<<dim3((k2 + 127) / 128, 1, nf), 128, 0, st>>>(n, k2, NB, A, ldf, swaps, kSwapLeft);
            trailing(k, 2 * NB);
            k += 2 * NB;
        } else {
            if (rest > 0) trailing(k, nb);
            k += NB;
        }
    }
    PXH_HIP(hipGetLastError());
}

// Assemble A(omega_f) for frequencies [f0, f0 + nf) into A.
void assemble(pxh_ctx *c, int f0, int nf, const double *freq, double2 *A, const double *M, DevBuf<double> &T,
              DevBuf<SysParam> &par, DevBuf<double2> &lam) {
    hipStream_t st = c->stream;
    if (c->target == PXH_BEM) {
        const Mesh &m = c->mesh;
        const int n = m.n;
        std::vector<double2> hl(nf);
        for (int f = 0; f < nf; ++f) {
            const double *e = c->eps.data() + 2 * (size_t)(f0 + f);
            const double es = m.eps_solvent;
            const double nr = es + e[0], ni = e[1], dr = es - e[0], di = -e[1];
            const double den = dr * dr + di * di;
            hl[f] = make_double2(2.0 * M_PI * (nr * dr + ni * di) / den, 2.0 * M_PI * (ni * dr - nr * di) / den);
        }
        lam.upload(hl.data(), nf, st);
        if (m.variant == PXH_DPCM)
            k_assemble_bem_dpcm<<<dim3((n + 255) / 256, n, nf), 256, 0, st>>>(n, m.D.p, lam.p, A, (size_t)n * n);
        else
            k_assemble_bem_ief<<<dim3((n + 63) / 64, (n + 63) / 64, nf), 256, 0, st>>>(n, m.D.p, m.SA.p, lam.p, A,
                                                                                        (size_t)n * n);
        PXH_HIP(hipStreamSynchronize(st));
        return;
    }
    Atoms &a = c->atoms;
    const int nv = a.n_var();
    std::vector<SysParam> hp(nf);
    for (int f = 0; f < nf; ++f) hp[f] = make_sysparam(c, freq[f0 + f], f0 + f);
    par.upload(hp.data(), nf, st);
    k_assemble_atoms<<<dim3((nv + 255) / 256, nv, nf), 256, 0, st>>>(
        a.n, nv, a.nt, a.model == PXH_WFQFMU, M, a.pos.p, a.d_nbfrac.p, par.p, A, (size_t)nv * nv);
    if (a.hetero)
        k_assemble_hetero_q<<<dim3((nv + 255) / 256, a.n, nf), 256, 0, st>>>(
            a.n, nv, a.nt, a.kb_rowptr.p, a.kb_col.p, a.kb_kij.p, a.kb_kji.p, a.pos.p, T.p, par.p, A,
            (size_t)nv * nv);
    PXH_HIP(hipGetLastError());
    PXH_HIP(hipStreamSynchronize(st));
}

// Dense T of the atoms (for the alloy q rows), n_var x n_var.
void dense_t(pxh_ctx *c, DevBuf<double> &T) {
    const size_t nv = c->atoms.n_var();
    T.alloc(nv * nv);
    dense_t_atoms(c, T.p, false);
}

// Frequencies factorised together: at most kMaxFreqChunk and
// kDenseChunkBytes of matrices, within half of the free HBM and within the
// context's budget.
size_t freq_chunk(pxh_ctx *c, int n, int n_freq) {
    size_t free_b = 0, total_b = 0;
    PXH_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t per = (size_t)n * n * sizeof(double2) * 2 + 1;
    size_t k = std::max<size_t>(1, (free_b / 2) / per);
    if (c->budget) k = std::max<size_t>(1, std::min(k, c->budget / per));
    k = std::min(k, std::max<size_t>(1, kDenseChunkBytes / ((size_t)n * n * sizeof(double2))));
    return std::min<size_t>(std::min<size_t>(k, kMaxFreqChunk), (size_t)n_freq);
}

}  // namespace

void constant_matrix_atoms(pxh_ctx *c, double *d_m);

int solve_dense_inversion(pxh_ctx *c, int n_freq, const double *freq, const double2 *d_rhs_all, int rhs_fstride,
                          int rhs_stride,
                          double2 *d_sol, int pol_mask) {
    hipStream_t st = c->stream;
    const bool bem = c->target == PXH_BEM;
    const int n = bem ? c->mesh.n : c->atoms.n_var();
    // homogeneous omegaFQ: one eigendecomposition for every frequency when it
    // is estimated faster than the LUs (the LU below when T is not positive
    // definite)
    if (!bem && rhs_fstride == 0 && spectral_applicable(c) && spectral_pays(c, n, n_freq)) {
        DeviceExclusive alone;  // no other work of the process on the device meanwhile
        const bool ok = spectral_decompose(c, d_rhs_all, pol_mask, true);
        if (ok) {
            spectral_charges(c, n_freq, freq, d_sol, (size_t)rhs_stride);
            c->spectral.release();
            return PXH_OK;
        }
    }
    DevBuf<double> Mtmp, T;
    const double *Mp = nullptr;
    if (!bem) {
        // the constant matrix depends on the target alone: kept with it up
        // to kKeepMconstBytes, else built for this call
        Atoms &At = c->atoms;
        if (!At.Mconst_valid) {
            const bool keep = (size_t)n * n * sizeof(double) <= kKeepMconstBytes;
            DevBuf<double> &dst = keep ? At.Mconst : Mtmp;
            dst.alloc((size_t)n * n);
            constant_matrix_atoms(c, dst.p);
            At.Mconst_valid = keep;
            Mp = dst.p;
        } else {
            Mp = At.Mconst.p;
        }
        if (c->atoms.hetero) dense_t(c, T);
    }
    const int chunk = (int)freq_chunk(c, n, n_freq);
    // the chunk's matrices and pivots stay with the context between calls
    // when they take at most kKeepLuBytes (released at the end otherwise)
    DevBuf<double2> &A = c->work_lu;
    DevBuf<int> info;
    DevBuf<SysParam> par;
    DevBuf<double2> lam, X1, Z, Z1;
    A.alloc((size_t)n * n * chunk);
    c->work_piv.alloc((size_t)2 * n * chunk + sizeof(SwapSet) / sizeof(int) * chunk);
    int *piv = c->work_piv.p, *perm = c->work_piv.p + (size_t)n * chunk;
    SwapSet *swaps = reinterpret_cast<SwapSet *>(c->work_piv.p + (size_t)2 * n * chunk);
    info.alloc(chunk);
    const bool cc = bem && c->mesh.charge_constraint;
    Z.alloc((size_t)3 * n * chunk);
    if (cc) {
        X1.alloc((size_t)n * chunk);
        Z1.alloc((size_t)n * chunk);
    }
    int status = 0;
    for (int f0 = 0; f0 < n_freq; f0 += chunk) {
        const int nf = std::min(chunk, n_freq - f0);
        assemble(c, f0, nf, freq, A.p, Mp, T, par, lam);
        PXH_HIP(hipMemsetAsync(info.p, 0, sizeof(int) * nf, st));
        lu_factor(c, n, nf, A.p, piv, swaps, info.p);
        double2 *X = d_sol + (size_t)rhs_stride * f0;
        // U^T (L^T z) = b, then x = P^T z
        k_perm_from_piv<<<nf, 256, 0, st>>>(n, piv, perm);
        k_copy_rhs<<<dim3((n + 255) / 256, 3, nf), 256, 0, st>>>(
            n, pol_mask, d_rhs_all + (size_t)rhs_fstride * f0, (size_t)rhs_fstride, Z.p, (size_t)3 * n);
        k_lu_solve<<<dim3(3, 1, nf), kSolveThreads, 0, st>>>(n, A.p, (size_t)n * n, Z.p, (size_t)3 * n, n, pol_mask);
        k_scatter_rows<<<dim3((n + 255) / 256, 3, nf), 256, 0, st>>>(n, perm, Z.p, (size_t)3 * n, X,
                                                                     (size_t)rhs_stride);
        if (cc) {
            k_fill_one<<<dim3((n + 255) / 256, 1, nf), 256, 0, st>>>(n, Z1.p, (size_t)n);
            k_lu_solve<<<dim3(1, 1, nf), kSolveThreads, 0, st>>>(n, A.p, (size_t)n * n, Z1.p, (size_t)n, n, 1);
            k_scatter_rows<<<dim3((n + 255) / 256, 1, nf), 256, 0, st>>>(n, perm, Z1.p, (size_t)n, X1.p, (size_t)n);
            k_charge_constraint<<<dim3(3, 1, nf), 256, 0, st>>>(n, pol_mask, X, (size_t)rhs_stride, X1.p, (size_t)n);
        }
        PXH_HIP(hipGetLastError());
        std::vector<int> hi(nf);
        PXH_HIP(hipMemcpyAsync(hi.data(), info.p, sizeof(int) * nf, hipMemcpyDeviceToHost, st));
        PXH_HIP(hipStreamSynchronize(st));
        for (int f = 0; f < nf; ++f)
            if (hi[f] != 0) status = PXH_ESINGULAR;
    }
    if (A.n * sizeof(double2) > kKeepLuBytes) {
        A.release();
        c->work_piv.release();
    }
    return status;
}

namespace {

// ---------------------------------------------------------------- groups
// Many targets solved together: the dense matrices of every (target,
// frequency) slot of a group of targets of similar size are padded to the
// group's largest n with an identity block, diag(A_g(w), I), and factorised
// together by the same LU launches (block-diagonal, so partial pivoting never
// leaves A_g and the padded unknowns come out zero).
struct GroupSlot {
    const double *M;       // frequency-independent matrix of the slot's target, nv x nv
    const double4 *pos;    // its atoms
    const double *nbfrac;  // its neighbour fractions (nt x n_at)
    const double2 *rhs;    // its right-hand sides, nv x 3
    int n_at, nv, nt, fmu;
    SysParam par;          // the slot's frequency
};

// A[s] = diag(A_g(w_s), I), N x N column-major; A_g as k_assemble_atoms.
__global__ void k_assemble_group(int N, const GroupSlot *__restrict__ slots, double2 *__restrict__ A) {
    const int s = blockIdx.z;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    if (r >= N) return;
    const GroupSlot &g = slots[s];
    double2 v = make_double2(r == col ? 1.0 : 0.0, 0.0);
    if (r < g.nv && col < g.nv) {
        v = make_double2(g.M[r + (size_t)g.nv * col], 0.0);
        if (r == col) {
            if (r < g.n_at) {
                const double2 z = g.par.zq[(int)g.pos[r].w];
                v.x -= z.x;
                v.y -= z.y;
            } else if (g.fmu) {
                const int i = (r - g.n_at) / 3;
                double2 lam = make_double2(0, 0);
                for (int t = 0; t < g.nt; ++t) {
                    lam.x += g.nbfrac[t + (size_t)g.nt * i] * g.par.inv_alpha[t].x;
                    lam.y += g.nbfrac[t + (size_t)g.nt * i] * g.par.inv_alpha[t].y;
                }
                v = lam;
            }
        }
    }
    A[(size_t)N * N * s + r + (size_t)N * col] = v;
}

// Z[s][k] = the slot target's right-hand side k (zero-padded to N), for k in mask.
__global__ void k_group_rhs(int N, int mask, const GroupSlot *__restrict__ slots, double2 *__restrict__ Z) {
    const int s = blockIdx.z, k = blockIdx.y;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const GroupSlot &g = slots[s];
    Z[(size_t)3 * N * s + (size_t)N * k + r] =
        ((mask >> k) & 1) && r < g.nv ? g.rhs[(size_t)g.nv * k + r] : make_double2(0.0, 0.0);
}

// out[s][a + 3 b] = sum_i x_b(q_i) r_a(i) (+ mu_a(i) for omegaFQFmu): the
// polarizability sums of wfq_module.f90:537-559 / wfqfmu_module.f90:894-938
// for slot s and field direction b = blockIdx.y.
__global__ __launch_bounds__(256) void k_group_polar(int N, const GroupSlot *__restrict__ slots,
                                                     const double2 *__restrict__ X, double2 *__restrict__ out) {
    const int s = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
    const GroupSlot &g = slots[s];
    const double2 *x = X + (size_t)3 * N * s + (size_t)N * b;
    double2 acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
    for (int i = tid; i < g.n_at; i += 256) {
        const double4 p = g.pos[i];
        const double2 q = x[i];
        const double rr[3] = {p.x, p.y, p.z};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            acc[a].x += q.x * rr[a];
            acc[a].y += q.y * rr[a];
            if (g.fmu) {
                const double2 mu = x[g.n_at + 3 * i + a];
                acc[a].x += mu.x;
                acc[a].y += mu.y;
            }
        }
    }
    __shared__ double2 sh[3][256];
    for (int a = 0; a < 3; ++a) sh[a][tid] = acc[a];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off)
            for (int a = 0; a < 3; ++a) {
                sh[a][tid].x += sh[a][tid + off].x;
                sh[a][tid].y += sh[a][tid + off].y;
            }
        __syncthreads();
    }
    if (tid < 3) out[(size_t)9 * s + tid + 3 * b] = sh[tid][0];
}

// Static FQ / FQFmu matrix (fq_module.f90:179-207, fqfmu_module.f90:81-140),
// unknowns [q_1..q_N, lambda_1..lambda_M, mu_1x, mu_1y, mu_1z, ...]: T (the
// dense atomistic T, unknowns [q, mu], n_t = N or 4N) in the q and mu rows
// and columns, with 1/alpha added on the dipole diagonal (matrix_module.f90:
// 280-288); the Lagrangian columns and rows of ones of each molecule
// (add_lagrangian_blocks, matrix_module.f90:351-382); zero elsewhere.  One
// work-item per element, columns over blockIdx.y.
struct InvAlpha {
    double v[kMaxTypes];
};

__global__ void k_assemble_static(int ns, int n, int m, int n_t, const double *__restrict__ T,
                                  const int *__restrict__ mol, const double4 *__restrict__ pos, InvAlpha ia,
                                  double2 *__restrict__ A) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    if (r >= ns) return;
    // index into T (q or mu unknown), or -1 - k for the multiplier of molecule k
    auto map = [&](int u) { return u < n ? u : (u < n + m ? -1 - (u - n) : u - m); };
    const int tr = map(r), tc = map(col);
    double v = 0.0;
    if (tr >= 0 && tc >= 0) {
        v = T[tr + (size_t)n_t * tc];
        if (tr == tc && tr >= n) v += ia.v[(int)pos[(tr - n) / 3].w];
    } else if (tr >= 0 && tr < n && tc < 0) {
        v = mol[tr] == -1 - tc ? 1.0 : 0.0;
    } else if (tc >= 0 && tc < n && tr < 0) {
        v = mol[tc] == -1 - tr ? 1.0 : 0.0;
    }
    A[r + (size_t)ns * col] = make_double2(v, 0.0);
}

}  // namespace

// Dense inversions of many atomistic targets together (the frequency loop
// of solve_w_inversion, algorithm_module.f90:189-265, for every target):
// targets sorted by size and grouped while the largest n of a group stays
// within kGroupPad of its smallest; each group's (target, frequency) slots are
// padded to that n and factorised together in chunks of at most
// kDenseChunkBytes (and half the free HBM), so one launch of each LU kernel
// covers many small targets and frequencies.
constexpr double kGroupPad = 1.12;
constexpr int kMaxGroupSlots = 8192;

int solve_group(pxh_ctx *const *ctxs, int n_ctx, int n_freq, const double *freq, double e0, int rhs_form,
                int pol_mask, double *polar, double *results) {
    pxh_ctx *c0 = ctxs[0];
    hipStream_t st = c0->stream;
    for (double &v : c0->stats.v) v = 0.0;  // the counters of this call (pxh_get_stats on ctxs[0])
    // per target: the constant matrix (kept with the target) and the
    // right-hand sides, built on its own stream, then joined
    std::vector<DevBuf<double2>> rhs(n_ctx);
    std::vector<char> by_spectral(n_ctx, 0);
    for (int t = 0; t < n_ctx; ++t) {
        pxh_ctx *c = ctxs[t];
        Atoms &A = c->atoms;
        const size_t nv = A.n_var();
        by_spectral[t] = spectral_applicable(c) && spectral_pays(c, (int)nv, n_freq);
        PXH_CHECK(by_spectral[t] || nv * nv * sizeof(double) <= kKeepMconstBytes, PXH_ENOMEM,
                  "a grouped target is larger than the constant matrix kept with it (2 GB)");
        if (!by_spectral[t] && !A.Mconst_valid) {
            A.Mconst.alloc(nv * nv);
            constant_matrix_atoms(c, A.Mconst.p);
            A.Mconst_valid = true;
        }
        rhs[t].alloc(nv * 3);
        dynamic_rhs_atoms(c, freq[0], e0, rhs_form, rhs[t].p);
        PXH_HIP(hipStreamSynchronize(c->stream));
    }
    // homogeneous omegaFQ targets for which one eigendecomposition is
    // estimated faster than their LUs: solved alone by the spectral route,
    // each on its own context (stream and rocBLAS handle), one after another,
    // with the device gate held exclusively
    std::vector<int> order, spec;
    for (int t = 0; t < n_ctx; ++t) (by_spectral[t] ? spec : order).push_back(t);
    std::vector<std::vector<double2>> hpol(n_ctx);
    std::unique_ptr<DeviceExclusive> alone;  // no other work of the process on the device meanwhile
    if (!spec.empty()) alone.reset(new DeviceExclusive);
    for (int t : spec) {
        pxh_ctx *c = ctxs[t];
        if (spectral_decompose(c, rhs[t].p, pol_mask, false)) {  // polarizabilities only
            DevBuf<double2> pol;
            pol.alloc((size_t)9 * n_freq);
            spectral_polar(c, n_freq, freq, pol.p);
            hpol[t].resize((size_t)9 * n_freq);
            PXH_HIP(hipMemcpyAsync(hpol[t].data(), pol.p, sizeof(double2) * hpol[t].size(), hipMemcpyDeviceToHost,
                                   c->stream));
            PXH_HIP(hipStreamSynchronize(c->stream));
        }
        c->spectral.release();
        timer_drain(c);
    }
    alone.reset();
    for (int t : spec) {
        pxh_ctx *c = ctxs[t];
        if (c != c0)
            for (int k = 24; k < 28; ++k) {
                c0->stats.v[k] += c->stats.v[k];
                c->stats.v[k] = 0.0;
            }
        if (hpol[t].empty()) {  // T not positive definite: the LU groups take it
            Atoms &A = c->atoms;
            const size_t nv = A.n_var();
            PXH_CHECK(nv * nv * sizeof(double) <= kKeepMconstBytes, PXH_ENOMEM,
                      "a grouped target is larger than the constant matrix kept with it (2 GB)");
            if (!A.Mconst_valid) {
                A.Mconst.alloc(nv * nv);
                constant_matrix_atoms(c, A.Mconst.p);
                A.Mconst_valid = true;
                PXH_HIP(hipStreamSynchronize(c->stream));
            }
            order.push_back(t);
            continue;
        }
        const std::vector<double2> &hp = hpol[t];
        for (int w = 0; w < n_freq; ++w) {
            double pl[18];
            for (int e = 0; e < 9; ++e) {
                pl[2 * e] = hp[(size_t)9 * w + e].x / e0;
                pl[2 * e + 1] = hp[(size_t)9 * w + e].y / e0;
            }
            if (polar) std::memcpy(polar + 18 * ((size_t)t * n_freq + w), pl, sizeof pl);
            if (results) {
                double r[20];
                results_row(freq[w], pl, r);
                double *res_t = results + (size_t)20 * n_freq * t;  // n_freq x 20, column-major
                for (int e = 0; e < 20; ++e) res_t[w + (size_t)n_freq * e] = r[e];
            }
        }
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return ctxs[a]->atoms.n_var() < ctxs[b]->atoms.n_var(); });
    DevBuf<double2> A, Z, X, pol;
    DevBuf<int> piv, info;
    DevBuf<SwapSet> swaps;
    DevBuf<GroupSlot> dslots;
    int status = PXH_OK;
    size_t free_b = 0, total_b = 0;
    PXH_HIP(hipMemGetInfo(&free_b, &total_b));
    for (size_t g0 = 0; g0 < order.size();) {
        // the group: targets from g0 while n stays within kGroupPad of the first
        const int n_first = ctxs[order[g0]]->atoms.n_var();
        size_t g1 = g0 + 1;
        while (g1 < order.size() && ctxs[order[g1]]->atoms.n_var() <= kGroupPad * n_first) ++g1;
        const int N = ctxs[order[g1 - 1]]->atoms.n_var();
        const size_t per = (size_t)N * N * sizeof(double2);
        size_t cap = std::max<size_t>(1, std::min(kDenseChunkBytes, free_b / 2) / per);
        if (c0->budget) cap = std::max<size_t>(1, std::min(cap, c0->budget / per));
        cap = std::min<size_t>(cap, kMaxGroupSlots);
        // slots target-major: (order[g0], w0), (order[g0], w1), ...
        const size_t n_slots = (g1 - g0) * (size_t)n_freq;
        for (size_t s0 = 0; s0 < n_slots; s0 += cap) {
            const int cnt = (int)std::min(cap, n_slots - s0);
            std::vector<GroupSlot> hs(cnt);
            for (int k = 0; k < cnt; ++k) {
                const size_t sl = s0 + k;
                pxh_ctx *c = ctxs[order[g0 + sl / n_freq]];
                const int w = (int)(sl % n_freq);
                Atoms &At = c->atoms;
                hs[k] = GroupSlot{At.Mconst.p, At.pos.p, At.d_nbfrac.p, rhs[order[g0 + sl / n_freq]].p, At.n,
                                  At.n_var(), At.nt, At.model == PXH_WFQFMU ? 1 : 0, make_sysparam(c, freq[w], w)};
            }
            dslots.upload(hs.data(), hs.size(), st);
            A.alloc((size_t)N * N * cnt);
            Z.alloc((size_t)3 * N * cnt);
            X.alloc((size_t)3 * N * cnt);
            piv.alloc((size_t)2 * N * cnt);
            swaps.alloc(cnt);
            info.alloc(cnt);
            pol.alloc((size_t)9 * cnt);
            k_assemble_group<<<dim3((N + 255) / 256, N, cnt), 256, 0, st>>>(N, dslots.p, A.p);
            PXH_HIP(hipMemsetAsync(info.p, 0, sizeof(int) * cnt, st));
            lu_factor(c0, N, cnt, A.p, piv.p, swaps.p, info.p);
            int *perm = piv.p + (size_t)N * cnt;
            k_perm_from_piv<<<cnt, 256, 0, st>>>(N, piv.p, perm);
            k_group_rhs<<<dim3((N + 255) / 256, 3, cnt), 256, 0, st>>>(N, pol_mask, dslots.p, Z.p);
            k_lu_solve<<<dim3(3, 1, cnt), kSolveThreads, 0, st>>>(N, A.p, (size_t)N * N, Z.p, (size_t)3 * N, N,
                                                                   pol_mask);
            k_scatter_rows<<<dim3((N + 255) / 256, 3, cnt), 256, 0, st>>>(N, perm, Z.p, (size_t)3 * N, X.p,
                                                                          (size_t)3 * N);
            k_group_polar<<<dim3(cnt, 3), 256, 0, st>>>(N, dslots.p, X.p, pol.p);
            PXH_HIP(hipGetLastError());
            std::vector<double2> hp((size_t)9 * cnt);
            std::vector<int> hi(cnt);
            PXH_HIP(hipMemcpyAsync(hp.data(), pol.p, sizeof(double2) * hp.size(), hipMemcpyDeviceToHost, st));
            PXH_HIP(hipMemcpyAsync(hi.data(), info.p, sizeof(int) * cnt, hipMemcpyDeviceToHost, st));
            PXH_HIP(hipStreamSynchronize(st));
            for (int k = 0; k < cnt; ++k) {
                const size_t sl = s0 + k;
                const int t = order[g0 + sl / n_freq], w = (int)(sl % n_freq);
                if (hi[k] != 0) status = PXH_ESINGULAR;
                double pl[18];
                for (int e = 0; e < 9; ++e) {
                    pl[2 * e] = hp[(size_t)9 * k + e].x / e0;
                    pl[2 * e + 1] = hp[(size_t)9 * k + e].y / e0;
                }
                if (polar) std::memcpy(polar + 18 * ((size_t)t * n_freq + w), pl, sizeof pl);
                if (results) {
                    double r[20];
                    results_row(freq[w], pl, r);
                    double *res_t = results + (size_t)20 * n_freq * t;  // n_freq x 20, column-major
                    for (int e = 0; e < 20; ++e) res_t[w + (size_t)n_freq * e] = r[e];
                }
            }
        }
        g0 = g1;
    }
    return status;
}

// Static FQ / FQFmu solve (algorithm_module.f90:469-504, DSYSV there; here the
// batched complex LU of the inversion path on the real symmetric system).
// what 0: ground state, one right-hand side -chi (fq_module.f90:211-240),
// energy -1/2 x^T b (:364-381); what 1: static field, three right-hand sides
// (rhs_module.f90:43-127), and for FQ the polarizability from three more
// solves at unit field: alpha_ab = -sum_i x_a(i) q_i^(b), the closed form of
// the constrained solve that fq_module.f90:385-597 evaluates through J^-1.
int solve_static(pxh_ctx *c, int what, double e0, int rhs_form, double *variables, double *polar, double *energy) {
    hipStream_t st = c->stream;
    Atoms &A = c->atoms;
    const int n = A.n, m = (int)A.mol_count.size(), n_t = A.n_var();
    const bool fmu = A.model == PXH_WFQFMU;
    const int ns = n + m + (fmu ? 3 * n : 0);
    const int ncol = what == 0 ? 1 : 3;
    const bool want_polar = what == 1 && !fmu;
    const int nz = want_polar ? 6 : ncol;  // right-hand sides solved
    // host right-hand sides, column-major ns x nz
    std::vector<double> b((size_t)ns * nz, 0.0);
    if (what == 0) {
        for (int i = 0; i < n; ++i) b[i] = -A.chi[A.h_type[i]];
    } else {
        double mn[3] = {0, 0, 0};
        if (rhs_form == PXH_RHS_POTENTIAL)
            for (int k = 0; k < 3; ++k) {
                mn[k] = A.h_coord[k];
                for (int i = 0; i < n; ++i) mn[k] = std::min(mn[k], A.h_coord[3 * i + k]);
            }
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < n; ++i) {
                const double x = A.h_coord[3 * i + k];
                // rhs_module.f90:43-90
                b[i + (size_t)ns * k] = rhs_form == PXH_RHS_POTENTIAL ? -e0 * std::fabs(x - mn[k]) : -e0 * x;
                if (want_polar) b[i + (size_t)ns * (3 + k)] = -x;
            }
        // dipole rows: 1 on component k when e0 is not zero, else 0
        // (rhs_module.f90:96-127, as written there)
        if (fmu)
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < 3; ++k) b[(n + m + 3 * i + k) + (size_t)ns * k] = e0 == 0.0 ? 0.0 : 1.0;
    }
    DevBuf<double> T;
    T.alloc((size_t)n_t * n_t);
    dense_t_atoms(c, T.p, false);
    std::vector<int> mol(n, 0);
    for (int k = 0, first = 0; k < m; first += A.mol_count[k], ++k)
        for (int i = first; i < first + A.mol_count[k]; ++i) mol[i] = k;
    DevBuf<int> dmol;
    dmol.upload(mol.data(), mol.size(), st);
    InvAlpha ia{};
    for (int t = 0; t < A.nt; ++t) ia.v[t] = fmu ? 1.0 / A.alpha_s[t] : 0.0;
    DevBuf<double2> M, Z, X;
    M.alloc((size_t)ns * ns);
    k_assemble_static<<<dim3((ns + 255) / 256, ns), 256, 0, st>>>(ns, n, m, n_t, T.p, dmol.p, A.pos.p, ia, M.p);
    PXH_HIP(hipGetLastError());
    DevBuf<int> piv, info;
    DevBuf<SwapSet> swaps;
    piv.alloc((size_t)2 * ns);
    swaps.alloc(1);
    info.alloc(1);
    PXH_HIP(hipMemsetAsync(info.p, 0, sizeof(int), st));
    lu_factor(c, ns, 1, M.p, piv.p, swaps.p, info.p);
    int *perm = piv.p + ns;
    k_perm_from_piv<<<1, 256, 0, st>>>(ns, piv.p, perm);
    // right-hand sides in groups of three (the solve kernel's width)
    const int ngroups = (nz + 2) / 3;
    std::vector<double2> hz((size_t)ns * 3 * ngroups, make_double2(0.0, 0.0));
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < ns; ++i) hz[i + (size_t)ns * k].x = b[i + (size_t)ns * k];
    Z.upload(hz.data(), hz.size(), st);
    X.alloc(hz.size());
    for (int g = 0; g < ngroups; ++g) {
        const int mask = (1 << std::min(3, nz - 3 * g)) - 1;
        double2 *zg = Z.p + (size_t)3 * ns * g, *xg = X.p + (size_t)3 * ns * g;
        k_lu_solve<<<dim3(3, 1, 1), kSolveThreads, 0, st>>>(ns, M.p, (size_t)ns * ns, zg, (size_t)3 * ns, ns, mask);
        k_scatter_rows<<<dim3((ns + 255) / 256, 3, 1), 256, 0, st>>>(ns, perm, zg, (size_t)3 * ns, xg,
                                                                     (size_t)3 * ns);
    }
    PXH_HIP(hipGetLastError());
    std::vector<double2> hx(hz.size());
    int hinfo = 0;
    PXH_HIP(hipMemcpyAsync(hx.data(), X.p, sizeof(double2) * hx.size(), hipMemcpyDeviceToHost, st));
    PXH_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
    PXH_HIP(hipStreamSynchronize(st));
    if (hinfo != 0) return PXH_ESINGULAR;
    if (variables)
        for (int k = 0; k < ncol; ++k)
            for (int i = 0; i < ns; ++i) variables[i + (size_t)ns * k] = hx[i + (size_t)ns * k].x;
    if (energy) {
        double e = 0.0;
        for (int i = 0; i < ns; ++i) e += hx[i].x * b[i];
        *energy = -0.5 * e;
    }
    if (polar && want_polar)
        for (int a = 0; a < 3; ++a)
            for (int bb = 0; bb < 3; ++bb) {
                double sum = 0.0;
                for (int i = 0; i < n; ++i) sum += A.h_coord[3 * i + a] * hx[i + (size_t)ns * (3 + bb)].x;
                polar[a + 3 * bb] = -sum;
            }
    return PXH_OK;
}

void release_blas(pxh_ctx *c) {
    if (c->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(c->blas));
    c->blas = nullptr;
}

}  // namespace pxh
