// Spectral route for the dense omegaFQ inversion of a homogeneous target.
//
// There A(w) = K T - z(w) I for every frequency (wfq_module.f90:214-254: the
// constant matrix K-bar T_qq of matrix_module.f90:491-524 with the Drude
// shift z_q(w) of :239-242 on the diagonal, the same for every atom), and the
// right-hand sides K (E0 r_k) do not depend on w (rhs_module.f90:140-186).
// K-bar is symmetric (one type: A_ij = A_t) and T_qq symmetric positive
// definite (Coulomb energy matrix of Gaussian charges), so with the Cholesky
// factor T = L L^T, K T = L^-T (L^T K L) L^T is similar to the symmetric
//     S = L^T K L = Q diag(lambda) Q^T,
// its eigenvalues are real, and with X = L^-T Q (X^T T X = I, X^-1 = Q^T L^T)
//     x_k(w) = L^-T Q diag(1 / (lambda - z(w))) Q^T L^T b_k
// for every frequency at O(n^2) each (O(n) for the polarizability), after one
// eigendecomposition in place of one complex LU (8/3 n^3 flop) per frequency.
// The reference factorises A(w) per frequency (ZGETRF / ZGETRS,
// algorithm_module.f90:510-541); the solutions are the same to rounding.
//
// Work on the device: T by the library's kernel, its Cholesky factor by
// rocSOLVER DPOTRF, P = K L by the sparse kernel below (K-bar in CSR: a few
// neighbours per row), S = L^T P by rocBLAS DTRMM, the eigendecomposition of
// S by rocSOLVER DSYEVD (LAPACK semantics: tridiagonal reduction,
// divide-and-conquer, back-transformation), the small products by rocBLAS
// GEMM / TRMM / TRSM, and the per-frequency coefficients, polarizability
// sums and the scatter into the solution layout by the kernels below.  A T
// that is not positive definite (other kernels or parameters) makes DPOTRF
// report it, and the caller falls back to the LU.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pxh_internal.h"

namespace pxh {

rocblas_handle blas_handle(pxh_ctx *c);

namespace {

// B6[:, k] = Re b_k, B6[:, 3 + k] = Im b_k for the selected directions (zero otherwise).
__global__ void k_rhs_split(int n, int mask, const double2 *__restrict__ rhs, double *__restrict__ b6) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (i >= n) return;
    const double2 v = ((mask >> k) & 1) ? rhs[(size_t)n * k + i] : make_double2(0.0, 0.0);
    b6[(size_t)n * k + i] = v.x;
    b6[(size_t)n * (3 + k) + i] = v.y;
}

// P = K L (n x n, column-major) for the lower-triangular L (entries above the
// diagonal of its storage are not part of it): with K-bar in CSR (off-diagonal
// entries, K_ii = -sum_j K_ij), P_ij = sum_k K_ik (L_kj - L_ij), the
// difference form of the K-bar products (matrix_module.f90:1011-1098).
// Work-items over rows i, grid rows over columns j.
__global__ void k_kbar_lower(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                             const double *__restrict__ kij, const double *__restrict__ L, double *__restrict__ P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int z0 = rowptr[i], z1 = rowptr[i + 1];
    for (int j = blockIdx.y; j < n; j += gridDim.y) {
        const double *Lj = L + (size_t)n * j;
        const double lij = i >= j ? Lj[i] : 0.0;
        double acc = 0.0;
        for (int z = z0; z < z1; ++z) {
            const int k = col[z];
            acc += kij[z] * ((k >= j ? Lj[k] : 0.0) - lij);
        }
        P[i + (size_t)n * j] = acc;
    }
}

// R3[:, a] = coordinate a of the atoms.
__global__ void k_pos_cols(int n, const double4 *__restrict__ pos, double *__restrict__ r3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = pos[i];
    r3[i] = p.x;
    r3[(size_t)n + i] = p.y;
    r3[2 * (size_t)n + i] = p.z;
}

__device__ __forceinline__ double2 spectral_weight(double lam, double2 z) {
    // 1 / (lambda - z)
    const double re = lam - z.x, im = -z.y;
    const double d = re * re + im * im;
    return make_double2(re / d, -im / d);
}

// Coefficients of the chunk's solutions in the eigenvector basis, split in
// real and imaginary columns: column j = 3 (f - f0) + k of Y holds
// c_k .* (1 / (lambda - z(w_f))), Re in Y[:, j], Im in Y[:, ncol + j].
__global__ void k_spectral_coef(int n, int ncol, int mask, const double *__restrict__ lam,
                                const double2 *__restrict__ z, const double *__restrict__ c6, double *__restrict__ y) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (m >= n) return;
    const int f = j / 3, k = j % 3;
    double2 v = make_double2(0.0, 0.0);
    if ((mask >> k) & 1) {
        const double2 w = spectral_weight(lam[m], z[f]);
        const double cr = c6[(size_t)n * k + m], ci = c6[(size_t)n * (3 + k) + m];
        v = make_double2(cr * w.x - ci * w.y, cr * w.y + ci * w.x);
    }
    y[(size_t)n * j + m] = v.x;
    y[(size_t)n * (ncol + j) + m] = v.y;
}

// x_k(w_f)[i] = Q[i, j] + i Q[i, ncol + j] into the solution layout
// (sol + stride f + n k).
__global__ void k_spectral_scatter(int n, int ncol, const double *__restrict__ q, double2 *__restrict__ sol,
                                   size_t stride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const int f = j / 3, k = j % 3;
    sol[stride * f + (size_t)n * k + i] = make_double2(q[(size_t)n * j + i], q[(size_t)n * (ncol + j) + i]);
}

// out[f][a + 3 b] = sum_i q_b(i) r_a(i) = sum_m W[m, a] c_b(m) / (lambda_m - z(w_f))
// (the polarizability sums of wfq_module.f90:537-559), one workgroup per frequency.
__global__ __launch_bounds__(256) void k_spectral_polar(int n, int mask, const double *__restrict__ lam,
                                                        const double2 *__restrict__ z, const double *__restrict__ w3,
                                                        const double *__restrict__ c6, double2 *__restrict__ out) {
    const int f = blockIdx.x, tid = threadIdx.x;
    const double2 zf = z[f];
    double2 acc[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[e] = make_double2(0.0, 0.0);
    for (int m = tid; m < n; m += 256) {
        const double2 wt = spectral_weight(lam[m], zf);
        const double r[3] = {w3[m], w3[(size_t)n + m], w3[2 * (size_t)n + m]};
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            if (!((mask >> b) & 1)) continue;
            const double cr = c6[(size_t)n * b + m], ci = c6[(size_t)n * (3 + b) + m];
            const double2 v = make_double2(cr * wt.x - ci * wt.y, cr * wt.y + ci * wt.x);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                acc[a + 3 * b].x += r[a] * v.x;
                acc[a + 3 * b].y += r[a] * v.y;
            }
        }
    }
    __shared__ double2 sh[9][256];
#pragma unroll
    for (int e = 0; e < 9; ++e) sh[e][tid] = acc[e];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off)
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                sh[e][tid].x += sh[e][tid + off].x;
                sh[e][tid].y += sh[e][tid + off].y;
            }
        __syncthreads();
    }
    if (tid < 9) out[(size_t)9 * f + tid] = sh[tid][0];
}

void check_blas(rocblas_status s, const char *what) {
    PXH_CHECK(s == rocblas_status_success, PXH_EHIP, what);
}

// B = op(L) B (trmm) or B = op(L)^-1 B (trsm) for the lower-triangular L, from the left.
void tri_left(rocblas_handle h, bool solve, bool trans, int m, int nr, const double *L, double *B) {
    const double one = 1.0;
    const rocblas_operation op = trans ? rocblas_operation_transpose : rocblas_operation_none;
    if (solve)
        check_blas(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, op, rocblas_diagonal_non_unit, m, nr, &one,
                                 L, m, B, m),
                   "rocblas_dtrsm failed");
    else
        check_blas(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_lower, op, rocblas_diagonal_non_unit, m, nr, &one,
                                 L, m, B, m, B, m),
                   "rocblas_dtrmm failed");
}

// C = op(A) B for column-major real matrices (no accumulation).
void dgemm(rocblas_handle h, bool trans_a, int m, int n, int k, const double *A, int lda, const double *B, int ldb,
           double *C, int ldc) {
    const double one = 1.0, zero = 0.0;
    check_blas(rocblas_dgemm(h, trans_a ? rocblas_operation_transpose : rocblas_operation_none,
                             rocblas_operation_none, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc),
               "rocblas_dgemm failed");
}

}  // namespace

// One atom type (K-bar symmetric, one z for every atom), omegaFQ.
bool spectral_applicable(const pxh_ctx *c) {
    if (c->target == PXH_BEM) return false;
    const Atoms &A = c->atoms;
    return A.model == PXH_WFQ && !A.hetero && !A.is_static && A.nt == 1 && A.n >= 2;
}

// Estimated seconds: the LU route (8/3 n^3 flop per frequency at the
// whole-path rate measured on the batched LU) against one decomposition
// (measured on MI355X, DSYEVD with the Cholesky and the products around it:
// ~0.15 s at n = 4000, 0.7 s at 8000, 1.9 s at 12000, ~n^2.3, and at least
// ~0.06 s below that, where the eigensolver's per-column launches dominate).
bool spectral_pays(const pxh_ctx *c, int n, int n_freq) {
    if (c->opt.spectral == 0) return false;
    if (c->opt.spectral == 2) return true;
    const double nn = (double)n;
    const double t_lu = n_freq * (8.0 / 3.0) * nn * nn * nn / 23e12;
    const double t_eig = std::max(0.06, 0.15 * std::pow(nn / 4000.0, 2.3));
    return t_eig < 0.8 * t_lu;
}

// The decomposition for the right-hand sides rhs (n x 3, the selected
// directions of mask); on success the context's spectral workspace holds
// lambda (n), L (n x n, lower), W = Q^T L^-1 r (n x 3),
// C = Q^T L^T b (n x 6: Re, Im of each direction) and, with need_q, Q
// (n x n; DSYEVD).  Without need_q (polarizabilities only) Q is never
// formed: S = H T H^T by DSYTRD, T = Z Lambda Z^T by DSTEDC, and
// Q^T v = Z^T (H^T v) for the 9 vectors (DORMTR on n x 9 instead of the
// back-transformation of all n eigenvectors).  Returns false when T is not
// positive definite or the eigensolver does not converge (the caller then
// uses the LU).
// The caller holds the device gate exclusively (DeviceExclusive) from here
// to the last use of the decomposition.
bool spectral_decompose(pxh_ctx *c, const double2 *rhs, int mask, bool need_q) {
    Atoms &At = c->atoms;
    const int n = At.n;
    hipStream_t st = c->stream;
    SpectralWs &S = c->spectral;
    S.n = 0;
    S.have_q = false;
    DevBuf<double> P, B6, R3, E;
    DevBuf<int> info;
    S.L.alloc((size_t)n * n);
    S.Q.alloc((size_t)n * n);
    S.lam.alloc(n);
    S.c6.alloc((size_t)6 * n);
    S.w3.alloc((size_t)3 * n);
    P.alloc((size_t)n * n);
    B6.alloc((size_t)6 * n);
    R3.alloc((size_t)3 * n);
    E.alloc(n);
    info.alloc(2);
    rocblas_handle h = blas_handle(c);
    const int tm = timer_begin(c);
    dense_t_atoms(c, S.L.p, false);
    k_rhs_split<<<dim3((n + 255) / 256, 3), 256, 0, st>>>(n, mask, rhs, B6.p);
    k_pos_cols<<<(n + 255) / 256, 256, 0, st>>>(n, At.pos.p, R3.p);
    PXH_HIP(hipGetLastError());
    check_blas(rocsolver_dpotrf(h, rocblas_fill_lower, n, S.L.p, n, info.p), "rocsolver_dpotrf failed");
    int hinfo[2] = {0, 0};
    PXH_HIP(hipMemcpyAsync(hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
    PXH_HIP(hipStreamSynchronize(st));
    if (hinfo[0] != 0) {  // T not positive definite
        timer_end(c, tm, 25);
        PXH_HIP(hipStreamSynchronize(st));
        S.release();
        return false;
    }
    // S = L^T (K L), symmetric; DSYEVD reads its lower triangle
    k_kbar_lower<<<dim3((n + 255) / 256, std::min(n, 65535)), 256, 0, st>>>(n, At.kb_rowptr.p, At.kb_col.p,
                                                                          At.kb_kij.p, S.L.p, P.p);
    PXH_HIP(hipGetLastError());
    const double one = 1.0;
    check_blas(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                             rocblas_diagonal_non_unit, n, n, &one, S.L.p, n, P.p, n, S.Q.p, n),
               "rocblas_dtrmm failed");
    DevBuf<double> tau;
    if (need_q) {
        check_blas(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, n, S.Q.p, n, S.lam.p, E.p,
                                    info.p + 1),
                   "rocsolver_dsyevd failed");
    } else {
        // reflectors of H into S.Q, the tridiagonal's eigenvectors Z into P
        tau.alloc(n);
        check_blas(rocsolver_dsytrd(h, rocblas_fill_lower, n, S.Q.p, n, S.lam.p, E.p, tau.p),
                   "rocsolver_dsytrd failed");
        check_blas(rocsolver_dstedc(h, rocblas_evect_tridiagonal, n, S.lam.p, E.p, P.p, n, info.p + 1),
                   "rocsolver_dstedc failed");
    }
    PXH_HIP(hipMemcpyAsync(hinfo + 1, info.p + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    PXH_HIP(hipStreamSynchronize(st));
    if (hinfo[1] != 0) {  // no convergence
        timer_end(c, tm, 25);
        PXH_HIP(hipStreamSynchronize(st));
        S.release();
        return false;
    }
    // C = Q^T (L^T b), W = Q^T (L^-1 r); without Q: Z^T (H^T v)
    tri_left(h, false, true, n, 6, S.L.p, B6.p);
    tri_left(h, true, false, n, 3, S.L.p, R3.p);
    if (need_q) {
        dgemm(h, true, n, 6, n, S.Q.p, n, B6.p, n, S.c6.p, n);
        dgemm(h, true, n, 3, n, S.Q.p, n, R3.p, n, S.w3.p, n);
    } else {
        check_blas(rocsolver_dormtr(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose, n, 6,
                                    S.Q.p, n, tau.p, B6.p, n),
                   "rocsolver_dormtr failed");
        check_blas(rocsolver_dormtr(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose, n, 3,
                                    S.Q.p, n, tau.p, R3.p, n),
                   "rocsolver_dormtr failed");
        dgemm(h, true, n, 6, n, P.p, n, B6.p, n, S.c6.p, n);
        dgemm(h, true, n, 3, n, P.p, n, R3.p, n, S.w3.p, n);
    }
    timer_end(c, tm, 25);
    PXH_HIP(hipStreamSynchronize(st));  // the temporaries are released on return
    c->stats.v[24] += 1;
    // nominal flops: Cholesky n^3/3, S n^3, tridiagonal reduction 4/3 n^3,
    // back-transformation of the eigenvectors 2 n^3 (with Q only)
    c->stats.v[26] += (1.0 / 3.0 + 1.0 + 4.0 / 3.0 + (need_q ? 2.0 : 0.0)) * (double)n * n * n;
    // algorithmic HBM bytes of the tridiagonal reduction: the lower triangle of
    // the trailing matrix read once per column, sum 8 (n - k)^2 / 2 = 4 n^3 / 3
    c->stats.v[27] += 4.0 / 3.0 * (double)n * n * n;
    S.n = n;
    S.mask = mask;
    S.have_q = need_q;
    return true;
}

// Solutions of every frequency into sol (n x 3 per frequency, frequency
// stride `stride`), in chunks of frequencies: Y = coefficients, then
// x = L^-T (Q Y).
void spectral_charges(pxh_ctx *c, int n_freq, const double *freq, double2 *sol, size_t stride) {
    SpectralWs &S = c->spectral;
    PXH_CHECK(S.have_q, PXH_EINVAL, "spectral charges need the eigenvectors (decomposition with need_q)");
    const int n = S.n;
    hipStream_t st = c->stream;
    rocblas_handle h = blas_handle(c);
    // frequencies per block: Y and Q within 1 GB each, and 3 per frequency
    // within the grid's y dimension
    const int chunk = (int)std::max<size_t>(
        1, std::min<size_t>(std::min<size_t>(n_freq, 16384), (size_t(1) << 30) / ((size_t)48 * n)));
    DevBuf<double> Y, Q;
    DevBuf<double2> z;
    std::vector<double2> hz(n_freq);
    for (int f = 0; f < n_freq; ++f) hz[f] = make_sysparam(c, freq[f], f).zq[0];
    z.upload(hz.data(), hz.size(), st);
    Y.alloc((size_t)6 * n * chunk);
    Q.alloc((size_t)6 * n * chunk);
    for (int f0 = 0; f0 < n_freq; f0 += chunk) {
        const int nf = std::min(chunk, n_freq - f0), ncol = 3 * nf;
        k_spectral_coef<<<dim3((n + 255) / 256, ncol), 256, 0, st>>>(n, ncol, S.mask, S.lam.p, z.p + f0, S.c6.p,
                                                                     Y.p);
        dgemm(h, false, n, 2 * ncol, n, S.Q.p, n, Y.p, n, Q.p, n);
        tri_left(h, true, true, n, 2 * ncol, S.L.p, Q.p);  // x = L^-T (Q y)
        k_spectral_scatter<<<dim3((n + 255) / 256, ncol), 256, 0, st>>>(n, ncol, Q.p, sol + stride * f0, stride);
        PXH_HIP(hipGetLastError());
    }
    PXH_HIP(hipStreamSynchronize(st));  // Y, Q and z are released on return
}

// Polarizability sums out[f][a + 3 b] (device, 9 per frequency).
void spectral_polar(pxh_ctx *c, int n_freq, const double *freq, double2 *out) {
    SpectralWs &S = c->spectral;
    hipStream_t st = c->stream;
    DevBuf<double2> z;
    std::vector<double2> hz(n_freq);
    for (int f = 0; f < n_freq; ++f) hz[f] = make_sysparam(c, freq[f], f).zq[0];
    z.upload(hz.data(), hz.size(), st);
    k_spectral_polar<<<n_freq, 256, 0, st>>>(S.n, S.mask, S.lam.p, z.p, S.w3.p, S.c6.p, out);
    PXH_HIP(hipGetLastError());
    PXH_HIP(hipStreamSynchronize(st));
}

}  // namespace pxh
