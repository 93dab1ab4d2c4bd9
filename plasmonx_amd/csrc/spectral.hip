// Spectral route for the dense omegaFQ inversion of a homogeneous target.
//
// There A(w) = K T - z(w) I for every frequency (wfq_module.f90:214-254: the
// constant matrix K-bar T_qq of matrix_module.f90:491-524 with the Drude
// shift z_q(w) of :239-242 on the diagonal, the same for every atom), and the
// right-hand sides K (E0 r_k) do not depend on w (rhs_module.f90:140-186).
// K-bar is symmetric (one type: A_ij = A_t) and T_qq symmetric positive
// definite (Coulomb energy matrix of Gaussian charges), so K T is similar to
// the symmetric L^T K L (T = L L^T): its eigenvalues are real and
//     T K T x = lambda T x        (generalized symmetric-definite problem)
// gives K T = X diag(lambda) X^-1 with X^T T X = I, i.e. X^-1 = X^T T, and
//     x_k(w) = X diag(1 / (lambda - z(w))) X^T T b_k
// for every frequency at O(n^2) each (O(n) for the polarizability), after one
// eigendecomposition in place of one complex LU (8/3 n^3 flop) per frequency.
// The reference factorises A(w) per frequency (ZGETRF / ZGETRS,
// algorithm_module.f90:510-541); the solutions are the same to rounding.
//
// Work on the device: T and M = K T by the library's kernels, G = T M and the
// small products by rocBLAS GEMMs, the eigendecomposition by rocSOLVER DSYGVD
// (LAPACK semantics: Cholesky of T, reduction, divide-and-conquer
// tridiagonal eigensolver, back-transformation), the per-frequency
// coefficients, polarizability sums and the scatter into the solution layout
// by the kernels below.  A T that is not positive definite (other kernels or
// parameters) makes DSYGVD report it, and the caller falls back to the LU.
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
// (DSYGVD times measured on MI355X: 0.33 s at n = 4000, 1.08 s at 8000,
// 2.61 s at 12000, ~n^1.9, and at least ~0.12 s below that).
bool spectral_pays(const pxh_ctx *c, int n, int n_freq) {
    if (c->opt.spectral == 0) return false;
    if (c->opt.spectral == 2) return true;
    const double nn = (double)n;
    const double t_lu = n_freq * (8.0 / 3.0) * nn * nn * nn / 23e12;
    const double t_eig = std::max(0.12, 0.33 * std::pow(nn / 4000.0, 1.9)) + 2.0 * nn * nn * nn / 40e12;
    return t_eig < 0.8 * t_lu;
}

size_t spectral_bytes(int n) { return ((size_t)3 * n * n + (size_t)24 * n) * sizeof(double); }

// The decomposition for the right-hand sides rhs (n x 3, the selected
// directions of mask); on success the context's spectral workspace holds
// lambda (n), X (n x n), W = X^T r (n x 3) and C = X^T T b (n x 6: Re, Im of
// each direction).  Returns false when T is not positive definite or the
// eigensolver does not converge (the caller then uses the LU).
bool spectral_decompose(pxh_ctx *c, const double2 *rhs, int mask) {
    Atoms &At = c->atoms;
    const int n = At.n;
    hipStream_t st = c->stream;
    SpectralWs &S = c->spectral;
    S.n = 0;
    const double *M = nullptr;
    DevBuf<double> Mtmp;
    if (!At.Mconst_valid) {
        const bool keep = (size_t)n * n * sizeof(double) <= kKeepMconstBytes;
        DevBuf<double> &dst = keep ? At.Mconst : Mtmp;
        dst.alloc((size_t)n * n);
        constant_matrix_atoms(c, dst.p);
        At.Mconst_valid = keep;
        M = dst.p;
    } else {
        M = At.Mconst.p;
    }
    DevBuf<double> T, B6, U, R3, E;
    DevBuf<int> info;
    T.alloc((size_t)n * n);
    dense_t_atoms(c, T.p, false);
    S.X.alloc((size_t)n * n);
    S.lam.alloc(n);
    S.c6.alloc((size_t)6 * n);
    S.w3.alloc((size_t)3 * n);
    B6.alloc((size_t)6 * n);
    U.alloc((size_t)6 * n);
    R3.alloc((size_t)3 * n);
    E.alloc(n);
    info.alloc(1);
    rocblas_handle h = blas_handle(c);
    const int tm = timer_begin(c);
    // G = T (K T) = T K T, symmetric (DSYGVD reads its upper triangle)
    dgemm(h, false, n, n, n, T.p, n, M, n, S.X.p, n);
    // U = T b, before T is overwritten by its Cholesky factor
    k_rhs_split<<<dim3((n + 255) / 256, 3), 256, 0, st>>>(n, mask, rhs, B6.p);
    k_pos_cols<<<(n + 255) / 256, 256, 0, st>>>(n, At.pos.p, R3.p);
    PXH_HIP(hipGetLastError());
    dgemm(h, false, n, 6, n, T.p, n, B6.p, n, U.p, n);
    // T K T x = lambda T x: eigenvalues ascending into lam, X (X^T T X = I) in place of G
    check_blas(rocsolver_dsygvd(h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, n, S.X.p, n, T.p,
                                n, S.lam.p, E.p, info.p),
               "rocsolver_dsygvd failed");
    int hinfo = 0;
    PXH_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
    PXH_HIP(hipStreamSynchronize(st));
    if (hinfo != 0) {  // > n: T not positive definite; 1..n: no convergence
        timer_end(c, tm, 25);
        S.release();
        return false;
    }
    // C = X^T (T b), W = X^T r
    dgemm(h, true, n, 6, n, S.X.p, n, U.p, n, S.c6.p, n);
    dgemm(h, true, n, 3, n, S.X.p, n, R3.p, n, S.w3.p, n);
    timer_end(c, tm, 25);
    c->stats.v[24] += 1;
    c->stats.v[26] += 13.0 * (double)n * n * n;  // nominal flops: G (2 n^3), DSYGVD (~11 n^3 with vectors)
    S.n = n;
    S.mask = mask;
    return true;
}

// Solutions of every frequency into sol (n x 3 per frequency, frequency
// stride `stride`), in chunks of frequencies: Y = coefficients, Q = X Y.
void spectral_charges(pxh_ctx *c, int n_freq, const double *freq, double2 *sol, size_t stride) {
    SpectralWs &S = c->spectral;
    const int n = S.n;
    hipStream_t st = c->stream;
    rocblas_handle h = blas_handle(c);
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(n_freq, (size_t(1) << 30) / ((size_t)96 * n)));
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
        dgemm(h, false, n, 2 * ncol, n, S.X.p, n, Y.p, n, Q.p, n);
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
