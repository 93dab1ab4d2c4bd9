// Internal declarations of libpxh (gfx950).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pxh.h"

namespace pxh {

// Error carried from deep helpers up to the C-ABI boundary.
struct Error {
    int code;
    std::string msg;
};

#define PXH_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::pxh::Error{PXH_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)};  \
    } while (0)

#define PXH_CHECK(cond, code, msg)                           \
    do {                                                     \
        if (!(cond)) throw ::pxh::Error{(code), (msg)};     \
    } while (0)

// Maximum number of atom types handled by the pair kernels.
constexpr int kMaxTypes = 8;
// Frequencies whose dense matrices are factorised together (inversion).
constexpr size_t kMaxFreqChunk = 128;                  // dense path: frequencies factorised together ...
constexpr size_t kDenseChunkBytes = size_t(32) << 30;  // ... within this many bytes of matrices
constexpr size_t kKeepLuBytes = size_t(4) << 30;       // their workspace kept with the context up to this
constexpr size_t kKeepMconstBytes = size_t(2) << 30;  // dense constant matrices kept with the target up to this

// Per type-pair constants of the interaction kernels (index ti + nt*tj).
struct PairConst {
    double inv_rqq;    // 1/sqrt(Rq_i^2 + Rq_j^2)           (T_qq, Gaussian)
    double inv_rqmu;   // 1/sqrt(Rq_i^2 + Rmu_j^2)          (T_qmu(i,j): charge on i, dipole on j)
    double inv_rmumu;  // 1/sqrt(Rmu_i^2 + Rmu_j^2)         (T_mumu)
    double eta_bar;    // (eta_i + eta_j)/2                  (Ohno)
};

// Device buffer with RAII.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count <= n && p) return;
        release();
        if (count == 0) return;
        PXH_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) PXH_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

// Atomistic system resident in HBM.
struct Atoms {
    int model = 0, kernel = 0, n = 0, nt = 0;  // model PXH_WFQ (charges) or PXH_WFQFMU (charges and dipoles)
    bool hetero = false;
    bool is_static = false;             // static FQ / FQFmu (no K-bar, no Drude parameters)
    std::vector<double> chi, alpha_s;   // static models: per type (pxh_set_static)
    std::vector<int> mol_count;         // static models: atoms per molecule, contiguous runs
    std::vector<double> h_coord;        // 3n, bohr (host copy)
    std::vector<int> h_type;            // n, 0-based
    std::vector<double> eta, rq, rmu, aij, tau, dens;
    std::vector<double> fd, fs, r0;     // nt*nt
    std::vector<double> h_nbfrac;       // nt*n: nb_counts/nb_total
    std::vector<PairConst> h_pair;      // nt*nt
    DevBuf<double4> pos;                // x, y, z, type (as double)
    DevBuf<int> tperm;                  // target order of the VALU pair kernels (Morton order of the atoms)
    DevBuf<float4> posf;                // centred single-precision copy
    DevBuf<PairConst> pair;
    DevBuf<double> d_eta;               // per type
    DevBuf<double> d_nbfrac;            // nt*n (row: type, column: atom)
    // K-bar in CSR: row i holds j != i with (1-f) above the cutoff.
    DevBuf<int> kb_rowptr, kb_col;
    DevBuf<double> kb_kij, kb_kji;      // K_ij = (1-f_ij) A_i / r, K_ji = (1-f_ji) A_j / r
    size_t kb_nnz = 0;
    // looser K-bar CSR (1 - f >= 2^-30) for the FP32-class products
    DevBuf<int> kbm_rowptr, kbm_col;
    DevBuf<double> kbm_kij, kbm_kji;
    size_t kbm_nnz = 0;
    bool kbm_built = false;
    DevBuf<double> Tstore;              // stored T for the in-memory method (n_var^2)
    DevBuf<double> Mconst;              // dense constant matrix of the inversion path, kept when small
    bool Mconst_valid = false;
    // diagonal of K-bar T_qq: S1_i = sum_j K_ij (T_qq(j,i) - eta_i), and for
    // alloys S2_i[t] = sum_{j of type t} K_ji (T_qq(j,i) - eta_i)
    DevBuf<double> diag_s1, diag_s2;
    int n_var() const { return model == PXH_WFQFMU ? 4 * n : n; }
};

struct Mesh {
    int n = 0, variant = 0, green = 0;
    bool charge_constraint = false;
    double normal_factor = 1.0, sphere_r = 0.0, eps_solvent = 1.0;
    std::vector<double> h_verts;        // 9n
    std::vector<double> h_coord, h_normal, h_t1, h_t2, h_area;
    DevBuf<double> D;                   // n*n column-major, D (refined)
    DevBuf<double> Dief;                // n*n, IEF: D (unrefined diagonal) - 2 pi I
    DevBuf<double> SA;                  // n*n, IEF: S A^-1
    DevBuf<float> Df, SAf;              // FP32 copies for the mixed path
    DevBuf<double4> ief_z;              // IEF: S^-1 c (x, y, z, 0) per tessera, once computed
    bool ief_z_valid = false;
    bool built = false;
};

struct Options {
    int kbar_cutoff = 1;
    int batch = 48;          // slots of the batched solver
    double inner_tol = 1e-5;  // relative tolerance of the FP32 inner GMRES
    int tile_i = 256;
    int splits = 0;          // 0 = automatic
    int kv = 4;              // systems per pair-kernel pass (FP32)
    int max_outer = 30;
    int precond = 1;         // diagonal right preconditioning in the batched path
    int mfma = 1;            // omegaFQFmu FP32-class products on MFMA (fp16 hi/lo split)
    int fp64_mfma = 1;       // omegaFQFmu FP64 products of >= 5 systems on the FP64 matrix cores
    int bem_gemv = 1;        // BEM FP64 products of <= 4 systems by the streaming pass (else the MFMA tiles)
    int mfma_levels = 2;     // FP32 accumulation levels of the MFMA pair kernel (1 or 2)
    int mfma_flush = 8;      // source steps (16 atoms each) summed in the MFMA registers between flushes
    int mfma_r4 = 1;         // one-column-block MFMA passes: four row blocks per wave (else two)
    int matvec_mixed = 0;    // pxh_matvec uses the mixed-precision product (testing)
    double reorth_kappa = 0.70710678118654752;  // batched CGS: second pass when ||w'|| < kappa ||w||
    int multishift = 1;      // scalar-shift frequency families: one Krylov basis for all frequencies
    int initial_guess = 0;   // iterative methods start from the caller's variables
    int dcgs2 = 1;           // batched solver: Gram-Schmidt with delayed re-orthogonalisation
    int res_batch = 16;      // batched solver: FP64 residual products wait for this many systems ...
    int res_wait = 5;        // ... for at most this many steps
    int autotune = 1;        // time the candidate launch shapes of a new size once per process
    int x0_projection = 8;   // batched solver: start from the best x in the span of this many converged neighbours
    int graphs = 1;          // multi-shift GMRES on BEM: Arnoldi steps replayed from a captured HIP graph
    int spectral = 1;        // homogeneous omegaFQ inversion by one eigendecomposition: 0 never, 1 when it
                             // is estimated faster than the LU, 2 always
};

// Spectral route of the homogeneous omegaFQ inversion (spectral.hip): with
// T = L L^T, the eigenvalues and eigenvectors Q of L^T K L, and the
// right-hand sides and coordinates in that basis.
struct SpectralWs {
    int n = 0, mask = 0;   // n = 0: no decomposition held
    bool have_q = false;   // Q holds the eigenvectors (else the tridiagonal reduction's reflectors)
    DevBuf<double> L;      // n x n Cholesky factor of T (lower triangle)
    DevBuf<double> Q;      // n x n orthonormal eigenvectors of L^T K L
    DevBuf<double> lam;    // n eigenvalues
    DevBuf<double> c6;     // Q^T L^T b: Re of the three directions, then Im (n x 6)
    DevBuf<double> w3;     // Q^T L^-1 r: coordinates x, y, z (n x 3)
    void release() {
        L.release();
        Q.release();
        lam.release();
        c6.release();
        w3.release();
        n = 0;
        have_q = false;
    }
};

struct Stats {
    static constexpr int kN = 32;
    double v[kN] = {};
};

// Kernel timers that do not stall the host: a ring of event pairs recorded
// on the context's stream.  An elapsed time is added into its stats entry
// once the pair is known complete: when the ring comes round to it again, or
// at drain (end of a solve, reading the stats).
struct TimerRing {
    static constexpr int kN = 64;
    hipEvent_t beg[kN] = {}, end[kN] = {};
    int stat[kN] = {};
    bool open[kN] = {}, pending[kN] = {};
    int head = 0;
};

}  // namespace pxh

struct pxh_ctx {
    int device = 0;
    size_t budget = 0;
    hipStream_t stream = nullptr;
    pxh::TimerRing timers;
    pxh::Atoms atoms;
    pxh::Mesh mesh;
    int target = 0;                       // PXH_WFQ / PXH_WFQFMU / PXH_BEM
    std::vector<double> alpha;            // nt x n_freq complex (interleaved)
    int alpha_nf = 0;
    std::vector<double> eps;              // n_freq complex
    int eps_nf = 0;
    pxh::Options opt;
    pxh::Stats stats;
    std::string last_error;
    pxh::DevBuf<double2> work_part, work_t;  // scratch of the operator products
    pxh::DevBuf<_Float16> work_xh, work_xl;   // fp16 hi / lo operand of the MFMA products
    pxh::DevBuf<float> work_scale;            // per-column scales of that operand
    pxh::DevBuf<double> work_colmax;          // chunk maxima behind those scales
    pxh::DevBuf<double> work_pack;            // packed operand of the BEM matrix products
    pxh::DevBuf<double> work_gemm;            // split-K partial sums of those products
    pxh::DevBuf<double2> work_lu;             // dense path: the factorised matrices of one frequency chunk
    pxh::DevBuf<int> work_piv;                // their row interchanges, and the permutations built from them
    pxh::DevBuf<double2> work_panel;          // dense path: column-major copy of the current panel (tall panels)
    pxh::SpectralWs spectral;                 // dense omegaFQ path: the eigendecomposition of the call
    void *blas = nullptr;                     // rocBLAS handle of the dense path (created on first use)
    void *solver_ws = nullptr;                // work buffers of the iterative solvers (solver.hip, grow-only)
    bool use_stored = false;                  // in-memory method: products use the stored T
    bool capturing = false;                   // a HIP graph is being captured on the stream: no timers, no launch stats
    // IEF operator form used by the products: 0 the system matrix
    // -(D + lambda I) S; 1 the right-preconditioned -(D + lambda I) (unknown
    // u = S q); 2 the matrix S alone
    int bem_op = 0;
};

namespace pxh {
// api.hip: timers on the context's stream (TimerRing); timer_begin returns -1
// when the ring has no free pair, and timer_end(-1) does nothing.
int timer_begin(pxh_ctx *c);
void timer_end(pxh_ctx *c, int t, int stat_index);
void timer_drain(pxh_ctx *c);
// dense.hip: releases the context's rocBLAS handle
void release_blas(pxh_ctx *c);
// solver.hip: releases the context's solver work buffers
void release_solver_ws(pxh_ctx *c);
// noalloc.hip: looks up this device's scratch array of pxh_matvec_noalloc
// (once per device, with the device current)
void scratch_setup(int device);
// kernels_atoms.hip
void build_atoms_device(pxh_ctx *c);
void build_kbar(pxh_ctx *c);
// Per-system description for the batched operator application.
struct SysParam {
    double2 zq[kMaxTypes];      // z_q(omega) per type
    double2 inv_alpha[kMaxTypes];
    double2 wratio[kMaxTypes * kMaxTypes];  // w_tj / w_ti (alloys), index ti + nt*tj
};
// y[s] = A(omega_s) x[s] for the listed systems; x, y are [nsys_total][nv]
// complex (double2); sys_list holds the systems to process (device).
void induced_field(pxh_ctx *c, int n_points, const double *d_pts, const double2 *d_var, int i_pol, double e0,
                   double2 *d_E);
void bem_induced_field(pxh_ctx *c, int n_tess, const double *h_centroids, const double2 *d_q, int n_points,
                       const double *d_pts, int i_pol, double e0, double2 *d_E);
void plasmon_density(pxh_ctx *c, int n_points, const double *d_pts, const double2 *d_var, double2 *d_rho,
                     double2 *d_rho_q, double2 *d_rho_mu);
// ext_part / ext_t (both or neither): caller-provided room for the split
// partials (fp64_product_partials(c, nsys) entries) and for t (nsys x n_var
// entries); null: the context's product buffers.
void apply_atoms_fp64(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x,
                      double2 *y, int ld, double2 *ext_part = nullptr, double2 *ext_t = nullptr);
size_t fp64_product_partials(const pxh_ctx *c, int nsys);
void apply_atoms_fp32(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x,
                      double2 *y, int ld);
void apply_atoms_stored(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x,
                        double2 *y, int ld);
void apply_atoms_mfma(pxh_ctx *c, int nsys, const int *d_sys_list, const SysParam *d_par, const double2 *x,
                      double2 *y, int ld);
void prepare_stored_t(pxh_ctx *c);
// squared radius beyond which the Gaussian damping is below FP64 / FP32 rounding
double far_field_r2(const Atoms &A, bool fp64);
float mfma_t_scale(const Atoms &A);
void dense_t_atoms(pxh_ctx *c, double *d_T, bool zero_qq_diag);
void dynamic_rhs_atoms(pxh_ctx *c, double freq, double e0, int rhs_form, double2 *d_rhs /* nv x 3 */);
void constant_matrix_atoms(pxh_ctx *c, double *d_m);
SysParam make_sysparam(const pxh_ctx *c, double freq, int i_freq);
// BEM shift lambda(omega) of entry i_freq of the permittivity table (solver.hip)
double2 bem_lambda(const pxh_ctx *c, int i_freq);
// dinv = 1 / diag(A(omega)) of one system (n_var complex, device)
void precond_diag_atoms(pxh_ctx *c, const SysParam *d_par_one, double2 *dinv);
void precond_diag_bem(pxh_ctx *c, const double2 *d_lambda_one, double2 *dinv);
// kernels_bem.hip
void build_mesh_device(pxh_ctx *c);
void apply_bem(pxh_ctx *c, int nsys, const int *d_sys_list, const double2 *d_lambda, const double2 *x, double2 *y,
               int ld, bool fp32);
void dynamic_rhs_bem(pxh_ctx *c, double e0, double2 *d_rhs);
// dense.hip: dense inversions of many atomistic targets together (all on
// one device, the same frequencies); polar n_ctx x n_freq x 9 complex,
// results n_ctx blocks of n_freq x 20 (column-major)
int solve_group(pxh_ctx *const *ctxs, int n_ctx, int n_freq, const double *freq, double e0, int rhs_form,
                int pol_mask, double *polar, double *results);
// api.hip: the device gate.  Every entry point that launches work holds the
// gate shared for its duration (DeviceShared, taken by the entry points'
// wrapper); the spectral route holds it exclusively (DeviceExclusive: the
// caller's shared hold is given up while waiting, new shared holders wait
// behind a waiting exclusive one, and the device is synchronised once the
// gate is held), so its rocSOLVER / rocBLAS work runs with no other work of
// the process on the device.  Decompositions that ran beside other
// contexts' work (LU or GMRES requests of the request loop, other
// decompositions) gave wrong eigenpairs on this ROCm release; alone they
// match the LU to 1e-13.  Holds taken by a thread inside its own exclusive
// hold (shared or exclusive) do nothing.
struct DeviceShared {
    DeviceShared();
    ~DeviceShared();
    DeviceShared(const DeviceShared &) = delete;
    DeviceShared &operator=(const DeviceShared &) = delete;
};
struct DeviceExclusive {
    DeviceExclusive();
    ~DeviceExclusive();
    DeviceExclusive(const DeviceExclusive &) = delete;
    DeviceExclusive &operator=(const DeviceExclusive &) = delete;

  private:
    int held_;  // shared holds of this thread given up while exclusive
};

// spectral.hip: the homogeneous omegaFQ inversion by one eigendecomposition
bool spectral_applicable(const pxh_ctx *c);
bool spectral_pays(const pxh_ctx *c, int n, int n_freq);
bool spectral_decompose(pxh_ctx *c, const double2 *d_rhs /* n x 3 */, int pol_mask, bool need_q);
void spectral_charges(pxh_ctx *c, int n_freq, const double *freq, double2 *d_sol, size_t stride);
void spectral_polar(pxh_ctx *c, int n_freq, const double *freq, double2 *d_out /* 9 per frequency */);
// solver.hip: the 20 result columns of one frequency from its 3 x 3 complex polarizability
void results_row(double freq, const double *pol, double *r);
// dense.hip: static FQ / FQFmu (what 0 ground state, 1 static field)
int solve_static(pxh_ctx *c, int what, double e0, int rhs_form, double *variables, double *polar, double *energy);
// solver.cpp
int solve_spectrum(pxh_ctx *c, int n_freq, const double *freq, double e0, int rhs_form, int pol_mask, int method,
                   double tol, int kdim, int precision, double *variables, double *polar, double *results,
                   int32_t *iters, double *true_res);
// HBM that one pxh_solve_spectrum call allocates beyond the context's
// resident data (upper estimate).
size_t estimate_solve_bytes(const pxh_ctx *c, int n_freq, int pol_mask, int method, int kdim, int precision);
}  // namespace pxh
