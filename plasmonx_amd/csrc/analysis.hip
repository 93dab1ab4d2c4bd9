// Induced field and plasmon density at points: the post-processing sums of
// plasmonX_analysis.
//
// For the variables of one frequency and one field direction, the field at
// a point p is the incident e0 along that direction plus, for every atom i
// with d = r_i - p and |d| > 0.1 bohr,
//   the field of its Gaussian charge (width R_q of atom i)
//     E -= (erf(|d|/R) - 2|d|/(sqrt(pi) R) exp(-|d|^2/R^2)) q_i d / |d|^3
//   (src/matrix/matrix_module.f90:1140-1172), and for omegaFQFmu of its
//   Gaussian dipole (width R_mu of atom i): E += T mu_i,
//     T_ab = (-3 d_a d_b / |d|^5 + delta_ab / |d|^3) fD1 + fD2 d_a d_b / |d|^2
//   (:1174-1246),
// summed over the atoms as src/wfq/wfq_module.f90:592-617 and
// src/wfq/wfqfmu_module.f90:973-1003.  The reference evaluates one point at
// a time on the host (grid_module.f90:200-242); here a work-item owns a
// point and the atoms stream through LDS in tiles, FP64 throughout.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pxh_internal.h"

namespace pxh {

namespace {

constexpr int kFieldTile = 256;
constexpr double kInvSqrtPi = 0.56418958354775628;  // 1 / sqrt(pi)
constexpr double kToBohr = 1.8897261254578281;       // bohr per Angstrom

struct TypeRadii {
    double rq[kMaxTypes], rmu[kMaxTypes];
};

__global__ __launch_bounds__(kFieldTile) void k_induced_field(int n_points, const double *__restrict__ pts, int n,
                                                              const double4 *__restrict__ pos, TypeRadii radii,
                                                              int fmu, const double2 *__restrict__ var, int i_pol,
                                                              double e0, int slice, size_t part_stride,
                                                              double2 *__restrict__ E) {
    __shared__ double4 sp[kFieldTile];
    __shared__ double2 sq[kFieldTile];
    __shared__ double2 smu[3][kFieldTile];
    // per-atom constants: 1/R_q, 1/R_q^2, 2/(sqrt(pi) R_q), 1/R_mu, 1/R_mu^2,
    // 2/(sqrt(pi) R_mu), 4/(sqrt(pi) R_mu^3): one division per pair remains
    __shared__ double sr[7][kFieldTile];
    const int tid = threadIdx.x;
    const int p = blockIdx.x * kFieldTile + tid;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (p < n_points) {
        px = pts[3 * (size_t)p];
        py = pts[3 * (size_t)p + 1];
        pz = pts[3 * (size_t)p + 2];
    }
    const double inc = blockIdx.y == 0 ? e0 : 0.0;  // the incident field, once
    double2 ex = make_double2(i_pol == 0 ? inc : 0.0, 0.0);
    double2 ey = make_double2(i_pol == 1 ? inc : 0.0, 0.0);
    double2 ez = make_double2(i_pol == 2 ? inc : 0.0, 0.0);
    const int a0 = blockIdx.y * slice, a1 = min(n, a0 + slice);
    for (int t0 = a0; t0 < a1; t0 += kFieldTile) {
        const int j = t0 + tid;
        if (j < a1) {
            const double4 a = pos[j];
            const int t = (int)a.w;
            sp[tid] = a;
            sq[tid] = var[j];
            const double iq = 1.0 / radii.rq[t], im = 1.0 / radii.rmu[t];
            sr[0][tid] = iq;
            sr[1][tid] = iq * iq;
            sr[2][tid] = 2.0 * kInvSqrtPi * iq;
            sr[3][tid] = im;
            sr[4][tid] = im * im;
            sr[5][tid] = 2.0 * kInvSqrtPi * im;
            sr[6][tid] = 4.0 * kInvSqrtPi * im * im * im;
            if (fmu)
                for (int c = 0; c < 3; ++c) smu[c][tid] = var[n + 3 * (size_t)j + c];
        }
        __syncthreads();
        const int m = min(kFieldTile, a1 - t0);
        for (int jj = 0; jj < m; ++jj) {
            const double4 a = sp[jj];
            const double dx = a.x - px, dy = a.y - py, dz = a.z - pz;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double r = sqrt(r2);
            if (!(r > 1.0e-1)) continue;
            const double inv_r = 1.0 / r, inv_r2 = inv_r * inv_r, inv_r3 = inv_r2 * inv_r;
            const double fq = erf(r * sr[0][jj]) - sr[2][jj] * r * exp(-r2 * sr[1][jj]);
            const double2 q = sq[jj];
            const double cq = fq * inv_r3;
            ex.x -= cq * dx * q.x;
            ex.y -= cq * dx * q.y;
            ey.x -= cq * dy * q.x;
            ey.y -= cq * dy * q.y;
            ez.x -= cq * dz * q.x;
            ez.y -= cq * dz * q.y;
            if (fmu) {
                const double em = exp(-r2 * sr[4][jj]);
                const double f1 = erf(r * sr[3][jj]) - sr[5][jj] * r * em;
                const double f2 = sr[6][jj] * em;
                // T = (f2 / r^2 - 3 f1 / r^5) d d^T + (f1 / r^3) I
                const double g = f2 * inv_r2 - 3.0 * f1 * inv_r3 * inv_r2;
                const double h = f1 * inv_r3;
                const double2 mx = smu[0][jj], my = smu[1][jj], mz = smu[2][jj];
                const double2 dm = make_double2(dx * mx.x + dy * my.x + dz * mz.x, dx * mx.y + dy * my.y + dz * mz.y);
                ex.x += g * dx * dm.x + h * mx.x;
                ex.y += g * dx * dm.y + h * mx.y;
                ey.x += g * dy * dm.x + h * my.x;
                ey.y += g * dy * dm.y + h * my.y;
                ez.x += g * dz * dm.x + h * mz.x;
                ez.y += g * dz * dm.y + h * mz.y;
            }
        }
        __syncthreads();
    }
    if (p < n_points) {
        double2 *out = E + blockIdx.y * part_stride;
        out[3 * (size_t)p] = ex;
        out[3 * (size_t)p + 1] = ey;
        out[3 * (size_t)p + 2] = ez;
    }
}

// Plasmon density at a point p: for every atom i with d = r_i - p, the
// normalised Gaussian of its charge, q_i exp(-|d|^2/R_q^2) / (pi^{3/2} R_q^3)
// (src/matrix/matrix_module.f90:1258-1284), and for omegaFQFmu minus the
// derivative of the Gaussian of R_mu along its dipole,
// -2 exp(-|d|^2/R_mu^2) / (pi^{3/2} R_mu^5) (mu_i . d) (:1295-1324), summed
// as src/wfq/wfq_module.f90:625-648 and src/wfq/wfqfmu_module.f90:1012-1081.
// The per-atom constants are formed once per tile in LDS.  rho_q and rho_mu
// (the two parts) are written when not null.
__global__ __launch_bounds__(kFieldTile) void k_plasmon_density(int n_points, const double *__restrict__ pts, int n,
                                                                const double4 *__restrict__ pos, TypeRadii radii,
                                                                int fmu, const double2 *__restrict__ var, int slice,
                                                                size_t part_stride, double2 *__restrict__ rho,
                                                                double2 *__restrict__ rho_q,
                                                                double2 *__restrict__ rho_mu) {
    __shared__ double4 sp[kFieldTile];
    __shared__ double2 sq[kFieldTile];
    __shared__ double2 smu[3][kFieldTile];
    __shared__ double sc[4][kFieldTile];  // 1/R_q^2, 1/(pi^1.5 R_q^3), 1/R_mu^2, 2/(pi^1.5 R_mu^5)
    const int tid = threadIdx.x;
    const int p = blockIdx.x * kFieldTile + tid;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (p < n_points) {
        px = pts[3 * (size_t)p];
        py = pts[3 * (size_t)p + 1];
        pz = pts[3 * (size_t)p + 2];
    }
    constexpr double kInvPi32 = 0.17958712212516656;  // 1 / pi^{3/2}
    double2 aq = make_double2(0.0, 0.0), am = make_double2(0.0, 0.0);
    const int a0 = blockIdx.y * slice, a1 = min(n, a0 + slice);
    for (int t0 = a0; t0 < a1; t0 += kFieldTile) {
        const int j = t0 + tid;
        if (j < a1) {
            const double4 a = pos[j];
            const int t = (int)a.w;
            sp[tid] = a;
            sq[tid] = var[j];
            const double R = radii.rq[t];
            sc[0][tid] = 1.0 / (R * R);
            sc[1][tid] = kInvPi32 / (R * R * R);
            if (fmu) {
                const double Rm = radii.rmu[t];
                sc[2][tid] = 1.0 / (Rm * Rm);
                sc[3][tid] = 2.0 * kInvPi32 / (Rm * Rm * Rm * Rm * Rm);
                for (int c = 0; c < 3; ++c) smu[c][tid] = var[n + 3 * (size_t)j + c];
            }
        }
        __syncthreads();
        const int m = min(kFieldTile, a1 - t0);
        for (int jj = 0; jj < m; ++jj) {
            const double4 a = sp[jj];
            const double dx = a.x - px, dy = a.y - py, dz = a.z - pz;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double gq = sc[1][jj] * exp(-r2 * sc[0][jj]);
            const double2 q = sq[jj];
            aq.x += gq * q.x;
            aq.y += gq * q.y;
            if (fmu) {
                const double gm = sc[3][jj] * exp(-r2 * sc[2][jj]);
                const double2 mx = smu[0][jj], my = smu[1][jj], mz = smu[2][jj];
                am.x -= gm * (dx * mx.x + dy * my.x + dz * mz.x);
                am.y -= gm * (dx * mx.y + dy * my.y + dz * mz.y);
            }
        }
        __syncthreads();
    }
    if (p < n_points) {
        if (rho) rho[p] = make_double2(aq.x + am.x, aq.y + am.y);
        if (rho_q) rho_q[blockIdx.y * part_stride + p] = aq;
        if (rho_mu) rho_mu[blockIdx.y * part_stride + p] = am;
    }
}

// Field of BEM surface charges (point charges at the tessera centroids) at a
// point p, as calculate_induced_field_at_point_bem
// (src/bem/bem_module.f90:1151-1213): the point is inside, and its field 0,
// unless |p - m| > |c_l - m| + 0.5 Angstrom, with m the mean of the
// centroids and l the first centroid nearest to p; outside, the incident e0
// minus sum_i q_i d / |d|^3 over |d| > 0.1 bohr, d = c_i - p
// (src/matrix/matrix_module.f90:1444-1464).  One pass: the nearest centroid
// and the sum are found together and the sum is dropped for inside points.
// ci[4] holds x, y, z and |c_i - m| + 0.5 Angstrom.
__global__ __launch_bounds__(kFieldTile) void k_bem_field(int n_points, const double *__restrict__ pts, int n,
                                                          const double4 *__restrict__ ci,
                                                          const double2 *__restrict__ q, double mx, double my,
                                                          double mz, int i_pol, double e0, double2 *__restrict__ E) {
    __shared__ double4 sc[kFieldTile];
    __shared__ double2 sq[kFieldTile];
    const int tid = threadIdx.x;
    const int p = blockIdx.x * kFieldTile + tid;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (p < n_points) {
        px = pts[3 * (size_t)p];
        py = pts[3 * (size_t)p + 1];
        pz = pts[3 * (size_t)p + 2];
    }
    double2 ex = make_double2(0.0, 0.0), ey = ex, ez = ex;
    double best = 1.0e6 * kToBohr, rl = 0.0;
    for (int t0 = 0; t0 < n; t0 += kFieldTile) {
        const int j = t0 + tid;
        if (j < n) {
            sc[tid] = ci[j];
            sq[tid] = q[j];
        }
        __syncthreads();
        const int m = min(kFieldTile, n - t0);
        for (int jj = 0; jj < m; ++jj) {
            const double4 a = sc[jj];
            const double dx = a.x - px, dy = a.y - py, dz = a.z - pz;
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (best > r) {
                best = r;
                rl = a.w;
            }
            if (!(r > 1.0e-1)) continue;
            const double c = 1.0 / (r * r * r);
            const double2 v = sq[jj];
            ex.x -= dx * v.x * c;
            ex.y -= dx * v.y * c;
            ey.x -= dy * v.x * c;
            ey.y -= dy * v.y * c;
            ez.x -= dz * v.x * c;
            ez.y -= dz * v.y * c;
        }
        __syncthreads();
    }
    if (p < n_points) {
        const double rp = sqrt((px - mx) * (px - mx) + (py - my) * (py - my) + (pz - mz) * (pz - mz));
        const bool out = rp > rl;
        if (out) {
            if (i_pol == 0) ex.x += e0;
            if (i_pol == 1) ey.x += e0;
            if (i_pol == 2) ez.x += e0;
        }
        const double2 z = make_double2(0.0, 0.0);
        E[3 * (size_t)p] = out ? ex : z;
        E[3 * (size_t)p + 1] = out ? ey : z;
        E[3 * (size_t)p + 2] = out ? ez : z;
    }
}

// Sum of the per-slice partial results, in slice order (deterministic):
// parts[s * count + i] for s < n_slices.
__global__ void k_sum_slices(size_t count, int n_slices, const double2 *__restrict__ parts,
                             double2 *__restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    double2 acc = parts[i];
    for (int s = 1; s < n_slices; ++s) {
        const double2 v = parts[s * count + i];
        acc.x += v.x;
        acc.y += v.y;
    }
    out[i] = acc;
}

// rho = q part + mu part, each summed over the slices; the parts on request.
__global__ void k_sum_density(int n_points, int n_slices, const double2 *__restrict__ pq,
                              const double2 *__restrict__ pm, double2 *__restrict__ rho,
                              double2 *__restrict__ rho_q, double2 *__restrict__ rho_mu) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_points) return;
    double2 q = pq[p], m = pm[p];
    for (int s = 1; s < n_slices; ++s) {
        const double2 a = pq[(size_t)s * n_points + p], b = pm[(size_t)s * n_points + p];
        q.x += a.x;
        q.y += a.y;
        m.x += b.x;
        m.y += b.y;
    }
    rho[p] = make_double2(q.x + m.x, q.y + m.y);
    if (rho_q) rho_q[p] = q;
    if (rho_mu) rho_mu[p] = m;
}

// Atom slices per point block so that a launch holds about 8 workgroups per
// CU: a 30^3 grid is 106 blocks of points, fewer than the 256 CUs.
int atom_slices(int point_blocks, int n) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    const int want = (8 * cus + point_blocks - 1) / point_blocks;
    const int max_slices = (n + kFieldTile - 1) / kFieldTile;  // at least one tile per slice
    return std::max(1, std::min(want, max_slices));
}

}  // namespace

void plasmon_density(pxh_ctx *c, int n_points, const double *d_pts, const double2 *d_var, double2 *d_rho,
                     double2 *d_rho_q, double2 *d_rho_mu) {
    const Atoms &A = c->atoms;
    TypeRadii radii{};
    for (int t = 0; t < A.nt; ++t) {
        radii.rq[t] = A.rq[t];
        radii.rmu[t] = A.model == PXH_WFQFMU ? A.rmu[t] : 1.0;
    }
    const int blocks = (n_points + kFieldTile - 1) / kFieldTile;
    if (blocks == 0) return;
    const int fmu = A.model == PXH_WFQFMU ? 1 : 0;
    const int S = A.n == 0 ? 1 : atom_slices(blocks, A.n);
    const int slice = A.n == 0 ? 1 : (A.n + S - 1) / S;
    if (S == 1) {
        k_plasmon_density<<<blocks, kFieldTile, 0, c->stream>>>(n_points, d_pts, A.n, A.pos.p, radii, fmu, d_var, slice,
                                                                0, d_rho, d_rho_q, d_rho_mu);
        PXH_HIP(hipGetLastError());
        return;
    }
    DevBuf<double2> part;
    const size_t np = (size_t)n_points;
    part.alloc(2 * (size_t)S * np);
    k_plasmon_density<<<dim3(blocks, S), kFieldTile, 0, c->stream>>>(n_points, d_pts, A.n, A.pos.p, radii, fmu, d_var,
                                                                     slice, np, nullptr, part.p, part.p + S * np);
    PXH_HIP(hipGetLastError());
    k_sum_density<<<blocks, kFieldTile, 0, c->stream>>>(n_points, S, part.p, part.p + S * np, d_rho, d_rho_q,
                                                        d_rho_mu);
    PXH_HIP(hipGetLastError());
    PXH_HIP(hipStreamSynchronize(c->stream));  // before the partials are freed
}

void bem_induced_field(pxh_ctx *c, int n_tess, const double *h_centroids, const double2 *d_q, int n_points,
                       const double *d_pts, int i_pol, double e0, double2 *d_E) {
    double m[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < n_tess; ++i)
        for (int k = 0; k < 3; ++k) m[k] += h_centroids[3 * i + k] / n_tess;
    std::vector<double4> ci(n_tess);
    for (int i = 0; i < n_tess; ++i) {
        const double *x = h_centroids + 3 * i;
        const double r = std::sqrt((x[0] - m[0]) * (x[0] - m[0]) + (x[1] - m[1]) * (x[1] - m[1]) +
                                   (x[2] - m[2]) * (x[2] - m[2]));
        ci[i] = make_double4(x[0], x[1], x[2], r + 0.5 * kToBohr);
    }
    DevBuf<double4> dci;
    dci.upload(ci.data(), (size_t)n_tess, c->stream);
    const int blocks = (n_points + kFieldTile - 1) / kFieldTile;
    if (blocks > 0)
        k_bem_field<<<blocks, kFieldTile, 0, c->stream>>>(n_points, d_pts, n_tess, dci.p, d_q, m[0], m[1], m[2], i_pol,
                                                          e0, d_E);
    PXH_HIP(hipGetLastError());
    PXH_HIP(hipStreamSynchronize(c->stream));  // before the centroids are freed
}

void induced_field(pxh_ctx *c, int n_points, const double *d_pts, const double2 *d_var, int i_pol, double e0,
                   double2 *d_E) {
    const Atoms &A = c->atoms;
    TypeRadii radii{};
    for (int t = 0; t < A.nt; ++t) {
        radii.rq[t] = A.rq[t];
        radii.rmu[t] = A.model == PXH_WFQFMU ? A.rmu[t] : 1.0;
    }
    const int blocks = (n_points + kFieldTile - 1) / kFieldTile;
    if (blocks == 0) return;
    const int fmu = A.model == PXH_WFQFMU ? 1 : 0;
    const int S = A.n == 0 ? 1 : atom_slices(blocks, A.n);
    const int slice = A.n == 0 ? 1 : (A.n + S - 1) / S;
    if (S == 1) {
        k_induced_field<<<blocks, kFieldTile, 0, c->stream>>>(n_points, d_pts, A.n, A.pos.p, radii, fmu, d_var, i_pol,
                                                              e0, slice, 0, d_E);
        PXH_HIP(hipGetLastError());
        return;
    }
    DevBuf<double2> part;
    const size_t count = 3 * (size_t)n_points;
    part.alloc((size_t)S * count);
    k_induced_field<<<dim3(blocks, S), kFieldTile, 0, c->stream>>>(n_points, d_pts, A.n, A.pos.p, radii, fmu, d_var,
                                                                   i_pol, e0, slice, count, part.p);
    PXH_HIP(hipGetLastError());
    k_sum_slices<<<(unsigned)((count + kFieldTile - 1) / kFieldTile), kFieldTile, 0, c->stream>>>(count, S, part.p,
                                                                                                   d_E);
    PXH_HIP(hipGetLastError());
    PXH_HIP(hipStreamSynchronize(c->stream));  // before the partials are freed
}

}  // namespace pxh
