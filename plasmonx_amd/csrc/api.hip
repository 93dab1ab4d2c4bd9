// C-ABI entry points of libpxh (declared in include/pxh.h).
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <mutex>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "pxh_internal.h"

using pxh::Error;

namespace {

thread_local std::string g_last_error;  // failures before a context exists

template <typename F>
int guarded(pxh_ctx *c, F &&f) {
    try {
        if (c) PXH_HIP(hipSetDevice(c->device));
        pxh::DeviceShared gate;
        f();
        return PXH_OK;
    } catch (const Error &e) {
        if (c) c->last_error = e.msg;
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        if (c) c->last_error = "host allocation failed";
        g_last_error = "host allocation failed";
        return PXH_ENOMEM;
    }
}

}  // namespace

namespace pxh {

namespace {

// Writer-preferring shared / exclusive gate over the process's device work.
struct Gate {
    std::mutex m;
    std::condition_variable cv;
    int readers = 0, writers_waiting = 0;
    bool writer = false;
} g_gate;
thread_local int t_shared = 0;  // nesting depth of this thread's shared holds
thread_local int t_excl = 0;    // nesting depth of this thread's exclusive holds

void gate_lock_shared() {
    std::unique_lock<std::mutex> lk(g_gate.m);
    g_gate.cv.wait(lk, [] { return !g_gate.writer && g_gate.writers_waiting == 0; });
    ++g_gate.readers;
}

void gate_unlock_shared() {
    std::lock_guard<std::mutex> lk(g_gate.m);
    if (--g_gate.readers == 0) g_gate.cv.notify_all();
}

}  // namespace

// Inside this thread's exclusive hold the device is already its own: shared
// and nested exclusive holds taken there do nothing (the holds nest LIFO, so
// each destructor sees the state its constructor saw).
DeviceShared::DeviceShared() {
    if (t_excl) return;
    if (t_shared++ == 0) gate_lock_shared();
}

DeviceShared::~DeviceShared() {
    if (t_excl) return;
    if (--t_shared == 0) gate_unlock_shared();
}

DeviceExclusive::DeviceExclusive() : held_(t_shared) {
    if (t_excl++) return;
    if (held_) gate_unlock_shared();
    t_shared = 0;
    {
        std::unique_lock<std::mutex> lk(g_gate.m);
        ++g_gate.writers_waiting;
        g_gate.cv.wait(lk, [] { return !g_gate.writer && g_gate.readers == 0; });
        --g_gate.writers_waiting;
        g_gate.writer = true;
    }
    (void)hipDeviceSynchronize();  // work left on other streams by calls that have returned
}

DeviceExclusive::~DeviceExclusive() {
    if (--t_excl) return;
    (void)hipDeviceSynchronize();
    {
        std::lock_guard<std::mutex> lk(g_gate.m);
        g_gate.writer = false;
    }
    g_gate.cv.notify_all();
    t_shared = held_;
    if (held_) gate_lock_shared();
}

namespace {
void timer_collect(pxh_ctx *c, int t) {
    TimerRing &R = c->timers;
    PXH_HIP(hipEventSynchronize(R.end[t]));
    float ms = 0.f;
    PXH_HIP(hipEventElapsedTime(&ms, R.beg[t], R.end[t]));
    c->stats.v[R.stat[t]] += ms;
    R.pending[t] = false;
}
}  // namespace

int timer_begin(pxh_ctx *c) {
    if (c->capturing) return -1;  // an event record would become a graph node
    TimerRing &R = c->timers;
    const int t = R.head;
    if (R.open[t]) return -1;  // as many timers open as the ring holds
    if (R.pending[t]) timer_collect(c, t);
    PXH_HIP(hipEventRecord(R.beg[t], c->stream));
    R.open[t] = true;
    R.head = (t + 1) % TimerRing::kN;
    return t;
}

void timer_end(pxh_ctx *c, int t, int stat_index) {
    if (t < 0) return;
    TimerRing &R = c->timers;
    PXH_HIP(hipEventRecord(R.end[t], c->stream));
    R.stat[t] = stat_index;
    R.open[t] = false;
    R.pending[t] = true;
}

void timer_drain(pxh_ctx *c) {
    TimerRing &R = c->timers;
    for (int k = 0; k < TimerRing::kN; ++k) {
        const int t = (R.head + k) % TimerRing::kN;  // oldest first
        if (R.pending[t]) timer_collect(c, t);
    }
}

}  // namespace pxh

extern "C" {

int pxh_ctx_create(int device, size_t hbm_budget_bytes, pxh_ctx **out) {
    if (!out) return PXH_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        g_last_error = "no HIP device available";
        return PXH_ENODEV;
    }
    if (device < 0 || device >= count) {
        g_last_error = "device index out of range";
        return PXH_EINVAL;
    }
    pxh_ctx *c = new (std::nothrow) pxh_ctx();
    if (!c) return PXH_ENOMEM;
    c->device = device;
    c->budget = hbm_budget_bytes;
    const int st = guarded(c, [&] {
        PXH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        // the scratch array of pxh_matvec_noalloc on this device; a failure
        // here is reported by that entry point, which looks it up again
        try {
            pxh::scratch_setup(device);
        } catch (const Error &) {
        }
        for (int k = 0; k < pxh::TimerRing::kN; ++k) {
            PXH_HIP(hipEventCreate(&c->timers.beg[k]));
            PXH_HIP(hipEventCreate(&c->timers.end[k]));
        }
    });
    if (st != PXH_OK) {
        delete c;
        return st;
    }
    *out = c;
    return PXH_OK;
}

int pxh_ctx_destroy(pxh_ctx *c) {
    if (!c) return PXH_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    pxh::release_blas(c);
    pxh::release_solver_ws(c);
    for (int k = 0; k < pxh::TimerRing::kN; ++k) {
        if (c->timers.beg[k]) (void)hipEventDestroy(c->timers.beg[k]);
        if (c->timers.end[k]) (void)hipEventDestroy(c->timers.end[k]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PXH_OK;
}

int pxh_set_atoms(pxh_ctx *c, int model, int kernel, int n_atoms, const double *coord_bohr,
                  const int32_t *type_of_atom, int n_types, const double *eta, const double *r_q, const double *r_mu,
                  const double *a_ij, const double *fermi_d, const double *fermi_s, const double *r0,
                  const double *tau, const double *density, const int32_t *nb_counts, const int32_t *nb_total) {
    if (!c) return PXH_EINVAL;
    return guarded(c, [&] {
        PXH_CHECK(model == PXH_WFQ || model == PXH_WFQFMU || model == PXH_FQ || model == PXH_FQFMU, PXH_EINVAL,
                  "model must be PXH_WFQ, PXH_WFQFMU, PXH_FQ or PXH_FQFMU");
        const bool stat = model == PXH_FQ || model == PXH_FQFMU;
        const bool dipoles = model == PXH_WFQFMU || model == PXH_FQFMU;
        PXH_CHECK(kernel >= PXH_GAUSSIAN && kernel <= PXH_OHNO, PXH_EINVAL, "unknown kernel");
        PXH_CHECK(n_atoms > 0 && n_types > 0 && n_types <= pxh::kMaxTypes, PXH_EINVAL,
                  "n_atoms must be > 0 and 1 <= n_types <= 8");
        PXH_CHECK(coord_bohr && type_of_atom && eta, PXH_EINVAL, "null input array");
        PXH_CHECK(stat || (a_ij && fermi_d && fermi_s && r0 && tau && density), PXH_EINVAL,
                  "the dynamic models need a_ij, the Fermi tables, tau and density");
        PXH_CHECK(kernel != PXH_GAUSSIAN || r_q, PXH_EINVAL, "the Gaussian kernel needs r_q");
        PXH_CHECK(!dipoles || (r_q && r_mu), PXH_EINVAL, "the dipole models need r_q and r_mu");
        // every check runs on local copies: a rejected call leaves the
        // context's current target as it was
        std::vector<int> type(n_atoms);
        std::vector<int> seen(n_types, 0);
        for (int i = 0; i < n_atoms; ++i) {
            PXH_CHECK(type_of_atom[i] >= 1 && type_of_atom[i] <= n_types, PXH_EINVAL, "type_of_atom out of range");
            type[i] = type_of_atom[i] - 1;
            seen[type[i]] = 1;
        }
        int used = 0;
        for (int t = 0; t < n_types; ++t) used += seen[t];
        const bool hetero = used > 1;
        // plain omegaFQ has no alloy form (wfq_module.f90:203-204)
        PXH_CHECK(!(hetero && model == PXH_WFQ), PXH_EINVAL, "heterogeneous omegaFQ is not supported");
        (void)dipoles;
        auto cp = [&](const double *p, size_t k) { return p ? std::vector<double>(p, p + k) : std::vector<double>(k, 0.0); };
        std::vector<double> v_eta = cp(eta, n_types), v_rq = cp(r_q, n_types), v_rmu = cp(r_mu, n_types),
                            v_aij = cp(a_ij, n_types), v_tau = cp(tau, n_types), v_dens = cp(density, n_types);
        for (int t = 0; t < n_types; ++t) {
            if (!stat) PXH_CHECK(v_tau[t] > 0.0 && v_dens[t] > 0.0, PXH_EINVAL, "tau and density must be positive");
            if (kernel == PXH_GAUSSIAN) PXH_CHECK(v_rq[t] > 0.0, PXH_EINVAL, "r_q must be positive");
            if (dipoles) PXH_CHECK(v_rq[t] > 0.0 && v_rmu[t] > 0.0, PXH_EINVAL, "r_q, r_mu must be positive");
        }
        std::vector<double> nbfrac((size_t)n_types * n_atoms, 0.0);
        for (int i = 0; i < n_atoms; ++i) {
            if (nb_counts && nb_total) {
                PXH_CHECK(nb_total[i] > 0, PXH_EINVAL, "nb_total must be positive");
                for (int t = 0; t < n_types; ++t)
                    nbfrac[t + (size_t)n_types * i] = (double)nb_counts[t + (size_t)n_types * i] / (double)nb_total[i];
            } else {
                nbfrac[type[i] + (size_t)n_types * i] = 1.0;
            }
        }
        // from here on the context's atoms change: a failure below (device
        // memory, HIP) leaves the context without a target
        c->target = 0;
        pxh::Atoms &A = c->atoms;
        A.n = 0;
        A.model = dipoles ? PXH_WFQFMU : PXH_WFQ;
        A.is_static = stat;
        A.chi.clear();
        A.alpha_s.clear();
        A.mol_count.clear();
        A.kernel = kernel;
        A.nt = n_types;
        A.hetero = hetero;
        A.h_coord.assign(coord_bohr, coord_bohr + 3 * (size_t)n_atoms);
        A.h_type = std::move(type);
        A.eta = std::move(v_eta);
        A.rq = std::move(v_rq);
        A.rmu = std::move(v_rmu);
        A.aij = std::move(v_aij);
        A.tau = std::move(v_tau);
        A.dens = std::move(v_dens);
        A.fd = cp(fermi_d, (size_t)n_types * n_types);
        A.fs = cp(fermi_s, (size_t)n_types * n_types);
        A.r0 = cp(r0, (size_t)n_types * n_types);
        A.h_nbfrac = std::move(nbfrac);
        A.n = n_atoms;
        try {
            pxh::build_atoms_device(c);
        } catch (...) {
            A.n = 0;
            throw;
        }
        c->target = model;
    });
}

int pxh_set_static(pxh_ctx *c, int n_types, const double *chi, const double *alpha, int n_mol,
                   const int32_t *atoms_per_molecule) {
    if (!c) return PXH_EINVAL;
    return guarded(c, [&] {
        PXH_CHECK(c->target == PXH_FQ || c->target == PXH_FQFMU, PXH_EINVAL, "no static FQ / FQFMu target set");
        pxh::Atoms &A = c->atoms;
        PXH_CHECK(n_types == A.nt && chi, PXH_EINVAL, "chi must be given for the n_types of the current atoms");
        PXH_CHECK(c->target == PXH_FQ || alpha, PXH_EINVAL, "FQFMu needs alpha");
        PXH_CHECK(n_mol >= 0 && (n_mol == 0 || atoms_per_molecule), PXH_EINVAL, "bad molecule table");
        std::vector<int> mc(atoms_per_molecule, atoms_per_molecule + n_mol);
        long total = 0;
        for (int k : mc) {
            PXH_CHECK(k > 0, PXH_EINVAL, "every molecule needs at least one atom");
            total += k;
        }
        PXH_CHECK(n_mol == 0 || total == A.n, PXH_EINVAL, "atoms_per_molecule must add up to n_atoms");
        std::vector<double> al(n_types, 0.0);
        if (alpha)
            for (int t = 0; t < n_types; ++t) {
                al[t] = alpha[t];
                PXH_CHECK(c->target != PXH_FQFMU || al[t] > 0.0, PXH_EINVAL, "alpha must be positive");
            }
        A.chi.assign(chi, chi + n_types);
        A.alpha_s = al;
        A.mol_count = mc;
    });
}

int pxh_solve_static(pxh_ctx *c, int what, double e0, int rhs_form, double *variables, double *polar,
                     double *energy) {
    if (!c) return PXH_EINVAL;
    int status = PXH_OK;
    const int g = guarded(c, [&] {
        PXH_CHECK(c->target == PXH_FQ || c->target == PXH_FQFMU, PXH_EINVAL, "no static FQ / FQFMu target set");
        PXH_CHECK((int)c->atoms.chi.size() == c->atoms.nt, PXH_EINVAL, "pxh_set_static has not been called");
        PXH_CHECK(what == PXH_ENERGY || what == PXH_STATIC_FIELD, PXH_EINVAL, "what must be ENERGY or STATIC_FIELD");
        PXH_CHECK(rhs_form == PXH_RHS_FIELD || rhs_form == PXH_RHS_POTENTIAL, PXH_EINVAL, "unknown rhs_form");
        // the reference stops here: fqfmu_module.f90:340-347
        PXH_CHECK(!(polar && what == PXH_STATIC_FIELD && c->target == PXH_FQFMU), PXH_EINVAL,
                  "the static polarizability of FQFMu is not implemented (as in the reference)");
        status = pxh::solve_static(c, what, e0, rhs_form, variables, polar, energy);
        if (status == PXH_ESINGULAR) c->last_error = "singular pivot in the static solve";
    });
    return g != PXH_OK ? g : status;
}

int pxh_set_alpha(pxh_ctx *c, int n_types, int n_freq, const double *alpha_w) {
    if (!c) return PXH_EINVAL;
    return guarded(c, [&] {
        PXH_CHECK(n_types == c->atoms.nt && n_freq >= 0 && (alpha_w || n_freq == 0), PXH_EINVAL,
                  "alpha table must be n_types x n_freq of the current atoms");
        c->alpha.assign(alpha_w, alpha_w + 2 * (size_t)n_types * n_freq);
        c->alpha_nf = n_freq;
    });
}

int pxh_set_mesh(pxh_ctx *c, int variant, int green, int n_tess, const double *verts9, double normal_factor,
                 double sphere_r_bohr, double eps_solvent, int charge_constraint) {
    if (!c) return PXH_EINVAL;
    return guarded(c, [&] {
        PXH_CHECK(variant == PXH_DPCM || variant == PXH_IEFPCM, PXH_EINVAL, "variant must be DPCM or IEFPCM");
        PXH_CHECK(green == PXH_APPROXIMATE || green == PXH_ACCURATE, PXH_EINVAL, "unknown green function");
        PXH_CHECK(n_tess > 0 && verts9, PXH_EINVAL, "empty mesh");
        PXH_CHECK(green == PXH_ACCURATE || sphere_r_bohr > 0.0, PXH_EINVAL,
                  "the approximate Green function needs a sphere radius");
        // a failure while the mesh is built leaves the context without a target
        c->target = 0;
        pxh::Mesh &M = c->mesh;
        M.n = n_tess;
        M.variant = variant;
        M.green = green;
        M.normal_factor = normal_factor;
        M.sphere_r = sphere_r_bohr;
        M.eps_solvent = eps_solvent;
        M.charge_constraint = charge_constraint != 0;
        M.h_verts.assign(verts9, verts9 + 9 * (size_t)n_tess);
        pxh::build_mesh_device(c);
        c->target = PXH_BEM;
    });
}

int pxh_set_permittivity(pxh_ctx *c, int n_freq, const double *eps_w) {
    if (!c) return PXH_EINVAL;
    return guarded(c, [&] {
        PXH_CHECK(n_freq >= 0 && (eps_w || n_freq == 0), PXH_EINVAL, "bad permittivity table");
        c->eps.assign(eps_w, eps_w + 2 * (size_t)n_freq);
        c->eps_nf = n_freq;
    });
}

int pxh_solve_spectrum(pxh_ctx *c, int n_freq, const double *freq_au, double e0, int rhs_form, int pol_mask,
                       int method, double tol_abs, int krylov_dim, int precision, double *variables, double *polar,
                       double *results, int32_t *iters, double *true_residual) {
    if (!c) return PXH_EINVAL;
    int status = PXH_OK;
    const int g = guarded(c, [&] {
        PXH_CHECK(c->target == PXH_WFQ || c->target == PXH_WFQFMU || c->target == PXH_BEM, PXH_EINVAL,
                  "no dynamic target set (static FQ / FQFMu targets take pxh_solve_static)");
        PXH_CHECK(n_freq > 0 && freq_au, PXH_EINVAL, "no frequencies");
        PXH_CHECK(pol_mask > 0 && pol_mask < 8, PXH_EINVAL, "pol_mask must select x, y and/or z");
        PXH_CHECK(method >= PXH_INVERSION && method <= PXH_ITERATIVE_OTF, PXH_EINVAL, "unknown method");
        PXH_CHECK(method == PXH_INVERSION || (tol_abs > 0.0 && krylov_dim > 0), PXH_EINVAL,
                  "GMRES needs tol_abs > 0 and krylov_dim > 0");
        PXH_CHECK(!(c->target == PXH_BEM && method == PXH_ITERATIVE_OTF), PXH_EINVAL,
                  "BEM has no on-the-fly product (bem_module.f90:91-135)");
        PXH_CHECK(rhs_form == PXH_RHS_FIELD || rhs_form == PXH_RHS_POTENTIAL, PXH_EINVAL, "unknown rhs_form");
        status = pxh::solve_spectrum(c, n_freq, freq_au, e0, rhs_form, pol_mask, method, tol_abs, krylov_dim,
                                     precision, variables, polar, results, iters, true_residual);
        if (status == PXH_ENOCONV) c->last_error = "GMRES not converged within the Krylov dimension";
        if (status == PXH_ESINGULAR) c->last_error = "singular pivot in the LU factorisation";
    });
    return g != PXH_OK ? g : status;
}

int pxh_solve_group(int n_ctx, pxh_ctx *const *ctxs, int n_freq, const double *freq_au, double e0, int rhs_form,
                    int pol_mask, double *polar, double *results) {
    if (n_ctx <= 0 || !ctxs || !ctxs[0]) return PXH_EINVAL;
    pxh_ctx *c0 = ctxs[0];
    int status = PXH_OK;
    const int g = guarded(c0, [&] {
        PXH_CHECK(n_freq > 0 && freq_au, PXH_EINVAL, "no frequencies");
        PXH_CHECK(pol_mask > 0 && pol_mask < 8, PXH_EINVAL, "pol_mask must select x, y and/or z");
        PXH_CHECK(rhs_form == PXH_RHS_FIELD || rhs_form == PXH_RHS_POTENTIAL, PXH_EINVAL, "unknown rhs_form");
        for (int t = 0; t < n_ctx; ++t) {
            const pxh_ctx *c = ctxs[t];
            PXH_CHECK(c != nullptr, PXH_EINVAL, "null context in the group");
            PXH_CHECK(c->device == c0->device, PXH_EINVAL, "the contexts of a group must share one device");
            PXH_CHECK(c->target == PXH_WFQ || c->target == PXH_WFQFMU, PXH_EINVAL,
                      "every context of a group needs an omegaFQ / omegaFQFmu target");
            PXH_CHECK(!c->atoms.hetero, PXH_EINVAL, "grouped solves take homogeneous targets (alloys: pxh_solve_spectrum)");
            PXH_CHECK(c->atoms.model != PXH_WFQFMU || c->alpha_nf >= n_freq, PXH_EINVAL,
                      "alpha(omega) table shorter than the frequency list");
            for (int u = 0; u < t; ++u) PXH_CHECK(ctxs[u] != c, PXH_EINVAL, "a context appears twice in the group");
        }
        status = pxh::solve_group(ctxs, n_ctx, n_freq, freq_au, e0, rhs_form, pol_mask, polar, results);
        if (status == PXH_ESINGULAR) c0->last_error = "singular pivot in the LU factorisation";
    });
    return g != PXH_OK ? g : status;
}

int pxh_induced_field(pxh_ctx *c, int n_points, const double *points_bohr, const double *variables, int i_pol,
                      double e0, double *field) {
    if (!c) return PXH_EINVAL;
    return guarded(c, [&] {
        PXH_CHECK(c->target == PXH_WFQ || c->target == PXH_WFQFMU, PXH_EINVAL, "no atomistic target set");
        PXH_CHECK(c->atoms.kernel == PXH_GAUSSIAN, PXH_EINVAL, "the induced field needs the Gaussian kernel");
        PXH_CHECK(n_points >= 0 && n_points <= INT_MAX - 256 && i_pol >= 0 && i_pol < 3, PXH_EINVAL,
                  "bad point count or field direction");
        PXH_CHECK(n_points == 0 || (points_bohr && variables && field), PXH_EINVAL, "null array");
        if (n_points == 0) return;
        const int nv = c->atoms.n_var();
        pxh::DevBuf<double> dp;
        pxh::DevBuf<double2> dv, de;
        dp.upload(points_bohr, (size_t)3 * n_points, c->stream);
        dv.upload(reinterpret_cast<const double2 *>(variables), (size_t)nv, c->stream);
        de.alloc((size_t)3 * n_points);
        pxh::induced_field(c, n_points, dp.p, dv.p, i_pol, e0, de.p);
        PXH_HIP(hipMemcpyAsync(field, de.p, sizeof(double2) * 3 * (size_t)n_points, hipMemcpyDeviceToHost,
                               c->stream));
        PXH_HIP(hipStreamSynchronize(c->stream));
    });
}

int pxh_bem_induced_field(pxh_ctx *c, int n_tess, const double *centroids_bohr, const double *charges, int n_points,
                          const double *points_bohr, int i_pol, double e0, double *field) {
    if (!c) return PXH_EINVAL;
    return guarded(c, [&] {
        PXH_CHECK(n_tess > 0 && n_points >= 0 && n_points <= INT_MAX - 256 && i_pol >= 0 && i_pol < 3, PXH_EINVAL,
                  "bad tessera or point count or field direction");
        PXH_CHECK(centroids_bohr && charges && (n_points == 0 || (points_bohr && field)), PXH_EINVAL, "null array");
        if (n_points == 0) return;
        const size_t np = (size_t)n_points;
        pxh::DevBuf<double> dp;
        pxh::DevBuf<double2> dq, de;
        dp.upload(points_bohr, 3 * np, c->stream);
        dq.upload(reinterpret_cast<const double2 *>(charges), (size_t)n_tess, c->stream);
        de.alloc(3 * np);
        pxh::bem_induced_field(c, n_tess, centroids_bohr, dq.p, n_points, dp.p, i_pol, e0, de.p);
        PXH_HIP(hipMemcpyAsync(field, de.p, sizeof(double2) * 3 * np, hipMemcpyDeviceToHost, c->stream));
        PXH_HIP(hipStreamSynchronize(c->stream));
    });
}

int pxh_plasmon_density(pxh_ctx *c, int n_points, const double *points_bohr, const double *variables, double *rho,
                        double *rho_q, double *rho_mu) {
    if (!c) return PXH_EINVAL;
    return guarded(c, [&] {
        PXH_CHECK(c->target == PXH_WFQ || c->target == PXH_WFQFMU, PXH_EINVAL, "no atomistic target set");
        PXH_CHECK(c->atoms.kernel == PXH_GAUSSIAN, PXH_EINVAL, "the plasmon density needs the Gaussian kernel");
        PXH_CHECK(n_points >= 0 && n_points <= INT_MAX - 256, PXH_EINVAL, "bad point count");
        PXH_CHECK(n_points == 0 || (points_bohr && variables && rho), PXH_EINVAL, "null array");
        if (n_points == 0) return;
        const int nv = c->atoms.n_var();
        const size_t np = (size_t)n_points;
        pxh::DevBuf<double> dp;
        pxh::DevBuf<double2> dv, dr, dq, dm;
        dp.upload(points_bohr, 3 * np, c->stream);
        dv.upload(reinterpret_cast<const double2 *>(variables), (size_t)nv, c->stream);
        dr.alloc(np);
        if (rho_q) dq.alloc(np);
        if (rho_mu) dm.alloc(np);
        pxh::plasmon_density(c, n_points, dp.p, dv.p, dr.p, rho_q ? dq.p : nullptr, rho_mu ? dm.p : nullptr);
        PXH_HIP(hipMemcpyAsync(rho, dr.p, sizeof(double2) * np, hipMemcpyDeviceToHost, c->stream));
        if (rho_q) PXH_HIP(hipMemcpyAsync(rho_q, dq.p, sizeof(double2) * np, hipMemcpyDeviceToHost, c->stream));
        if (rho_mu) PXH_HIP(hipMemcpyAsync(rho_mu, dm.p, sizeof(double2) * np, hipMemcpyDeviceToHost, c->stream));
        PXH_HIP(hipStreamSynchronize(c->stream));
    });
}

int pxh_matvec(pxh_ctx *c, double freq_au, int i_freq, int n_vec, const double *x, double *y) {
    if (!c) return PXH_EINVAL;
    return guarded(c, [&] {
        PXH_CHECK(n_vec > 0 && x && y, PXH_EINVAL, "bad vectors");
        PXH_CHECK(c->target == PXH_WFQ || c->target == PXH_WFQFMU || c->target == PXH_BEM, PXH_EINVAL,
                  "no dynamic target set");
        const bool bem = c->target == PXH_BEM;
        const int nv = bem ? c->mesh.n : c->atoms.n_var();
        PXH_CHECK(nv > 0, PXH_EINVAL, "no target set");
        pxh::DevBuf<double2> dx, dy;
        dx.upload(reinterpret_cast<const double2 *>(x), (size_t)nv * n_vec, c->stream);
        dy.alloc((size_t)nv * n_vec);
        std::vector<int> list(n_vec);
        for (int k = 0; k < n_vec; ++k) list[k] = k;
        pxh::DevBuf<int> dl;
        dl.upload(list.data(), n_vec, c->stream);
        if (bem) {
            PXH_CHECK(i_freq >= 0 && i_freq < c->eps_nf, PXH_EINVAL, "i_freq outside the permittivity table");
            std::vector<double2> hl(n_vec, pxh::bem_lambda(c, i_freq));
            pxh::DevBuf<double2> dlam;
            dlam.upload(hl.data(), n_vec, c->stream);
            pxh::apply_bem(c, n_vec, dl.p, dlam.p, dx.p, dy.p, nv, c->opt.matvec_mixed != 0);
        } else {
            std::vector<pxh::SysParam> hp(n_vec, pxh::make_sysparam(c, freq_au, i_freq));
            pxh::DevBuf<pxh::SysParam> dp;
            dp.upload(hp.data(), n_vec, c->stream);
            if (c->opt.matvec_mixed)
                pxh::apply_atoms_fp32(c, n_vec, dl.p, dp.p, dx.p, dy.p, nv);
            else
                pxh::apply_atoms_fp64(c, n_vec, dl.p, dp.p, dx.p, dy.p, nv);
        }
        PXH_HIP(hipMemcpyAsync(y, dy.p, sizeof(double2) * nv * n_vec, hipMemcpyDeviceToHost, c->stream));
        PXH_HIP(hipStreamSynchronize(c->stream));
    });
}

int pxh_constant_matrix(pxh_ctx *c, double *m) {
    if (!c) return PXH_EINVAL;
    return guarded(c, [&] {
        PXH_CHECK(m, PXH_EINVAL, "null output");
        PXH_CHECK(c->target == PXH_WFQ || c->target == PXH_WFQFMU || c->target == PXH_BEM, PXH_EINVAL,
                  "no dynamic target set");
        if (c->target == PXH_BEM) {
            const size_t nn = (size_t)c->mesh.n * c->mesh.n;
            PXH_HIP(hipMemcpyAsync(m, c->mesh.D.p, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
        } else {
            PXH_CHECK(c->atoms.n > 0, PXH_EINVAL, "no target set");
            const size_t nv = c->atoms.n_var();
            pxh::DevBuf<double> dm;
            dm.alloc(nv * nv);
            pxh::constant_matrix_atoms(c, dm.p);
            PXH_HIP(hipMemcpyAsync(m, dm.p, sizeof(double) * nv * nv, hipMemcpyDeviceToHost, c->stream));
            PXH_HIP(hipStreamSynchronize(c->stream));
        }
        PXH_HIP(hipStreamSynchronize(c->stream));
    });
}

int pxh_dynamic_rhs(pxh_ctx *c, double freq_au, double e0, int rhs_form, double *rhs) {
    if (!c) return PXH_EINVAL;
    return guarded(c, [&] {
        PXH_CHECK(rhs, PXH_EINVAL, "null output");
        PXH_CHECK(c->target == PXH_WFQ || c->target == PXH_WFQFMU || c->target == PXH_BEM, PXH_EINVAL,
                  "no dynamic target set");
        PXH_CHECK(rhs_form == PXH_RHS_FIELD || rhs_form == PXH_RHS_POTENTIAL, PXH_EINVAL, "unknown rhs_form");
        const bool bem = c->target == PXH_BEM;
        const int nv = bem ? c->mesh.n : c->atoms.n_var();
        PXH_CHECK(nv > 0, PXH_EINVAL, "no target set");
        pxh::DevBuf<double2> d;
        d.alloc((size_t)nv * 3);
        if (bem)
            pxh::dynamic_rhs_bem(c, e0, d.p);
        else
            pxh::dynamic_rhs_atoms(c, freq_au, e0, rhs_form, d.p);
        PXH_HIP(hipMemcpyAsync(rhs, d.p, sizeof(double2) * nv * 3, hipMemcpyDeviceToHost, c->stream));
        PXH_HIP(hipStreamSynchronize(c->stream));
    });
}

int pxh_last_error(pxh_ctx *c, char *buf, int len) {
    if (!buf || len <= 0) return PXH_EINVAL;
    const std::string &s = c ? c->last_error : g_last_error;
    const int k = std::min<int>(len - 1, (int)s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = '\0';
    return PXH_OK;
}

int pxh_set_option(pxh_ctx *c, const char *key, double value) {
    if (!c || !key) return PXH_EINVAL;
    return guarded(c, [&] {
        const std::string k(key);
        if (k == "kbar_cutoff") {
            c->opt.kbar_cutoff = value != 0.0;
            if (c->atoms.n > 0) pxh::build_kbar(c);
        } else if (k == "batch") {
            PXH_CHECK(value >= 1, PXH_EINVAL, "batch must be >= 1");
            c->opt.batch = (int)value;
        } else if (k == "inner_tol") {
            PXH_CHECK(value > 0 && value < 1, PXH_EINVAL, "inner_tol must be in (0, 1)");
            c->opt.inner_tol = value;
        } else if (k == "splits") {
            c->opt.splits = (int)value;
        } else if (k == "x0_projection") {
            PXH_CHECK(value >= 0 && value <= 16, PXH_EINVAL, "x0_projection must be 0..16");
            c->opt.x0_projection = (int)value;
        } else if (k == "res_batch") {
            PXH_CHECK(value >= 1, PXH_EINVAL, "res_batch must be >= 1");
            c->opt.res_batch = (int)value;
        } else if (k == "res_wait") {
            PXH_CHECK(value >= 0, PXH_EINVAL, "res_wait must be >= 0");
            c->opt.res_wait = (int)value;
        } else if (k == "autotune") {
            c->opt.autotune = value != 0.0;
        } else if (k == "graphs") {
            c->opt.graphs = value != 0.0;
        } else if (k == "spectral") {
            PXH_CHECK(value == 0 || value == 1 || value == 2, PXH_EINVAL, "spectral must be 0, 1 or 2");
            c->opt.spectral = (int)value;
        } else if (k == "dcgs2") {
            c->opt.dcgs2 = value != 0.0;
        } else if (k == "initial_guess") {
            c->opt.initial_guess = value != 0.0;
        } else if (k == "multishift") {
            c->opt.multishift = value != 0.0;
        } else if (k == "matvec_mixed") {
            c->opt.matvec_mixed = value != 0.0;
        } else if (k == "mfma") {
            c->opt.mfma = value != 0.0;
        } else if (k == "bem_gemv") {
            c->opt.bem_gemv = value != 0.0;
        } else if (k == "fp64_mfma") {
            c->opt.fp64_mfma = value != 0.0;
        } else if (k == "mfma_r4") {
            c->opt.mfma_r4 = value != 0.0;
        } else if (k == "mfma_flush") {
            PXH_CHECK(value >= 1 && value <= 1024, PXH_EINVAL, "mfma_flush must be 1..1024");
            c->opt.mfma_flush = (int)value;
        } else if (k == "mfma_levels") {
            PXH_CHECK(value == 1 || value == 2, PXH_EINVAL, "mfma_levels must be 1 or 2");
            c->opt.mfma_levels = (int)value;
        } else if (k == "precond") {
            c->opt.precond = value != 0.0;
        } else if (k == "max_outer") {
            c->opt.max_outer = std::max(1, (int)value);
        } else if (k == "reorth_kappa") {
            PXH_CHECK(value >= 0, PXH_EINVAL, "reorth_kappa must be >= 0");
            c->opt.reorth_kappa = value;
        } else {
            throw Error{PXH_EINVAL, "unknown option: " + k};
        }
    });
}

int pxh_get_stats(pxh_ctx *c, double *stats, int n) {
    if (!c || !stats) return PXH_EINVAL;
    const int st = guarded(c, [&] { pxh::timer_drain(c); });
    if (st != PXH_OK) return st;
    for (int k = 0; k < n && k < pxh::Stats::kN; ++k) stats[k] = c->stats.v[k];
    return PXH_OK;
}

}  // extern "C"
