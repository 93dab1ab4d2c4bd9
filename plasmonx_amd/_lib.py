"""ctypes binding of libpxh (include/pxh.h).

The HIP library is the only compute path: importing this module does not
touch the GPU, but every call that computes goes through libpxh.so and
raises PxhError when the library, the device or the call fails.  There is no
CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpxh.so")

PXH_OK, PXH_EINVAL, PXH_ENOCONV, PXH_ESINGULAR, PXH_ENOMEM, PXH_EHIP, PXH_ENODEV = range(7)
WFQ, WFQFMU, BEM, FQ, FQFMU = 1, 2, 3, 4, 5
ENERGY, STATIC_FIELD = 0, 1
GAUSSIAN, COULOMB, OHNO = 1, 2, 3
INVERSION, ITERATIVE, ITERATIVE_OTF = 0, 1, 2
FP64, MIXED = 0, 1
RHS_FIELD, RHS_POTENTIAL = 1, 2
DPCM, IEFPCM = 1, 2
APPROXIMATE, ACCURATE = 1, 2

EXPORTS = [
    "pxh_ctx_create", "pxh_ctx_destroy", "pxh_set_atoms", "pxh_set_alpha", "pxh_set_mesh",
    "pxh_set_permittivity", "pxh_solve_spectrum", "pxh_matvec", "pxh_constant_matrix",
    "pxh_dynamic_rhs", "pxh_last_error", "pxh_set_option", "pxh_get_stats",
    "pxh_server_create", "pxh_server_submit", "pxh_server_wait", "pxh_server_last_error", "pxh_server_destroy",
    "pxh_induced_field",
    "pxh_plasmon_density",
    "pxh_bem_induced_field",
    "pxh_matvec_noalloc",
    "pxh_set_static", "pxh_solve_static",
    "pxh_solve_group",
]


class PxhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libpxh status {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libpxh.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PxhError(PXH_ENODEV, f"{LIB_PATH} not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    D = ctypes.c_double
    I = ctypes.c_int
    lib.pxh_ctx_create.argtypes = [I, ctypes.c_size_t, ctypes.POINTER(P)]
    lib.pxh_ctx_destroy.argtypes = [P]
    lib.pxh_set_atoms.argtypes = [P, I, I, I, P, P, I, P, P, P, P, P, P, P, P, P, P, P]
    lib.pxh_set_alpha.argtypes = [P, I, I, P]
    lib.pxh_set_mesh.argtypes = [P, I, I, I, P, D, D, D, I]
    lib.pxh_set_permittivity.argtypes = [P, I, P]
    lib.pxh_solve_spectrum.argtypes = [P, I, P, D, I, I, I, D, I, I, P, P, P, P, P]
    lib.pxh_matvec.argtypes = [P, D, I, I, P, P]
    lib.pxh_matvec_noalloc.argtypes = [P, D, I, I, P, P]
    lib.pxh_set_static.argtypes = [P, I, P, P, I, P]
    lib.pxh_solve_group.argtypes = [I, P, I, P, D, I, I, P, P]
    lib.pxh_solve_static.argtypes = [P, I, D, I, P, P, P]
    lib.pxh_induced_field.argtypes = [P, I, P, P, I, D, P]
    lib.pxh_plasmon_density.argtypes = [P, I, P, P, P, P, P]
    lib.pxh_bem_induced_field.argtypes = [P, I, P, P, I, P, I, D, P]
    lib.pxh_constant_matrix.argtypes = [P, P]
    lib.pxh_dynamic_rhs.argtypes = [P, D, D, I, P]
    lib.pxh_last_error.argtypes = [P, ctypes.c_char_p, I]
    lib.pxh_set_option.argtypes = [P, ctypes.c_char_p, D]
    lib.pxh_get_stats.argtypes = [P, P, I]
    lib.pxh_server_create.argtypes = [I, ctypes.c_size_t, I, ctypes.POINTER(P)]
    lib.pxh_server_submit.argtypes = [P, P, I, P, D, I, I, I, D, I, I, ctypes.POINTER(ctypes.c_int64)]
    lib.pxh_server_wait.argtypes = [P, ctypes.c_int64, P, P, P, P, P]
    lib.pxh_server_last_error.argtypes = [P, ctypes.c_char_p, I]
    lib.pxh_server_destroy.argtypes = [P]
    for name in EXPORTS:
        getattr(lib, name).restype = I
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One libpxh context (one HIP device, resident geometry)."""

    def __init__(self, device=0, hbm_budget=0):
        self.lib = load()
        self._h = ctypes.c_void_p()
        self._check(self.lib.pxh_ctx_create(device, hbm_budget, ctypes.byref(self._h)), null=True)

    def _check(self, st, null=False):
        if st != PXH_OK:
            buf = ctypes.create_string_buffer(512)
            self.lib.pxh_last_error(None if null else self._h, buf, 512)
            raise PxhError(st, buf.value.decode())

    def close(self):
        if self._h:
            self.lib.pxh_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- inputs
    def set_atoms(self, sysm):
        """sysm: plasmonx_amd.model.AtomSystem."""
        c = np.ascontiguousarray(sysm.coord_bohr, dtype=np.float64)
        t = np.ascontiguousarray(sysm.type_idx + 1, dtype=np.int32)
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        fd = f(np.asarray(sysm.fermi_d).T.reshape(-1))  # column-major nt x nt
        fs = f(np.asarray(sysm.fermi_s).T.reshape(-1))
        r0 = f(np.asarray(sysm.r0).T.reshape(-1))
        nbc = np.ascontiguousarray(sysm.nb_counts, dtype=np.int32)  # (n, nt) row-major == nt x n col-major
        nbt = np.ascontiguousarray(sysm.nb_total, dtype=np.int32)
        arrs = [f(sysm.eta), f(sysm.r_q), f(sysm.r_mu), f(sysm.a_ij), fd, fs, r0, f(sysm.tau), f(sysm.density)]
        self._keep = (c, t, nbc, nbt, arrs)
        st = self.lib.pxh_set_atoms(self._h, sysm.model_id, sysm.kernel_id, sysm.n_atoms, _ptr(c), _ptr(t),
                                    sysm.n_types, _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]), _ptr(arrs[3]),
                                    _ptr(fd), _ptr(fs), _ptr(r0), _ptr(arrs[7]), _ptr(arrs[8]), _ptr(nbc),
                                    _ptr(nbt))
        self._check(st)
        self.n_var = sysm.n_var

    def set_static(self, sysm, mol_count):
        """Static FQ / FQFMu data of the current target (set_atoms of a model
        'fq' / 'fqfmu'): chi and alpha of its types, atoms per molecule."""
        chi = np.ascontiguousarray(sysm.arr("chi"), dtype=np.float64)
        alpha = np.ascontiguousarray(sysm.arr("alpha"), dtype=np.float64)
        mc = np.ascontiguousarray(mol_count, dtype=np.int32)
        self._check(self.lib.pxh_set_static(self._h, sysm.n_types, _ptr(chi), _ptr(alpha), mc.shape[0], _ptr(mc)))
        self.n_static = sysm.n_atoms + mc.shape[0] + (3 * sysm.n_atoms if sysm.model == "fqfmu" else 0)

    def solve_static(self, what, e0=0.0, rhs_form=RHS_FIELD, want_polar=True):
        """Static solve: what ENERGY or STATIC_FIELD.  Returns a dict with
        'variables' (n_var, ncol), 'energy' and, for FQ static field,
        'polar' (3, 3)."""
        ncol = 1 if what == ENERGY else 3
        var = np.zeros((ncol, self.n_static), dtype=np.float64)  # column-major n_var x ncol
        pol = np.zeros((3, 3), dtype=np.float64) if (want_polar and what == STATIC_FIELD) else None
        en = np.zeros(1)
        self._check(self.lib.pxh_solve_static(self._h, int(what), float(e0), int(rhs_form), _ptr(var), _ptr(pol),
                                              _ptr(en)))
        out = {"variables": var.T.copy(), "energy": float(en[0])}
        if pol is not None:
            out["polar"] = pol.T.copy()  # [a][b] = alpha_ab (column-major on the C side)
        return out

    def set_alpha(self, alpha_w):
        """alpha_w: complex (n_types, n_freq)."""
        a = np.asfortranarray(np.asarray(alpha_w, dtype=np.complex128))
        self._check(self.lib.pxh_set_alpha(self._h, a.shape[0], a.shape[1], _ptr(a)))

    def set_mesh(self, mesh):
        v = np.ascontiguousarray(mesh.verts_bohr, dtype=np.float64)
        self._check(self.lib.pxh_set_mesh(self._h, mesh.variant_id, mesh.green_id, mesh.n, _ptr(v),
                                          float(mesh.normal_factor), float(mesh.sphere_r_bohr),
                                          float(mesh.eps_solvent), int(mesh.charge_constraint)))
        self.n_var = mesh.n

    def set_permittivity(self, eps_w):
        e = np.ascontiguousarray(np.asarray(eps_w, dtype=np.complex128))
        self._check(self.lib.pxh_set_permittivity(self._h, e.shape[0], _ptr(e)))

    def set_option(self, key, value):
        self._check(self.lib.pxh_set_option(self._h, key.encode(), float(value)))

    # ---- compute
    def solve_spectrum(self, n_var, freq_au, e0, rhs_form=RHS_FIELD, pol_mask=7, method=ITERATIVE_OTF, tol_abs=1e-8,
                       krylov_dim=1000, precision=FP64, want_variables=False, x0=None):
        """x0: starting iterates [n_freq][3][n_var] for the iterative methods
        (sets the "initial_guess" option for this call); the solutions are
        returned as "variables"."""
        freq = np.ascontiguousarray(freq_au, dtype=np.float64)
        nf = freq.shape[0]
        var = np.zeros((nf, 3, n_var), dtype=np.complex128) if want_variables else None
        if x0 is not None:
            var = np.array(x0, dtype=np.complex128, order="C").reshape(nf, 3, n_var)
            want_variables = True
            self.set_option("initial_guess", 1)
        polar = np.zeros((nf, 3, 3), dtype=np.complex128)  # [f][b][a] = alpha_ab
        res = np.zeros((20, nf), dtype=np.float64)  # column-major n_freq x 20
        iters = np.zeros((nf, 3), dtype=np.int32)
        tres = np.zeros((nf, 3), dtype=np.float64)
        try:
            st = self.lib.pxh_solve_spectrum(self._h, nf, _ptr(freq), float(e0), rhs_form, pol_mask, method,
                                             float(tol_abs), int(krylov_dim), precision, _ptr(var), _ptr(polar),
                                             _ptr(res), _ptr(iters), _ptr(tres))
        finally:
            if x0 is not None:
                self.set_option("initial_guess", 0)
        self._check(st)
        out = {"polar": np.transpose(polar, (0, 2, 1)), "results": res.T.copy(), "iters": iters,
               "true_residual": tres}
        if want_variables:
            out["variables"] = var
        return out

    def matvec(self, freq_au, i_freq, x):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.complex128)
        y = np.zeros_like(x)
        self._check(self.lib.pxh_matvec(self._h, float(freq_au), int(i_freq), x.shape[0], _ptr(x), _ptr(y)))
        return y

    def induced_field(self, points_bohr, variables, i_pol, e0):
        """Field (n_points, 3) complex at points (n_points, 3) bohr for the
        variables (n_var complex) of field direction i_pol (0, 1, 2)."""
        pts = np.ascontiguousarray(points_bohr, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(variables, dtype=np.complex128).reshape(-1)
        e = np.zeros((pts.shape[0], 3), dtype=np.complex128)
        self._check(self.lib.pxh_induced_field(self._h, pts.shape[0], _ptr(pts), _ptr(v), int(i_pol), float(e0),
                                               _ptr(e)))
        return e

    def bem_induced_field(self, centroids_bohr, charges, points_bohr, i_pol, e0):
        """Field (n_points, 3) complex of BEM charges at the tessera centroids
        (n_tess, 3) bohr, zero inside the particle."""
        c = np.ascontiguousarray(centroids_bohr, dtype=np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(charges, dtype=np.complex128).reshape(-1)
        if q.shape[0] != c.shape[0]:
            raise ValueError("one charge per centroid")
        pts = np.ascontiguousarray(points_bohr, dtype=np.float64).reshape(-1, 3)
        e = np.zeros((pts.shape[0], 3), dtype=np.complex128)
        self._check(self.lib.pxh_bem_induced_field(self._h, c.shape[0], _ptr(c), _ptr(q), pts.shape[0], _ptr(pts),
                                                   int(i_pol), float(e0), _ptr(e)))
        return e

    def plasmon_density(self, points_bohr, variables, parts=False):
        """Plasmon density (n_points,) complex at points (n_points, 3) bohr for
        the variables (n_var complex); with parts=True, (rho, rho_q, rho_mu)."""
        pts = np.ascontiguousarray(points_bohr, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(variables, dtype=np.complex128).reshape(-1)
        out = [np.zeros(pts.shape[0], dtype=np.complex128) for _ in range(3 if parts else 1)]
        extra = [_ptr(o) for o in out[1:]] + [None] * (3 - len(out))
        self._check(self.lib.pxh_plasmon_density(self._h, pts.shape[0], _ptr(pts), _ptr(v), _ptr(out[0]), *extra))
        return tuple(out) if parts else out[0]

    def constant_matrix(self, n_var):
        m = np.zeros((n_var, n_var), dtype=np.float64, order="F")
        self._check(self.lib.pxh_constant_matrix(self._h, _ptr(m)))
        return m

    def dynamic_rhs(self, n_var, freq_au, e0, rhs_form=RHS_FIELD):
        r = np.zeros((3, n_var), dtype=np.complex128)
        self._check(self.lib.pxh_dynamic_rhs(self._h, float(freq_au), float(e0), rhs_form, _ptr(r)))
        return r

    def stats(self):
        s = np.zeros(32)
        self._check(self.lib.pxh_get_stats(self._h, _ptr(s), 32))
        return {"matvecs": s[0], "fp32_products": s[1], "fp64_products": s[2], "sweep_ms_fp32": s[3],
                "sweep_ms_fp64": s[4], "wall_ms": s[5], "pair_evaluations": s[6], "outer": s[7],
                "sweep_launches_fp32": s[8], "sweep_launches_fp64": s[9], "sweep_flops_fp32": s[10],
                "sweep_flops_fp64": s[11], "mfma_kernel_ms": s[12], "mfma_launches": s[13],
                "latency_p50_ms": s[14], "latency_p99_ms": s[15], "bem_gemm_ms": s[16], "bem_gemm_launches": s[17],
                "bem_gemm_flops": s[18], "lu_update_ms": s[19], "lu_update_launches": s[20],
                "lu_update_flops": s[21], "bem_gemm_bytes": s[22], "mfma_splits": s[23],
                "spectral_targets": s[24], "spectral_ms": s[25], "spectral_flops": s[26],
                "spectral_bytes": s[27]}


def solve_group(contexts, freq_au, e0, rhs_form=RHS_FIELD, pol_mask=7):
    """Dense inversion of the targets of many contexts at once
    (pxh_solve_group): per context a dict with 'polar' (n_freq, 3, 3) and
    'results' (n_freq, 20)."""
    lib = load()
    freq = np.ascontiguousarray(freq_au, dtype=np.float64)
    nf, nc = freq.shape[0], len(contexts)
    handles = (ctypes.c_void_p * nc)(*[c._h.value for c in contexts])
    polar = np.zeros((nc, nf, 3, 3), dtype=np.complex128)  # [t][f][b][a]
    res = np.zeros((nc, 20, nf), dtype=np.float64)
    st = lib.pxh_solve_group(nc, ctypes.cast(handles, ctypes.c_void_p), nf, _ptr(freq), float(e0), rhs_form,
                             pol_mask, _ptr(polar), _ptr(res))
    contexts[0]._check(st)
    return [{"polar": np.transpose(polar[t], (0, 2, 1)), "results": res[t].T.copy()} for t in range(nc)]


class Server:
    """The request loop of libpxh (pxh_server_*): solves on prepared contexts,
    admitted by HBM footprint and run concurrently on one device."""

    def __init__(self, device=0, hbm_budget=0, max_inflight=4):
        self.lib = load()
        self._h = ctypes.c_void_p()
        st = self.lib.pxh_server_create(int(device), int(hbm_budget), int(max_inflight), ctypes.byref(self._h))
        if st != PXH_OK:
            raise PxhError(st, "pxh_server_create failed")
        self._pending = {}

    def submit(self, ctx, freq_au, e0, rhs_form=RHS_FIELD, pol_mask=7, method=ITERATIVE_OTF, tol_abs=1e-8,
               krylov_dim=1000, precision=FP64):
        freq = np.ascontiguousarray(freq_au, dtype=np.float64)
        rid = ctypes.c_int64()
        st = self.lib.pxh_server_submit(self._h, ctx._h, freq.shape[0], _ptr(freq), float(e0), rhs_form, pol_mask,
                                        method, float(tol_abs), int(krylov_dim), precision, ctypes.byref(rid))
        if st != PXH_OK:
            raise PxhError(st, "pxh_server_submit failed")
        self._pending[rid.value] = freq.shape[0]
        return rid.value

    def wait(self, rid):
        nf = self._pending.pop(rid)
        polar = np.zeros((nf, 3, 3), dtype=np.complex128)
        res = np.zeros((20, nf), dtype=np.float64)
        iters = np.zeros((nf, 3), dtype=np.int32)
        tres = np.zeros((nf, 3), dtype=np.float64)
        times = np.zeros(2, dtype=np.float64)
        st = self.lib.pxh_server_wait(self._h, ctypes.c_int64(rid), _ptr(polar), _ptr(res), _ptr(iters), _ptr(tres),
                                      _ptr(times))
        if st != PXH_OK:
            buf = ctypes.create_string_buffer(512)
            self.lib.pxh_server_last_error(self._h, buf, 512)
            raise PxhError(st, buf.value.decode(errors="replace"))
        return {"polar": np.transpose(polar, (0, 2, 1)), "results": res.T.copy(), "iters": iters,
                "true_residual": tres, "queue_ms": times[0], "latency_ms": times[1]}

    def close(self):
        if self._h:
            self.lib.pxh_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
