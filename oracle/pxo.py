"""ctypes wrapper of the FP64 CPU oracle (oracle/build/libpxo.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  Targets are
passed duck-typed (the attributes of plasmonx_amd.model.AtomSystem / Mesh).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libpxo.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.pxo_solve_spectrum.restype = ctypes.c_int
        _lib.pxo_bem_solve_spectrum.restype = ctypes.c_int
        _lib.pxo_zgesv.restype = ctypes.c_int
        _lib.pxo_gmres_otf.restype = ctypes.c_int
        _lib.pxo_gmres_dense.restype = ctypes.c_int
        _lib.pxo_fermi.restype = ctypes.c_double
        _lib.pxo_fermi.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    return _lib


def P(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Atoms(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("kernel", ctypes.c_int), ("n_atoms", ctypes.c_int),
                ("n_types", ctypes.c_int), ("heterogeneous", ctypes.c_int)] + \
               [(k, ctypes.c_void_p) for k in ("coord", "type", "eta", "r_q", "r_mu", "a_ij", "tau", "density",
                                               "fermi_d", "fermi_s", "r0", "nb_counts", "nb_total")]


class OracleAtoms:
    """Holds the arrays alive and the C struct."""

    def __init__(self, s):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.arrs = dict(
            coord=f(s.coord_bohr), type=np.ascontiguousarray(s.type_idx, dtype=np.int32), eta=f(s.eta),
            r_q=f(s.r_q), r_mu=f(s.r_mu), a_ij=f(s.a_ij), tau=f(s.tau), density=f(s.density),
            fermi_d=f(np.asarray(s.fermi_d).T.reshape(-1)), fermi_s=f(np.asarray(s.fermi_s).T.reshape(-1)),
            r0=f(np.asarray(s.r0).T.reshape(-1)), nb_counts=np.ascontiguousarray(s.nb_counts, dtype=np.int32),
            nb_total=np.ascontiguousarray(s.nb_total, dtype=np.int32))
        kern = {"gaussian": 1, "coulomb": 2, "ohno": 3}[s.kernel]
        self.st = Atoms(2 if s.model in ("wfqfmu", "fqfmu") else 1, kern, s.n_atoms, s.n_types, int(s.heterogeneous),
                        *[P(self.arrs[k]) for k in ("coord", "type", "eta", "r_q", "r_mu", "a_ij", "tau", "density",
                                                    "fermi_d", "fermi_s", "r0", "nb_counts", "nb_total")])
        self.n_var = s.n_var
        self.n = s.n_atoms
        self.nt = s.n_types
        self.ref = ctypes.byref(self.st)


def solve_spectrum(s, freq_au, e0, alpha_w=None, rhs_form=1, pol_mask=7, method=0, tol=1e-9, m=1000,
                   want_variables=False):
    lib = load()
    oa = OracleAtoms(s)
    freq = np.ascontiguousarray(freq_au, dtype=np.float64)
    nf = len(freq)
    aw = None if alpha_w is None else np.asfortranarray(np.asarray(alpha_w, dtype=np.complex128))
    var = np.zeros((nf, 3, oa.n_var), dtype=np.complex128) if want_variables else None
    pol = np.zeros((nf, 3, 3), dtype=np.complex128)
    res = np.zeros((20, nf))
    it = np.zeros((nf, 3), dtype=np.int32)
    st = lib.pxo_solve_spectrum(oa.ref, ctypes.c_int(nf), P(freq), P(aw), ctypes.c_double(e0), ctypes.c_int(rhs_form),
                                ctypes.c_int(pol_mask), ctypes.c_int(method), ctypes.c_double(tol), ctypes.c_int(m),
                                P(var), P(pol), P(res), P(it))
    out = {"status": st, "polar": np.transpose(pol, (0, 2, 1)), "results": res.T.copy(), "iters": it}
    if want_variables:
        out["variables"] = var
    return out


def matvec_otf(s, freq, alpha_t, x):
    lib = load()
    oa = OracleAtoms(s)
    at = np.ascontiguousarray(np.asarray(alpha_t, dtype=np.complex128).reshape(-1))
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.zeros_like(x)
    lib.pxo_matvec_otf(oa.ref, ctypes.c_double(freq), P(at), P(x), P(y))
    return y


def induced_field(s, points_bohr, var, i_pol, e0):
    """Field (n_points, 3) complex at points (n_points, 3) bohr for the
    variables var (n_var complex) of the field direction i_pol (0..2)."""
    lib = load()
    oa = OracleAtoms(s)
    pts = np.ascontiguousarray(points_bohr, dtype=np.float64)
    v = np.ascontiguousarray(var, dtype=np.complex128)
    assert v.shape == (oa.n_var,)
    E = np.zeros((pts.shape[0], 3), dtype=np.complex128)
    lib.pxo_induced_field(oa.ref, ctypes.c_int(pts.shape[0]), P(pts), P(v), ctypes.c_int(i_pol), ctypes.c_double(e0),
                          P(E))
    return E


def plasmon_density(s, points_bohr, var):
    """(rho, rho_q, rho_mu), each (n_points,) complex, at points (n_points, 3)
    bohr for the variables var (n_var complex); rho = rho_q + rho_mu."""
    lib = load()
    oa = OracleAtoms(s)
    pts = np.ascontiguousarray(points_bohr, dtype=np.float64)
    v = np.ascontiguousarray(var, dtype=np.complex128)
    assert v.shape == (oa.n_var,)
    out = [np.zeros(pts.shape[0], dtype=np.complex128) for _ in range(3)]
    lib.pxo_plasmon_density(oa.ref, ctypes.c_int(pts.shape[0]), P(pts), P(v), *[P(o) for o in out])
    return tuple(out)


def constant_matrix(s):
    lib = load()
    oa = OracleAtoms(s)
    M = np.zeros((oa.n_var, oa.n_var), order="F")
    lib.pxo_constant_matrix(oa.ref, P(M))
    return M


def dynamic_matrix(s, M, freq, alpha_t):
    lib = load()
    oa = OracleAtoms(s)
    A = np.zeros((oa.n_var, oa.n_var), dtype=np.complex128, order="F")
    at = np.ascontiguousarray(np.asarray(alpha_t, dtype=np.complex128).reshape(-1))
    lib.pxo_dynamic_matrix(oa.ref, P(np.asfortranarray(M)), ctypes.c_double(freq), P(at), P(A))
    return A


def rhs(s, e0, freq=0.0, rhs_form=1):
    lib = load()
    oa = OracleAtoms(s)
    r = np.zeros((3, oa.n_var), dtype=np.complex128)
    lib.pxo_rhs(oa.ref, ctypes.c_double(e0), ctypes.c_int(rhs_form), ctypes.c_double(freq), ctypes.c_int(1), P(r))
    return r


def t_qq(s):
    """The charge-charge block T_qq (n x n) of the target's kernel."""
    lib = load()
    oa = OracleAtoms(s)
    T = np.zeros((oa.n, oa.n), order="F")
    lib.pxo_t_qq(oa.ref, P(T))
    return T


def static_rhs_q(s, e0, rhs_form=1):
    """The static potential V (n, 3) of the external field, 'field' (1) or
    'potential' (2) form."""
    lib = load()
    oa = OracleAtoms(s)
    V = np.zeros((oa.n, 3), order="F")
    lib.pxo_static_rhs_q(oa.ref, ctypes.c_double(e0), ctypes.c_int(rhs_form), P(V))
    return V


def kbar(s):
    lib = load()
    oa = OracleAtoms(s)
    K = np.zeros((oa.n, oa.n), order="F")
    lib.pxo_kbar(oa.ref, P(K))
    return K


def zgesv(A, B):
    lib = load()
    A = np.asfortranarray(A, dtype=np.complex128).copy(order="F")
    B = np.asfortranarray(B, dtype=np.complex128).copy(order="F")
    nrhs = 1 if B.ndim == 1 else B.shape[1]
    info = lib.pxo_zgesv(ctypes.c_int(A.shape[0]), ctypes.c_int(nrhs), P(A), P(B))
    return info, B


def gmres_dense(A, b, tol, m):
    lib = load()
    A = np.asfortranarray(A, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.zeros_like(b)
    it = ctypes.c_int(0)
    hist = np.zeros(m)
    tr = ctypes.c_double(0)
    st = lib.pxo_gmres_dense(ctypes.c_int(len(b)), P(A), P(b), P(x), ctypes.c_double(tol), ctypes.c_int(m),
                             ctypes.byref(it), P(hist), ctypes.byref(tr))
    return st, x, hist[:it.value], tr.value


def gmres_otf(s, freq, alpha_t, b, tol, m):
    lib = load()
    oa = OracleAtoms(s)
    at = np.ascontiguousarray(np.asarray(alpha_t, dtype=np.complex128).reshape(-1))
    b = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.zeros_like(b)
    it = ctypes.c_int(0)
    hist = np.zeros(m)
    tr = ctypes.c_double(0)
    st = lib.pxo_gmres_otf(oa.ref, ctypes.c_double(freq), P(at), P(b), P(x), ctypes.c_double(tol), ctypes.c_int(m),
                           ctypes.byref(it), P(hist), ctypes.byref(tr))
    return st, x, hist[:it.value], tr.value


def etchegoin(metal, freq_au, polar):
    lib = load()
    f = np.ascontiguousarray(freq_au, dtype=np.float64)
    out = np.zeros(len(f), dtype=np.complex128)
    fn = lib.pxo_silver_etchegoin if metal == "ag" else lib.pxo_gold_etchegoin
    fn(ctypes.c_int(int(polar)), ctypes.c_int(len(f)), P(f), P(out))
    return out


class Mesh(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int)] + [(k, ctypes.c_void_p) for k in ("coord", "normal", "tang1", "tang2", "area",
                                                                       "verts")]


class OracleMesh:
    def __init__(self, mesh):
        lib = load()
        n = mesh.n
        self.verts = np.ascontiguousarray(np.asarray(mesh.verts_bohr).reshape(n, 9), dtype=np.float64)
        self.coord = np.zeros((n, 3))
        self.normal = np.zeros((n, 3))
        self.t1 = np.zeros((n, 3))
        self.t2 = np.zeros((n, 3))
        self.area = np.zeros(n)
        lib.pxo_tesserae(ctypes.c_int(n), P(self.verts), ctypes.c_double(mesh.normal_factor), P(self.coord),
                         P(self.normal), P(self.t1), P(self.t2), P(self.area))
        self.st = Mesh(n, P(self.coord), P(self.normal), P(self.t1), P(self.t2), P(self.area), P(self.verts))
        self.ref = ctypes.byref(self.st)
        self.n = n


def bem_d(mesh):
    lib = load()
    om = OracleMesh(mesh)
    n = om.n
    D = np.zeros((n, n), order="F")
    Di = np.zeros((n, n), order="F")
    SA = np.zeros((n, n), order="F")
    lib.pxo_bem_d(om.ref, ctypes.c_int(int(mesh.green == "accurate")), ctypes.c_double(mesh.sphere_r_bohr), P(D),
                  P(Di), P(SA))
    return D, Di, SA


def bem_solve_spectrum(mesh, freq_au, eps_w, e0, method=0, tol=1e-9, m=1000, want_variables=False):
    lib = load()
    om = OracleMesh(mesh)
    freq = np.ascontiguousarray(freq_au, dtype=np.float64)
    nf = len(freq)
    eps = np.ascontiguousarray(eps_w, dtype=np.complex128)
    var = np.zeros((nf, 3, om.n), dtype=np.complex128) if want_variables else None
    pol = np.zeros((nf, 3, 3), dtype=np.complex128)
    res = np.zeros((20, nf))
    it = np.zeros((nf, 3), dtype=np.int32)
    st = lib.pxo_bem_solve_spectrum(om.ref, ctypes.c_int(int(mesh.variant == "iefpcm")),
                                    ctypes.c_int(int(mesh.green == "accurate")), ctypes.c_double(mesh.sphere_r_bohr),
                                    ctypes.c_double(mesh.eps_solvent), ctypes.c_int(nf), P(freq), P(eps),
                                    ctypes.c_double(e0), ctypes.c_int(int(mesh.charge_constraint)),
                                    ctypes.c_int(method), ctypes.c_double(tol), ctypes.c_int(m), P(var), P(pol),
                                    P(res), P(it))
    out = {"status": st, "polar": np.transpose(pol, (0, 2, 1)), "results": res.T.copy(), "iters": it}
    if want_variables:
        out["variables"] = var
    return out


def bem_induced_field(centroids_bohr, charges, points_bohr, i_pol, e0):
    """Field (n_points, 3) complex of the BEM charges at the centroids, zero
    inside the particle."""
    lib = load()
    c = np.ascontiguousarray(centroids_bohr, dtype=np.float64)
    q = np.ascontiguousarray(charges, dtype=np.complex128)
    pts = np.ascontiguousarray(points_bohr, dtype=np.float64)
    assert c.shape == (q.shape[0], 3)
    E = np.zeros((pts.shape[0], 3), dtype=np.complex128)
    lib.pxo_bem_induced_field(ctypes.c_int(q.shape[0]), P(c), P(q), ctypes.c_int(pts.shape[0]), P(pts),
                              ctypes.c_int(i_pol), ctypes.c_double(e0), P(E))
    return E



def static_matrix(s, mol_count):
    """The static FQ / FQFmu matrix (n_var x n_var, unknowns [q, lambda, mu])."""
    lib = load()
    oa = OracleAtoms(s)
    mc = np.ascontiguousarray(mol_count, dtype=np.int32)
    alpha = np.ascontiguousarray(s.arr("alpha"), dtype=np.float64)
    nv = lib.pxo_static_n_var(oa.ref, ctypes.c_int(len(mc)))
    A = np.zeros((nv, nv), order="F")
    lib.pxo_static_matrix(oa.ref, P(alpha), ctypes.c_int(len(mc)), P(mc), P(A))
    return A


def static_solve(s, mol_count, what, e0=0.0, rhs_form=1):
    """Static FQ / FQFmu solve: what 0 = ground state, 1 = static field.
    Returns (variables (n_var, ncol), polar (3, 3) or None, energy)."""
    lib = load()
    oa = OracleAtoms(s)
    mc = np.ascontiguousarray(mol_count, dtype=np.int32)
    chi = np.ascontiguousarray(s.arr("chi"), dtype=np.float64)
    alpha = np.ascontiguousarray(s.arr("alpha"), dtype=np.float64)
    nv = lib.pxo_static_n_var(oa.ref, ctypes.c_int(len(mc)))
    ncol = 1 if what == 0 else 3
    var = np.zeros((nv, ncol), order="F")
    pol = np.zeros((3, 3), order="F")
    en = ctypes.c_double(0.0)
    st = lib.pxo_static_solve(oa.ref, P(chi), P(alpha), ctypes.c_int(len(mc)), P(mc), ctypes.c_int(what),
                              ctypes.c_double(e0), ctypes.c_int(rhs_form), P(var), P(pol), ctypes.byref(en))
    if st != 0:
        raise RuntimeError(f"pxo_static_solve status {st}")
    return var, (pol if what == 1 and s.model != "fqfmu" else None), en.value
