"""The oracle's less used branches against what mathematics fixes on its own:
the Coulomb and Ohno charge-charge kernels (matrix_module.f90:79-94), the
'potential' form of the field right-hand side (rhs_module.f90:52-76) and
the BEM charge constraint (algorithm_module.f90:266-319).

Each check is chosen so that a plausible slip fails it: a missing diagonal,
a wrong eta average, max for min in the potential, a correction along the
wrong vector, a wrong sign of the constraint.
"""
import numpy as np
import pytest

from oracle import pxo
from plasmonx_amd import geometry
from plasmonx_amd import permittivity as pm
from plasmonx_amd.model import TOBOHR, AtomSystem, AtomType, Mesh
from plasmonx_amd.workloads import c1_sodium_sphere, core_shell_sphere, silver_type


def two_type_system(coord, kernel, model="wfq"):
    """Two Na-like types with different eta (a 'hetero' wfq system is not
    solvable, but the T_qq block and the static potential are defined)."""
    s, _ = c1_sodium_sphere()
    na = s.types[0]
    t2 = AtomType(name="X", eta=2.0 * na.eta, a_ij=na.a_ij, tau=na.tau, density=na.density, fermi_d=na.fermi_d,
                  fermi_s=na.fermi_s, r_q=1.5 * na.r_q, r_mu=na.r_mu)
    coord = np.asarray(coord, dtype=np.float64)
    sym = ["Na", "X"] * (len(coord) // 2) + ["Na"] * (len(coord) % 2)
    r0 = np.full((2, 2), s.r0[0, 0])
    return AtomSystem(model=model, kernel=kernel, symbols=sym, coord_bohr=coord, types=[na, t2],
                      fermi_d=np.full((2, 2), s.fermi_d[0, 0]), fermi_s=np.full((2, 2), s.fermi_s[0, 0]), r0=r0)


def na_cluster(kernel, n=120, model="wfq"):
    s, _ = c1_sodium_sphere()
    return AtomSystem(model=model, kernel=kernel, symbols=s.symbols[:n], coord_bohr=s.coord_bohr[:n],
                      types=s.types, fermi_d=s.fermi_d, fermi_s=s.fermi_s, r0=s.r0)


def pair_distances(coord):
    d = coord[:, None, :] - coord[None, :, :]
    return np.sqrt((d ** 2).sum(-1))


def test_coulomb_kernel_is_bare_coulomb_with_eta_diagonal():
    """Coulomb: T_ij = 1/r_ij off the diagonal, T_ii = eta_i."""
    rng = np.random.default_rng(11)
    coord = rng.uniform(-20, 20, size=(9, 3))
    s = two_type_system(coord, "coulomb")
    T = pxo.t_qq(s)
    r = pair_distances(coord)
    off = ~np.eye(9, dtype=bool)
    assert np.abs(T[off] * r[off] - 1.0).max() < 1e-14
    assert np.array_equal(np.diag(T), s.eta[s.type_idx])


def test_ohno_kernel_limits():
    """Ohno: T(r = 0) = eta_bar = (eta_i + eta_j)/2 (so T_ii = eta_i); T is
    symmetric, decreases with r and tends to the bare Coulomb 1/r, with
    r T = 1 - 1/(2 eta_bar^2 r^2) + O(r^-4) far away."""
    s = two_type_system([[0, 0, 0], [0, 0, 1e-9]], "ohno")
    T = pxo.t_qq(s)
    eta = s.eta
    assert T[0, 0] == pytest.approx(eta[0], rel=1e-15)
    assert T[1, 1] == pytest.approx(eta[1], rel=1e-15)
    eb = 0.5 * (eta[0] + eta[1])
    assert T[0, 1] == pytest.approx(eb, rel=1e-12)
    assert T[1, 0] == T[0, 1]
    for r in (50.0, 200.0, 1000.0):
        s = two_type_system([[0, 0, 0], [r, 0, 0]], "ohno")
        t = pxo.t_qq(s)[0, 1]
        assert abs(r * t - (1.0 - 1.0 / (2 * eb * eb * r * r))) < 2.0 / (eb * r) ** 4
    ts = [pxo.t_qq(two_type_system([[0, 0, 0], [r, 0, 0]], "ohno"))[0, 1] for r in (0.5, 1, 2, 4, 8)]
    assert all(a > b for a, b in zip(ts, ts[1:]))


def test_gaussian_kernel_limits():
    """Gaussian: T_ii = eta_i; T -> 1/r for r >> R (erf -> 1)."""
    s = two_type_system([[0, 0, 0], [60.0, 0, 0]], "gaussian")
    T = pxo.t_qq(s)
    assert T[0, 0] == s.eta[0] and T[1, 1] == s.eta[1]
    assert abs(T[0, 1] * 60.0 - 1.0) < 1e-15


@pytest.mark.parametrize("kernel", ["coulomb", "ohno"])
def test_on_the_fly_product_equals_assembled_matrix_other_kernels(kernel):
    """The on-the-fly product with the Coulomb / Ohno kernel against the
    assembled A(omega) x (brute force), and A's charge block against
    K-bar T_qq built from the checked T_qq above."""
    s = na_cluster(kernel, 90)
    freq = 3.0 / 27.211386245981
    M = pxo.constant_matrix(s)
    K = pxo.kbar(s)
    T = pxo.t_qq(s)
    assert np.abs(M - K @ T).max() < 1e-12 * np.abs(M).max()
    A = pxo.dynamic_matrix(s, M, freq, np.ones(1))
    rng = np.random.default_rng(5)
    x = rng.normal(size=s.n_var) + 1j * rng.normal(size=s.n_var)
    y = pxo.matvec_otf(s, freq, np.ones(1), x)
    ref = A @ x
    assert np.linalg.norm(y - ref) < 1e-12 * np.linalg.norm(ref)


def test_potential_form_static_rhs():
    """'potential': V_k(i) = -E0 |x_k(i) - min_j x_k(j)|: zero at the lowest
    atom along each axis, non-positive, and the field form plus a constant."""
    s = na_cluster("gaussian", 150)
    e0 = 1e-4
    Vp = pxo.static_rhs_q(s, e0, 2)
    Vf = pxo.static_rhs_q(s, e0, 1)
    assert np.abs(Vf + e0 * s.coord_bohr).max() == 0.0
    for k in range(3):
        assert Vp[:, k].max() == 0.0
        assert Vp[np.argmin(s.coord_bohr[:, k]), k] == 0.0
        shift = Vp[:, k] - Vf[:, k]
        assert np.ptp(shift) < 1e-12 * np.abs(Vf[:, k]).max()


@pytest.mark.parametrize("which", ["wfq", "wfqfmu", "alloy"])
def test_potential_form_dynamic_rhs_equals_field_form(which):
    """K-bar 1 = 0 (and H-bar 1 = 0 for alloys), so the two static forms,
    which differ by a constant potential, give the same dynamic right-hand
    side."""
    if which == "alloy":
        s = core_shell_sphere(4.0, 7.0)
    else:
        s = na_cluster("gaussian", 150, model=which)
        if which == "wfqfmu":
            s = AtomSystem(model="wfqfmu", kernel="gaussian", symbols=["Ag"] * 150,
                           coord_bohr=geometry.fcc_sphere(4.09, 9.0)[:150] * TOBOHR, types=[silver_type()],
                           fermi_d=np.array([[12.0]]), fermi_s=np.array([[1.1]]),
                           r0=np.array([[2.885 * TOBOHR]]))
    freq = 3.1 / 27.211386245981
    f = pxo.rhs(s, 1e-4, freq, rhs_form=1)
    p = pxo.rhs(s, 1e-4, freq, rhs_form=2)
    assert np.abs(p - f).max() < 1e-12 * np.abs(f).max()


def constraint_mesh(variant):
    """An egg-shaped surface (no inversion symmetry, so the unconstrained
    charges of a field direction do not sum to zero by symmetry); each vertex
    moves as a function of its position, so shared vertices stay shared."""
    v = geometry.icosphere(20.0, 2, seed=3)
    v = v + 0.15 * v ** 2 / 20.0
    return Mesh(verts_bohr=v * TOBOHR, variant=variant, green="approximate", sphere_r_bohr=20.0 * TOBOHR,
                charge_constraint=True)


@pytest.mark.parametrize("variant", ["dpcm", "iefpcm"])
def test_charge_constraint_closed_form(variant):
    """With the constraint the charges of each field direction sum to zero,
    and they differ from the unconstrained solution by a multiple of
    A(omega)^-1 1 (solved here by numpy from the oracle's checked operator)."""
    import dataclasses
    m = constraint_mesh(variant)
    freq = np.array([3.3, 3.6]) / 27.211386245981
    eps = pm.silver_etchegoin(freq)
    con = pxo.bem_solve_spectrum(m, freq, eps, 1e-4, method=0, want_variables=True)
    free = pxo.bem_solve_spectrum(dataclasses.replace(m, charge_constraint=False), freq, eps, 1e-4, method=0,
                                  want_variables=True)
    D, Di, SA = pxo.bem_d(m)
    for f in range(2):
        lam = 2 * np.pi * (m.eps_solvent + eps[f]) / (m.eps_solvent - eps[f])
        A = D + lam * np.eye(m.n)
        if variant == "iefpcm":
            A = -A @ SA
        q1 = np.linalg.solve(A, np.ones(m.n))
        for k in range(3):
            qc, q = con["variables"][f, k], free["variables"][f, k]
            assert abs(qc.sum()) < 1e-12 * np.abs(q).sum()
            dq = qc - q
            c = np.vdot(q1, dq) / np.vdot(q1, q1)
            assert np.linalg.norm(dq - c * q1) < 1e-10 * np.linalg.norm(q)
            # and the correction is not empty: the free charges do not sum to 0
            assert abs(q.sum()) > 1e-6 * np.abs(q).sum()
