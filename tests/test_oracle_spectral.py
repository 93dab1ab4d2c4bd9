"""The premises of the spectral route of the homogeneous omegaFQ inversion,
on the oracle's own matrices (CPU): K-bar symmetric, T_qq symmetric positive
definite (Cholesky T = L L^T), the eigenvalues of S = L^T K L (those of
K-bar T) non-positive, and x(w) = L^-T Q diag(1 / (lambda - z(w))) Q^T L^T b
(S = Q diag(lambda) Q^T) equal to the oracle's LU solve of (K T - z(w)) x = b
(wfq_module.f90:214-254, algorithm_module.f90:510-541), the steps the library
takes on the device.  Also the case the route must refuse: a Coulomb kernel
whose T is indefinite."""
import dataclasses

import numpy as np
import scipy.linalg as sl

from oracle import pxo
from plasmonx_amd.model import AtomSystem, system_from_fixture
from plasmonx_amd.workloads import c1_sodium_sphere
from conftest import load_fixture


def eigen_route(s, freq, e0):
    T, K = pxo.t_qq(s), pxo.kbar(s)
    assert np.array_equal(K, K.T) and np.array_equal(T, T.T)
    L = np.linalg.cholesky(T)
    S = L.T @ K @ L
    lam, Q = np.linalg.eigh(0.5 * (S + S.T))
    assert lam.max() < 1e-10 * np.abs(lam).max()  # K-bar is negative semidefinite
    b = np.asarray(pxo.rhs(s, e0, freq[0]))
    assert np.abs(b.imag).max() == 0.0
    C = Q.T @ (L.T @ b.real.T)
    t = s.types[0]
    out = []
    for w in freq:
        z = -(w / (2.0 * t.density * t.tau)) * (w * t.tau + 1j)
        out.append(sl.solve_triangular(L, Q @ (C / (lam - z)[:, None]), trans="T", lower=True))
    return np.array(out)  # [freq][atom][field]


def test_eigen_route_equals_the_oracle_lu():
    fx = load_fixture("wfq_graphene_disk_inversion")
    for s, freq, e0 in [(system_from_fixture(fx), np.array(fx["freq_au"])[:6], fx["e0"]),
                        (load_sodium(), None, 1e-4)]:
        if freq is None:
            freq = c1_sodium_sphere()[1][[0, 30, 60]]
        x = eigen_route(s, freq, e0)
        ref = pxo.solve_spectrum(s, freq, e0, method=0, want_variables=True)
        v = np.asarray(ref["variables"])  # [freq][field][n_var]
        got = np.transpose(x, (0, 2, 1))
        assert np.max(np.abs(got - v)) / np.max(np.abs(v)) < 1e-10


def load_sodium():
    s, _ = c1_sodium_sphere()
    n = 400
    return AtomSystem(model="wfq", kernel="gaussian", symbols=s.symbols[:n], coord_bohr=s.coord_bohr[:n],
                      types=s.types, fermi_d=s.fermi_d, fermi_s=s.fermi_s, r0=s.r0)


def test_coulomb_kernel_with_small_eta_is_indefinite():
    s, _ = c1_sodium_sphere()
    n = 300
    t = dataclasses.replace(s.types[0], eta=0.05)
    c = AtomSystem(model="wfq", kernel="coulomb", symbols=s.symbols[:n], coord_bohr=s.coord_bohr[:n], types=[t],
                   fermi_d=s.fermi_d, fermi_s=s.fermi_s, r0=s.r0)
    assert np.linalg.eigvalsh(pxo.t_qq(c)).min() < 0.0
