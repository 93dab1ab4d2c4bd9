"""Host-side models and workload generators: closed forms and library
routines pin them (no GPU)."""
import numpy as np
import pytest
from scipy.special import wofz

from plasmonx_amd import geometry, permittivity as pm
from plasmonx_amd.workloads import c1_sodium_sphere, c2_silver_sphere


def test_faddeeva_matches_scipy_wofz():
    rng = np.random.default_rng(11)
    z = rng.uniform(-8, 8, 400) + 1j * rng.uniform(1e-3, 6, 400)
    assert np.abs(pm.faddeeva(z) - wofz(z)).max() < 1e-10 * np.abs(wofz(z)).max()


def test_brendel_bormann_reduces_to_lorentz_as_broadening_vanishes():
    """With the Gaussian width s -> 0 a Brendel-Bormann oscillator becomes the
    Lorentz term f wp^2 / (w0^2 - w^2 - i g w); the Drude term is
    -f0 wp^2 / (w (w + i g0))."""
    ev = np.linspace(1.5, 2.5, 21)  # away from w0 = 3 eV, where the limit is smooth
    freq = ev / 27.211386245981
    f, w0, g, wp, f0, g0 = 0.5, 3.0, 0.2, 9.0, 0.8, 0.05
    eps = pm.brendel_bormann(freq, 1.0, wp, f0, g0, [(f, w0, g, 1e-4)])
    ref = 1.0 - f0 * wp ** 2 / (ev * (ev + 1j * g0)) + f * wp ** 2 / (w0 ** 2 - ev ** 2 - 1j * g * ev)
    assert np.abs(eps - ref).max() < 1e-6 * np.abs(ref).max()


def test_workload_sizes():
    assert len(c1_sodium_sphere()[0].symbols) == 1745
    s, f = c2_silver_sphere(n_freq=201)
    assert s.n_atoms == 15683 and s.n_var == 62732 and len(f) == 201
    assert geometry.icosphere(100.0, 4, seed=0).shape == (5120, 3, 3)


def test_icosphere_is_closed_and_outward():
    v = geometry.icosphere(10.0, 3, seed=0)
    c = v.mean(axis=1)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    assert np.all((n * c).sum(axis=1) > 0)
    # divergence theorem: sum of (area vector . centroid) / 3 = enclosed volume
    vol = (n * c).sum() / 6.0
    assert vol == pytest.approx(4 / 3 * np.pi * 10.0 ** 3, rel=0.02)


def test_lorentz_drude_interband_limits():
    """A Lorentz oscillator f wp^2 / (w0^2 - w^2 - i g w) is real and equal to
    f wp^2 / w0^2 at w -> 0 and purely imaginary, f wp^2 / (g w0), at
    resonance; the silver table is passive (Im alpha > 0) over the C2
    spectrum and of the Etchegoin interband term's size there."""
    from plasmonx_amd import permittivity as pm
    n0 = pm.AG_DENSITY
    one = [(0.3, 4.0, 0.5)]
    wp = 9.0
    a0 = pm.lorentz_drude_interband(np.array([1e-9 / 27.211386245981]), wp_ev=wp, osc=one)[0]
    assert abs(a0.real - 0.3 * wp ** 2 / 16.0 / (4 * np.pi * n0)) < 1e-12 * abs(a0) and abs(a0.imag) < 1e-9 * abs(a0)
    ar = pm.lorentz_drude_interband(np.array([4.0 / 27.211386245981]), wp_ev=wp, osc=one)[0]
    assert abs(ar.real) < 1e-12 * abs(ar) and abs(ar.imag - 0.3 * wp ** 2 / (0.5 * 4.0) / (4 * np.pi * n0)) < 1e-12 * abs(ar)
    w = np.linspace(2.8, 4.2, 201) / 27.211386245981
    ld, et = pm.lorentz_drude_interband(w), pm.silver_etchegoin(w, polar=True)
    assert np.all(ld.imag > 0)
    assert np.all(np.abs(ld) > 0.5 * np.abs(et)) and np.all(np.abs(ld) < 2.0 * np.abs(et))


def test_alpha_table_round_trip(tmp_path):
    """write_table then read_table (the 'wfqfmu file' route of C2) returns
    the model's values exactly at every frequency of the spectrum."""
    from plasmonx_amd.workloads import c2_alpha_table, c2_silver_sphere
    s, freq = c2_silver_sphere(radius=8.0, n_freq=201)
    aw = c2_alpha_table(freq, str(tmp_path / "ag.csv"))
    assert np.array_equal(aw, s.alpha_table(freq))
