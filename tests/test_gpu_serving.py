"""The request loop (pxh_server_*): requests of different kinds and sizes,
admitted by HBM footprint and run concurrently, give the same results as
each solved alone; a tight budget serialises them without changing them."""
import numpy as np
import pytest

from conftest import load_fixture
from plasmonx_amd import _lib
from plasmonx_amd import permittivity as pm
from plasmonx_amd.model import mesh_from_fixture
from plasmonx_amd.workloads import c2_silver_sphere, sodium_rod

pytestmark = pytest.mark.gpu


def _targets():
    """(context, solve kwargs) for an omegaFQFmu Ag sphere (mixed), an
    omegaFQ Na rod (FP64 GMRES) and a BEM DPCM mesh (mixed, stored D)."""
    out = []
    s, freq = c2_silver_sphere(radius=12.0, n_freq=6)
    c = _lib.Context(0)
    c.set_atoms(s)
    c.set_alpha(s.alpha_table(freq))
    out.append((c, dict(freq_au=freq, e0=1e-4, method=_lib.ITERATIVE_OTF, tol_abs=1e-11 * np.sqrt(s.n_var),
                        krylov_dim=200, precision=_lib.MIXED)))
    na = sodium_rod(30.0, 9.0)
    c = _lib.Context(0)
    c.set_atoms(na)
    f = np.linspace(2.5, 4.0, 5) / 27.211386245981
    out.append((c, dict(freq_au=f, e0=1e-4, pol_mask=1, method=_lib.ITERATIVE_OTF,
                        tol_abs=1e-10 * np.sqrt(na.n_var), krylov_dim=300, precision=_lib.FP64)))
    fx = load_fixture("bem_dpcm_accurate_inversion")
    m = mesh_from_fixture(fx)
    fb = np.linspace(3.0, 4.0, 4) / 27.211386245981
    c = _lib.Context(0)
    c.set_mesh(m)
    c.set_permittivity(pm.silver_etchegoin(fb))
    out.append((c, dict(freq_au=fb, e0=fx["e0"], method=_lib.ITERATIVE, tol_abs=1e-10 * np.sqrt(m.n),
                        krylov_dim=200, precision=_lib.MIXED)))
    return out


@pytest.mark.parametrize("budget", [0, 1])
def test_results_identical_alone_and_in_the_mix(budget):
    targets = _targets()
    alone = [c.solve_spectrum(c.n_var, **kw) for c, kw in targets]
    srv = _lib.Server(0, hbm_budget=budget, max_inflight=3)
    try:
        ids = [srv.submit(c, **kw) for c, kw in targets]
        mixed = [srv.wait(i) for i in ids]
    finally:
        srv.close()
    for a, m in zip(alone, mixed):
        assert np.array_equal(a["polar"], m["polar"])
        assert np.array_equal(a["iters"], m["iters"])
        assert m["latency_ms"] >= m["queue_ms"] >= 0.0
    for c, _ in targets:
        c.close()


def test_resubmission_on_one_context_runs_in_order():
    """Two requests on the same context are never run at once."""
    (c, kw), = _targets()[:1]
    ref = c.solve_spectrum(c.n_var, **kw)
    srv = _lib.Server(0, max_inflight=4)
    try:
        ids = [srv.submit(c, **kw) for _ in range(3)]
        outs = [srv.wait(i) for i in ids]
    finally:
        srv.close()
    for o in outs:
        assert np.array_equal(o["polar"], ref["polar"])
    c.close()
