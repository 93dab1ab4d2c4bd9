"""Multi-shift GMRES for the scalar-shift frequency families (BEM DPCM,
homogeneous omegaFQ): one Krylov basis per field direction serves every
frequency.  Bars: each system's FP64 true residual below tol; polarizabilities
equal to the dense inversion (same library, FP64 LU) within 1e-7; the same
answers with the feature off; far fewer operator products; and a basis too
small to converge hands its systems to the batched solver, warm-started."""
import numpy as np
import pytest

from conftest import load_fixture
from plasmonx_amd import _lib
from plasmonx_amd import permittivity as pm
from plasmonx_amd.model import mesh_from_fixture, system_from_fixture

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _sodium(ctx, case="wfq_sodium_iterative_on_memory", nf=None):
    fx = load_fixture(case)
    s = system_from_fixture(fx)
    ctx.set_option("kbar_cutoff", 0)
    ctx.set_atoms(s)
    return s.n_var, np.array(fx["freq_au"])[:nf], fx["e0"], 1


def _graphene(ctx):
    fx = load_fixture("wfq_graphene_disk_inversion")
    s = system_from_fixture(fx)
    ctx.set_option("kbar_cutoff", 0)
    ctx.set_atoms(s)
    return s.n_var, np.array(fx["freq_au"])[::4], fx["e0"], 3


def _dpcm(ctx):
    fx = load_fixture("bem_dpcm_external_brendel_bormann_ag")
    m = mesh_from_fixture(fx)
    freq = np.linspace(3.0, 4.0, 40) / 27.211386245981
    ctx.set_mesh(m)
    ctx.set_permittivity(pm.silver_etchegoin(freq))
    return m.n, freq, fx["e0"], 7


def _ief(ctx):
    fx = load_fixture("bem_ief_accurate_inversion")
    m = mesh_from_fixture(fx)
    freq = np.linspace(3.0, 4.0, 40) / 27.211386245981
    ctx.set_mesh(m)
    ctx.set_permittivity(pm.silver_etchegoin(freq))
    return m.n, freq, fx["e0"], 7


CASES = {"sodium_450w": _sodium, "graphene_disk": _graphene, "bem_dpcm": _dpcm, "bem_ief": _ief}


@pytest.mark.parametrize("case,method", [(k, m) for k in sorted(CASES) for m in (_lib.ITERATIVE, _lib.ITERATIVE_OTF)
                                         if not (k.startswith("bem") and m == _lib.ITERATIVE_OTF)])
def test_multishift_matches_inversion(gpu_ctx, case, method):
    n, freq, e0, mask = CASES[case](gpu_ctx)
    tol = 1e-11 * np.sqrt(n)
    ref = gpu_ctx.solve_spectrum(n, freq, e0, pol_mask=mask, method=_lib.INVERSION)
    gpu_ctx.set_option("batch", 48)
    gpu_ctx.set_option("multishift", 1)
    got = gpu_ctx.solve_spectrum(n, freq, e0, pol_mask=mask, method=method, tol_abs=tol, krylov_dim=1000,
                                 precision=_lib.MIXED)
    st_ms = gpu_ctx.stats()
    assert np.all(got["true_residual"] < tol)
    assert relerr(got["polar"], ref["polar"]) < 1e-7
    gpu_ctx.set_option("multishift", 0)
    off = gpu_ctx.solve_spectrum(n, freq, e0, pol_mask=mask, method=method, tol_abs=tol, krylov_dim=1000,
                                 precision=_lib.MIXED)
    st_off = gpu_ctx.stats()
    gpu_ctx.set_option("multishift", 1)
    # the batched solver stops on the same residual test; on the weakly
    # damped graphene flake that bounds its polarizability error by ~1e-6
    assert relerr(got["polar"], off["polar"]) < 1e-6
    # one basis for all frequencies: far fewer system-products
    assert st_ms["matvecs"] < 0.5 * st_off["matvecs"], (st_ms["matvecs"], st_off["matvecs"])


def test_small_basis_falls_back_to_batched(gpu_ctx):
    n, freq, e0, mask = _sodium(gpu_ctx, nf=60)
    tol = 1e-11 * np.sqrt(n)
    ref = gpu_ctx.solve_spectrum(n, freq, e0, pol_mask=mask, method=_lib.INVERSION)
    gpu_ctx.set_option("batch", 48)
    got = gpu_ctx.solve_spectrum(n, freq, e0, pol_mask=mask, method=_lib.ITERATIVE_OTF, tol_abs=tol,
                                 krylov_dim=8, precision=_lib.FP64)
    assert np.all(got["true_residual"] < tol)
    assert relerr(got["polar"], ref["polar"]) < 1e-7
    assert np.all(got["iters"][:, 0] > 8)  # the batched solver continued every system


def test_zero_field_direction_has_zero_solution(gpu_ctx):
    """Planar flake, field along z: the right-hand side vanishes, and so does
    the solution, without an Arnoldi process for it."""
    n, freq, e0, _ = _graphene(gpu_ctx)
    got = gpu_ctx.solve_spectrum(n, freq[:5], e0, pol_mask=7, method=_lib.ITERATIVE_OTF, tol_abs=1e-10 * np.sqrt(n),
                                 krylov_dim=500, precision=_lib.FP64, want_variables=True)
    assert np.all(got["variables"][:, 2, :] == 0)
    assert np.all(got["iters"][:, 2] == 0)


def test_ief_with_variables_solves_for_the_charges(gpu_ctx):
    """IEF through the multi-shift path returns polarizabilities only (it
    solves for u = S q); asking for the charges takes the batched solver on
    the IEF matrix itself, with the same polarizabilities."""
    n, freq, e0, mask = _ief(gpu_ctx)
    tol = 1e-11 * np.sqrt(n)
    ref = gpu_ctx.solve_spectrum(n, freq[:6], e0, method=_lib.INVERSION, want_variables=True)
    got = gpu_ctx.solve_spectrum(n, freq[:6], e0, method=_lib.ITERATIVE, tol_abs=tol, krylov_dim=500,
                                 precision=_lib.MIXED, want_variables=True)
    assert relerr(got["polar"], ref["polar"]) < 1e-7
    assert relerr(got["variables"], ref["variables"]) < 1e-7


@pytest.mark.parametrize("case", ["bem_dpcm", "bem_ief"])
def test_graph_replays_equal_direct_launches(gpu_ctx, case):
    """The multi-shift Arnoldi step captured once as a HIP graph and replayed
    for the following steps gives bit-identical spectra, iteration counts and
    residuals to the same steps launched directly; three solves with graphs
    (each replaying the step many times), with direct-launch work and stream
    synchronisations between them."""
    n, freq, e0, mask = CASES[case](gpu_ctx)
    tol = 1e-11 * np.sqrt(n)
    want_var = case == "bem_dpcm"
    gpu_ctx.set_option("batch", 48)
    gpu_ctx.set_option("multishift", 1)

    def solve():
        return gpu_ctx.solve_spectrum(n, freq, e0, pol_mask=mask, method=_lib.ITERATIVE, tol_abs=tol,
                                      krylov_dim=1000, precision=_lib.MIXED, want_variables=want_var)

    gpu_ctx.set_option("graphs", 0)
    direct = solve()
    assert np.all(direct["true_residual"] < tol)
    assert direct["iters"].max() >= 3  # enough steps for replays
    gpu_ctx.set_option("graphs", 1)
    for rep in range(3):
        got = solve()
        for key in ("polar", "iters", "true_residual") + (("variables",) if want_var else ()):
            assert np.array_equal(got[key], direct[key]), (rep, key)
        # direct launches in between: a dense inversion at two frequencies
        gpu_ctx.solve_spectrum(n, freq[:2], e0, pol_mask=mask, method=_lib.INVERSION)
