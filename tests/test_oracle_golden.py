"""The FP64 oracle against the reference's regression logs (the golden
fixtures under tests/golden, written by tests/golden/make_fixtures.py from
the reference's YAML inputs and reference/*.log outputs).

Tolerances follow the printed precision: the 20 result columns are printed
with 6 significant digits (E14.6) -> rtol 1e-5; RHS / charges / dipoles with
3 (E10.3) -> rtol 5e-3; matrices with 4 (E11.4) -> rtol 5e-4; runtest itself
compares at 1e-3.
"""
import numpy as np
import pytest

from conftest import all_cases, load_fixture, method_id, rel_close
from oracle import pxo
from plasmonx_amd import permittivity as pm
from plasmonx_amd.model import mesh_from_fixture, system_from_fixture

ATOM_CASES = all_cases("atoms")
BEM_CASES = all_cases("bem")


def oracle_atoms(fx, nf=None, want_variables=False):
    s = system_from_fixture(fx)
    freq = np.array(fx["freq_au"])[:nf]
    aw = s.alpha_table(freq) if s.model == "wfqfmu" else None
    r = pxo.solve_spectrum(s, freq, fx["e0"], alpha_w=aw, method=method_id(fx["method"]),
                           tol=fx["tol"] * np.sqrt(s.n_var), m=fx["gmres_dim"], want_variables=want_variables)
    return s, r


@pytest.mark.parametrize("case", ATOM_CASES)
def test_results_match_log(case):
    fx = load_fixture(case)
    nf = min(len(fx["freq_au"]), 60)
    s, r = oracle_atoms(fx, nf)
    assert r["status"] == 0
    exp = np.array(fx["expected"]["results"])[:nf]
    got = r["results"]
    # polarizabilities and absorption/extinction at 6 printed digits
    assert rel_close(got[:, :12], exp[:, :12], 1e-5, 1e-6).all(), np.max(np.abs(got[:, :12] - exp[:, :12]))
    assert rel_close(got[:, 12:], exp[:, 12:], 1e-5, 1e-5).all()


@pytest.mark.parametrize("case", [c for c in ATOM_CASES if "rhs" in load_fixture(c)["expected"]])
def test_verbose_blocks_match_log(case):
    fx = load_fixture(case)
    s, r = oracle_atoms(fx, 1, want_variables=True)
    ex = fx["expected"]
    n = s.n_atoms
    rhs = pxo.rhs(s, fx["e0"], fx["freq_au"][0])
    got_rhs = np.stack([rhs.real, rhs.imag], axis=-1).transpose(1, 0, 2).reshape(s.n_var, 6)
    assert rel_close(got_rhs, np.array(ex["rhs"]), 5e-3, 2e-3).all()
    var = r["variables"][0]  # (3, n_var)
    q = np.stack([var[:, :n].real, var[:, :n].imag], -1).transpose(1, 0, 2).reshape(n, 6)
    assert rel_close(q, np.array(ex["charges"]), 5e-3, 2e-3).all()
    if "dipoles" in ex:
        mu = np.stack([var[:, n:].real, var[:, n:].imag], -1).transpose(1, 0, 2).reshape(3 * n, 6)
        assert rel_close(mu, np.array(ex["dipoles"]), 5e-3, 2e-3).all()
    # charge conservation (ChErr): ~1e-14 after inversion, at the GMRES
    # threshold after an iterative solve
    cherr = np.max(np.abs(var[:, :n].sum(axis=1)))
    if fx["method"] == "inversion":
        assert cherr < 1e-9 * np.max(np.abs(var[:, :n]))
    else:
        assert cherr < 10 * fx["tol"] * np.sqrt(s.n_var)
    assert np.max(np.abs(np.array(ex["cherr"]))) < 1e-6 * np.max(np.abs(np.array(ex["charges"])))
    pol = r["polar"][0]
    assert rel_close(pol.real, np.array(ex["polar_re"]), 5e-4, 1e-6).all()
    assert rel_close(pol.imag, np.array(ex["polar_im"]), 5e-4, 1e-6).all()


@pytest.mark.parametrize("case", [c for c in ATOM_CASES if "constant" in load_fixture(c)["expected"]])
def test_printed_matrices_match_log(case):
    fx = load_fixture(case)
    s = system_from_fixture(fx)
    ex = fx["expected"]
    M = pxo.constant_matrix(s)
    freq = fx["freq_au"][0]
    aw = s.alpha_table(np.array([freq]))[:, 0]
    A = pxo.dynamic_matrix(s, M, freq, aw)
    if not s.heterogeneous:
        smp = np.array(ex["constant"])
        got = M[smp[:, 0].astype(int), smp[:, 1].astype(int)]
        assert rel_close(got, smp[:, 2], 5e-4, 1e-5).all()
    for key, part in (("dyn_re", A.real), ("dyn_im", A.imag)):
        if key in ex:
            smp = np.array(ex[key])
            got = part[smp[:, 0].astype(int), smp[:, 1].astype(int)]
            assert rel_close(got, smp[:, 2], 5e-4, 1e-5).all()


@pytest.mark.parametrize("case", [c for c in ATOM_CASES if "gmres" in load_fixture(c)["expected"]])
def test_gmres_history_matches_log(case):
    fx = load_fixture(case)
    s = system_from_fixture(fx)
    tol = fx["tol"] * np.sqrt(s.n_var)
    freqs = np.array(fx["freq_au"])
    hist = fx["expected"]["gmres"][:6]
    for h in hist:
        f = int(np.argmin(np.abs(freqs * 27.211386245981 - h["ev"])))
        assert abs(h["threshold"] - tol) <= 1e-4 * tol
        aw = s.alpha_table(freqs[f:f + 1])[:, 0]
        b = pxo.rhs(s, fx["e0"], freqs[f])["XYZ".index(h["pol"])]
        if fx["method"] == "iterative on the fly":
            st, x, got, tr = pxo.gmres_otf(s, freqs[f], aw, b, tol, fx["gmres_dim"])
        else:
            A = pxo.dynamic_matrix(s, pxo.constant_matrix(s), freqs[f], aw)
            st, x, got, tr = pxo.gmres_dense(A, b, tol, fx["gmres_dim"])
        assert st == 0
        assert abs(len(got) - len(h["hist"])) <= 1
        ref = np.array(h["hist"])
        k = min(len(got), len(ref))
        # residual estimates at the printed 5 digits (E13.5) while they are
        # far above rounding; the late iterations amplify summation-order
        # differences between BLAS libraries
        early = ref[:k] > 1e-2 * ref[0]
        assert rel_close(got[:k][early], ref[:k][early], 2e-4, 1e-9).all()
        assert (np.abs(np.log10(got[:k][~early] / ref[:k][~early])) < 1.0).all()


@pytest.mark.parametrize("case", BEM_CASES)
def test_bem_results_match_log(case):
    fx = load_fixture(case)
    m = mesh_from_fixture(fx)
    freq = np.array(fx["freq_au"])
    eps = np.array([complex(a, b) for a, b in fx["eps_w"]]) if "eps_w" in fx else pm.silver_etchegoin(freq)
    r = pxo.bem_solve_spectrum(m, freq, eps, fx["e0"], method=method_id(fx["method"]), tol=fx["tol"] * np.sqrt(m.n),
                               m=fx["gmres_dim"], want_variables=True)
    assert r["status"] == 0
    exp = np.array(fx["expected"]["results"])
    assert rel_close(r["results"][:, :12], exp[:, :12], 1e-5, 1e-6).all()
    ex = fx["expected"]
    if "charges" in ex:
        var = r["variables"][0]  # the log prints the unknowns as solved (bem_module.f90:1028-1036)
        q = np.stack([var.real, var.imag], -1).transpose(1, 0, 2).reshape(m.n, 6)
        assert rel_close(q, np.array(ex["charges"]), 5e-3, 2e-3).all()
    if "gmres" in ex:
        assert [len(h["hist"]) for h in ex["gmres"]] == list(r["iters"][0][:len(ex["gmres"])])


def test_etchegoin_matches_host_tables():
    freq = np.linspace(1.0, 5.0, 41) / 27.211386245981
    for metal, fn in (("ag", pm.silver_etchegoin), ("au", pm.gold_etchegoin)):
        for polar in (False, True):
            np.testing.assert_allclose(pxo.etchegoin(metal, freq, polar), fn(freq, polar=polar), rtol=1e-13)
