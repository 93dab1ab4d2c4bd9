"""The spectral route of the homogeneous omegaFQ inversion: one generalized
symmetric eigendecomposition T K T x = lambda T x serves every frequency,
x(w) = X diag(1 / (lambda - z(w))) X^T T b.  Checked against the LU route on
the same context ("spectral" 0) and against the oracle's LU (the ZGESV
restatement of algorithm_module.f90:510-541).  Bars as for the LU:
polarizabilities 1e-10, charges 1e-9 (relative to the largest)."""
import dataclasses

import numpy as np
import pytest

from oracle import pxo
from plasmonx_amd import Context, _lib
from plasmonx_amd.model import AtomSystem, system_from_fixture
from plasmonx_amd.workloads import c1_sodium_sphere, c4_flakes
from conftest import load_fixture

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def solve(ctx, s, freq, e0, spectral, pol_mask=7):
    ctx.set_option("spectral", spectral)
    out = ctx.solve_spectrum(s.n_var, freq, e0, pol_mask=pol_mask, method=_lib.INVERSION, want_variables=True)
    out["spectral_targets"] = ctx.stats()["spectral_targets"]
    return out


@pytest.mark.parametrize("case", ["wfq_sodium_inversion", "wfq_graphene_disk_inversion",
                                  "wfq_graphene_ribbon_inversion", "wfq_graphene_triangle_inversion"])
def test_spectral_matches_lu_and_oracle(case):
    fx = load_fixture(case)
    s = system_from_fixture(fx)
    freq = np.array(fx["freq_au"])[:20]
    print(case, "n", s.n_var, "freq", len(freq), flush=True)
    with Context(0) as ctx:
        ctx.set_option("kbar_cutoff", 0)
        ctx.set_atoms(s)
        sp = solve(ctx, s, freq, fx["e0"], 2)
        lu = solve(ctx, s, freq, fx["e0"], 0)
    assert sp["spectral_targets"] == 1 and lu["spectral_targets"] == 0
    ref = pxo.solve_spectrum(s, freq, fx["e0"], method=0, want_variables=True)
    assert relerr(sp["polar"], ref["polar"]) < 1e-10
    assert relerr(sp["variables"], ref["variables"]) < 1e-9
    assert relerr(sp["polar"], lu["polar"]) < 1e-10
    assert relerr(sp["results"], lu["results"]) < 1e-10


def na_cluster(kernel, n=300, eta=None):
    s, freq = c1_sodium_sphere()
    t = s.types[0] if eta is None else dataclasses.replace(s.types[0], eta=eta)
    return AtomSystem(model="wfq", kernel=kernel, symbols=s.symbols[:n], coord_bohr=s.coord_bohr[:n], types=[t],
                      fermi_d=s.fermi_d, fermi_s=s.fermi_s, r0=s.r0), freq


@pytest.mark.parametrize("kernel", ["coulomb", "ohno"])
def test_spectral_other_kernels(kernel):
    s, freq = na_cluster(kernel)
    pick = freq[[0, 20, 40, 60]]
    with Context(0) as ctx:
        ctx.set_atoms(s)
        sp = solve(ctx, s, pick, 1e-4, 2, pol_mask=1)
    assert sp["spectral_targets"] == 1
    ref = pxo.solve_spectrum(s, pick, 1e-4, pol_mask=1, method=0, want_variables=True)
    assert relerr(sp["polar"], ref["polar"]) < 1e-10
    assert relerr(sp["variables"], ref["variables"]) < 1e-9


def test_spectral_falls_back_to_lu_when_t_is_indefinite():
    # Coulomb kernel with eta = 0.05: T has a negative eigenvalue (checked on
    # the CPU in test_oracle_spectral), so the Cholesky step of the
    # eigensolver fails and the LU solves the system
    s, freq = na_cluster("coulomb", eta=0.05)
    pick = freq[[0, 30]]
    with Context(0) as ctx:
        ctx.set_atoms(s)
        sp = solve(ctx, s, pick, 1e-4, 2, pol_mask=1)
        lu = solve(ctx, s, pick, 1e-4, 0, pol_mask=1)
    assert sp["spectral_targets"] == 0
    assert np.array_equal(sp["polar"], lu["polar"])
    ref = pxo.solve_spectrum(s, pick, 1e-4, pol_mask=1, method=0)
    assert relerr(sp["polar"], ref["polar"]) < 1e-10


def test_group_with_spectral_targets():
    flakes, freq = c4_flakes(6, hi=2500)
    pick = freq[[3, 50, 100, 127]]
    print("flakes", [f.n_var for f in flakes], flush=True)
    ctxs = []
    for k, f in enumerate(flakes):
        c = Context(0)
        c.set_atoms(f)
        c.set_option("spectral", 2 if k % 2 == 0 else 0)
        ctxs.append(c)
    mixed = _lib.solve_group(ctxs, pick, 1e-4, pol_mask=3)
    n_spec = ctxs[0].stats()["spectral_targets"]
    for c in ctxs:
        c.set_option("spectral", 0)
    lu = _lib.solve_group(ctxs, pick, 1e-4, pol_mask=3)
    assert n_spec == 3
    for f, a, b in zip(flakes, mixed, lu):
        assert relerr(a["polar"], b["polar"]) < 1e-10
        assert relerr(a["results"], b["results"]) < 1e-10
    ref = pxo.solve_spectrum(flakes[0], pick, 1e-4, pol_mask=3, method=0)
    assert relerr(mixed[0]["polar"], ref["polar"]) < 1e-10
    for c in ctxs:
        c.close()


def test_large_flake_takes_the_spectral_route_by_itself():
    # a 128-frequency spectrum of a ~6000-atom flake: one decomposition is
    # estimated (and is) faster than 128 LUs, so the default option picks it;
    # 4 of the frequencies solved by the LU agree
    flakes, freq = c4_flakes(1, lo=5500, hi=6500)
    f = flakes[0]
    print("flake atoms", f.n_var, flush=True)
    with Context(0) as ctx:
        ctx.set_atoms(f)
        full = ctx.solve_spectrum(f.n_var, freq, 1e-4, pol_mask=3, method=_lib.INVERSION)
        assert ctx.stats()["spectral_targets"] == 1
        pick = [0, 42, 85, 127]
        lu = ctx.solve_spectrum(f.n_var, freq[pick], 1e-4, pol_mask=3, method=_lib.INVERSION)
        assert ctx.stats()["spectral_targets"] == 0
    assert relerr(full["polar"][pick], lu["polar"]) < 1e-9


def test_spectral_rods_in_the_request_loop_match_alone():
    # three Na rods large enough for the spectral route, served concurrently
    # with an LU request: each result equals the one solved alone
    from plasmonx_amd.workloads import sodium_rod
    rods = [sodium_rod(L, 14.0) for L in (110.0, 120.0, 130.0)]
    freq = np.linspace(2.5, 4.0, 120) / 27.211386245981
    jobs = []
    for r in rods:
        c = Context(0)
        c.set_atoms(r)
        jobs.append((c, r, dict(freq_au=freq, e0=1e-4, pol_mask=1, method=_lib.INVERSION)))
    small = sodium_rod(30.0, 9.0)
    c = Context(0)
    c.set_atoms(small)
    jobs.append((c, small, dict(freq_au=freq[:6], e0=1e-4, pol_mask=1, method=_lib.INVERSION)))
    print("rod atoms", [r.n_var for _, r, _ in jobs], flush=True)
    alone, routes = [], []
    for c, r, kw in jobs:
        alone.append(c.solve_spectrum(r.n_var, **kw))
        routes.append(c.stats()["spectral_targets"])
    assert routes == [1, 1, 1, 0]
    srv = _lib.Server(0, hbm_budget=0, max_inflight=4)
    try:
        ids = [srv.submit(c, **kw) for c, _, kw in jobs]
        mixed = [srv.wait(i) for i in ids]
    finally:
        srv.close()
    for a, m in zip(alone, mixed):
        assert np.array_equal(a["polar"], m["polar"])
    ref = pxo.solve_spectrum(jobs[0][1], freq[[0, 60]], 1e-4, pol_mask=1, method=0)
    assert relerr(alone[0]["polar"][[0, 60]], ref["polar"]) < 1e-10
    for c, _, _ in jobs:
        c.close()


def test_spectral_beside_other_contexts_matches_alone():
    # a spectral decomposition while other host threads run LU and GMRES-IR
    # solves on their own contexts: the device gate keeps the decomposition
    # alone on the device, and the result equals the one computed alone
    import threading
    from plasmonx_amd.workloads import core_shell_sphere, sodium_rod
    big = sodium_rod(200.0, 14.0)
    freq = np.linspace(2.5, 4.0, 120) / 27.211386245981
    small = sodium_rod(60.0, 12.0)
    fs = np.linspace(2.5, 4.0, 40) / 27.211386245981
    cs = core_shell_sphere(0.6 * 24.0, 24.0)
    fc = np.linspace(2.8, 4.2, 4) / 27.211386245981
    print("atoms", big.n_var, small.n_var, cs.n_var, flush=True)
    with Context(0) as ca, Context(0) as cb, Context(0) as cc:
        ca.set_atoms(big)
        cb.set_atoms(small)
        cb.set_option("spectral", 0)
        cc.set_atoms(cs)
        cc.set_alpha(cs.alpha_table(fc))
        alone = ca.solve_spectrum(big.n_var, freq, 1e-4, pol_mask=7, method=_lib.INVERSION)["polar"]
        assert ca.stats()["spectral_targets"] == 1
        stop = threading.Event()

        def load(ctx, fn):
            while not stop.is_set():
                fn(ctx)

        def lu(ctx):
            ctx.solve_spectrum(small.n_var, fs, 1e-4, pol_mask=7, method=_lib.INVERSION)

        def gmres(ctx):
            ctx.solve_spectrum(cs.n_var, fc, 1e-4, pol_mask=7, method=_lib.ITERATIVE_OTF,
                               tol_abs=1e-12 * np.sqrt(cs.n_var), krylov_dim=200, precision=_lib.MIXED)

        th = [threading.Thread(target=load, args=(cb, lu)), threading.Thread(target=load, args=(cc, gmres))]
        for t in th:
            t.start()
        try:
            mixed = [ca.solve_spectrum(big.n_var, freq, 1e-4, pol_mask=7, method=_lib.INVERSION)["polar"]
                     for _ in range(3)]
        finally:
            stop.set()
            for t in th:
                t.join()
    for m in mixed:
        assert np.array_equal(m, alone)
