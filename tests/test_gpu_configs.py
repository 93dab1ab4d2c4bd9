"""The benchmark configurations at their real sizes, against the oracle:
C1 (the 1745-atom Na sphere, dense LU), C2 (the 15683-atom Ag sphere: the
FP64 product and a few frequencies of the mixed solve) and C4 (seeded
graphene flakes solved together by the request loop, as the bench does).

Bars: LU spectra 1e-10 against the oracle's LU; FP64 products 1e-12;
mixed solves: the FP64 true residual below tol and charge conservation."""
import numpy as np
import pytest

from oracle import pxo
from plasmonx_amd import Context, _lib
from plasmonx_amd._lib import Server

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_c1_full_sphere_lu_matches_oracle(gpu_ctx):
    from plasmonx_amd.workloads import c1_sodium_sphere
    s, freq = c1_sodium_sphere()
    pick = freq[[0, 30, 60]]
    print(f"C1: {s.n_atoms} atoms, {len(pick)} w", flush=True)
    gpu_ctx.set_option("kbar_cutoff", 1)
    gpu_ctx.set_atoms(s)
    got = gpu_ctx.solve_spectrum(s.n_var, pick, 1e-4, pol_mask=1, method=_lib.INVERSION, want_variables=True)
    ref = pxo.solve_spectrum(s, pick, 1e-4, pol_mask=1, method=0, want_variables=True)
    assert relerr(got["polar"], ref["polar"]) < 1e-10
    assert relerr(got["variables"][:, 0], ref["variables"][:, 0]) < 1e-9


def test_c2_full_sphere_product_and_mixed_solve(gpu_ctx):
    """15683 atoms (62732 unknowns): one FP64 product against the oracle's
    on-the-fly product (exact K-bar on both sides), then the bench's mixed
    solve at 3 frequencies: FP64 true residual below tol, and the induced
    charges summing to zero (1^T K-bar = 0) for every system."""
    from plasmonx_amd.workloads import c2_silver_sphere
    s, freq = c2_silver_sphere()
    assert s.n_atoms == 15683
    pick = freq[[0, 100, 200]]
    aw = s.alpha_table(pick)
    print(f"C2: {s.n_atoms} atoms", flush=True)
    gpu_ctx.set_option("kbar_cutoff", 0)
    gpu_ctx.set_atoms(s)
    gpu_ctx.set_alpha(aw)
    rng = np.random.default_rng(15683)
    x = rng.normal(size=s.n_var) + 1j * rng.normal(size=s.n_var)
    got = gpu_ctx.matvec(pick[1], 1, x[None, :])[0]
    ref = pxo.matvec_otf(s, pick[1], aw[:, 1], x)
    assert relerr(got, ref) < 1e-12
    gpu_ctx.set_option("kbar_cutoff", 1)
    gpu_ctx.set_atoms(s)
    gpu_ctx.set_alpha(aw)
    gpu_ctx.set_option("batch", 48)
    tol = 1e-12 * np.sqrt(s.n_var)
    out = gpu_ctx.solve_spectrum(s.n_var, pick, 1e-4, method=_lib.ITERATIVE_OTF, tol_abs=tol, krylov_dim=120,
                                 precision=_lib.MIXED, want_variables=True)
    gpu_ctx.set_option("batch", 1)
    assert np.all(out["true_residual"] < tol)
    # 1^T K-bar = 0 turns the charge rows into -z_q(w) sum(q) = -1^T r_q, so
    # |sum q| <= sqrt(N) ||r|| / |z_q| with ||r|| < tol (the true residual)
    t = s.types[0]
    zq = np.abs(-(pick / (2 * t.density * t.tau)) * (pick * t.tau + 1j))
    q = out["variables"][:, :, :s.n_atoms]
    bound = np.sqrt(s.n_atoms) * out["true_residual"] / zq[:, None]
    assert np.all(np.abs(q.sum(axis=2)) <= bound)
    assert np.all(np.abs(q.sum(axis=2)) < 1e-8 * np.abs(q).sum(axis=2))
    assert np.all(out["results"][:, 1] > 0)  # absorption (Im alpha_iso) positive


def test_c4_flakes_together_match_oracle_and_alone():
    """Six seeded C4 flakes (500-1500 atoms) submitted together to the
    request loop (dense LU, 4 in flight), as the bench runs the batch: each
    flake's spectrum at 3 frequencies against the oracle's LU (1e-10) and
    against the same flake solved alone (bit for bit)."""
    from plasmonx_amd.workloads import c4_flakes
    flakes, freq = c4_flakes(6, hi=1500)
    pick = freq[[10, 64, 120]]
    print("C4 flakes:", [f.n_atoms for f in flakes], flush=True)
    srv = Server(0, max_inflight=4)
    ctxs, ids = [], []
    for s in flakes:
        c = Context(0)
        c.set_atoms(s)
        ctxs.append(c)
        ids.append(srv.submit(c, pick, 1e-4, pol_mask=3, method=_lib.INVERSION))
    outs = [srv.wait(i) for i in ids]
    srv.close()
    for s, c, o in zip(flakes, ctxs, outs):
        alone = c.solve_spectrum(s.n_var, pick, 1e-4, pol_mask=3, method=_lib.INVERSION)
        ref = pxo.solve_spectrum(s, pick, 1e-4, pol_mask=3, method=0)
        assert np.array_equal(o["polar"], alone["polar"])
        assert relerr(o["polar"], ref["polar"]) < 1e-10
        c.close()
