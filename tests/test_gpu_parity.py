"""libpxh (HIP, gfx950) against the FP64 oracle and the reference logs.

Bars: FP64 products and inversions agree with the oracle to 1e-10
(relative, rounding-order differences only); GMRES results to the solver
tolerance; mixed precision (FP32 interaction tiles inside FP64 refinement)
to <= 1e-6 relative error of the polarizability tensor.
"""
import numpy as np
import pytest

from conftest import all_cases, load_fixture, method_id, rel_close
from oracle import pxo
from plasmonx_amd import _lib
from plasmonx_amd import permittivity as pm
from plasmonx_amd.model import mesh_from_fixture, system_from_fixture

pytestmark = pytest.mark.gpu

SMALL = ["wfqfmu_ag_inversion", "wfq_sodium_iterative_on_the_fly", "wfqfmu_hetero_agau_inversion",
         "wfq_graphene_triangle_inversion", "wfqfmu_hetero_agau_read_1param_inversion"]


def setup(ctx, fx, nf=None):
    s = system_from_fixture(fx)
    freq = np.array(fx["freq_au"])[:nf]
    ctx.set_option("kbar_cutoff", 0)
    ctx.set_atoms(s)
    aw = None
    if s.model == "wfqfmu":
        aw = s.alpha_table(freq)
        ctx.set_alpha(aw)
    return s, freq, aw


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("case", SMALL)
def test_rhs_matches_oracle(gpu_ctx, case):
    fx = load_fixture(case)
    s, freq, aw = setup(gpu_ctx, fx, 1)
    got = gpu_ctx.dynamic_rhs(s.n_var, freq[0], fx["e0"])
    ref = pxo.rhs(s, fx["e0"], freq[0])
    assert relerr(got, ref) < 1e-12


@pytest.mark.parametrize("case", SMALL)
def test_matvec_fp64_matches_oracle(gpu_ctx, case):
    fx = load_fixture(case)
    s, freq, aw = setup(gpu_ctx, fx, 1)
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, s.n_var)) + 1j * rng.normal(size=(2, s.n_var))
    got = gpu_ctx.matvec(freq[0], 0, x)
    for k in range(2):
        ref = pxo.matvec_otf(s, freq[0], aw[:, 0] if aw is not None else np.ones(1), x[k])
        assert relerr(got[k], ref) < 1e-12


def test_kbar_cutoff_is_exact_to_rounding(gpu_ctx):
    fx = load_fixture("wfq_sodium_iterative_on_the_fly")
    s, freq, aw = setup(gpu_ctx, fx, 1)
    x = np.random.default_rng(2).normal(size=(1, s.n_var)) + 0j
    full = gpu_ctx.matvec(freq[0], 0, x)
    gpu_ctx.set_option("kbar_cutoff", 1)
    cut = gpu_ctx.matvec(freq[0], 0, x)
    gpu_ctx.set_option("kbar_cutoff", 0)
    assert relerr(cut, full) < 1e-13


@pytest.mark.parametrize("case", ["wfqfmu_ag_inversion", "wfq_sodium_iterative_on_the_fly"])
def test_constant_matrix_matches_oracle(gpu_ctx, case):
    fx = load_fixture(case)
    s, freq, aw = setup(gpu_ctx, fx, 1)
    got = gpu_ctx.constant_matrix(s.n_var)
    ref = pxo.constant_matrix(s)
    assert relerr(got, ref) < 1e-12


@pytest.mark.parametrize("case", [c for c in all_cases("atoms") if load_fixture(c)["method"] == "inversion"])
def test_inversion_matches_oracle_and_log(gpu_ctx, case):
    fx = load_fixture(case)
    nf = min(len(fx["freq_au"]), 20)
    s, freq, aw = setup(gpu_ctx, fx, nf)
    got = gpu_ctx.solve_spectrum(s.n_var, freq, fx["e0"], method=_lib.INVERSION, want_variables=True)
    ref = pxo.solve_spectrum(s, freq, fx["e0"], alpha_w=aw, method=0, want_variables=True)
    assert relerr(got["polar"], ref["polar"]) < 1e-10
    assert relerr(got["variables"], ref["variables"]) < 1e-9
    exp = np.array(fx["expected"]["results"])[:nf]
    assert rel_close(got["results"][:, :12], exp[:, :12], 1e-5, 1e-6).all()


@pytest.mark.parametrize("case", ["wfq_sodium_iterative_on_the_fly", "wfqfmu_ag_iterative_on_the_fly",
                                  "wfqfmu_hetero_agau_iterative_on_the_fly", "wfqfmu_ag_iterative_on_memory",
                                  "wfqfmu_hetero_agau_iterative_on_memory"])
def test_gmres_fp64_matches_oracle(gpu_ctx, case):
    fx = load_fixture(case)
    nf = min(len(fx["freq_au"]), 4)
    s, freq, aw = setup(gpu_ctx, fx, nf)
    tol = fx["tol"] * np.sqrt(s.n_var)
    gpu_ctx.set_option("batch", 1)
    got = gpu_ctx.solve_spectrum(s.n_var, freq, fx["e0"], method=method_id(fx["method"]), tol_abs=tol,
                                 krylov_dim=fx["gmres_dim"])
    ref = pxo.solve_spectrum(s, freq, fx["e0"], alpha_w=aw, method=0)
    assert np.all(got["true_residual"] < tol)
    assert relerr(got["polar"], ref["polar"]) < 1e-4
    exp = np.array(fx["expected"]["results"])[:nf]
    assert rel_close(got["results"][:, :12], exp[:, :12], 1e-4, 1e-6).all()
    # the reference's iteration counts (printed GMRES histories), within one
    if "gmres" in fx["expected"]:
        ev = freq * 27.211386245981
        for h in fx["expected"]["gmres"]:
            f = int(np.argmin(np.abs(ev - h["ev"])))
            if abs(ev[f] - h["ev"]) < 1e-6:
                assert abs(got["iters"][f, "XYZ".index(h["pol"])] - len(h["hist"])) <= 1


@pytest.mark.parametrize("case", ["wfqfmu_ag_inversion", "wfq_sodium_inversion", "wfqfmu_hetero_agau_inversion"])
def test_mixed_batched_within_1e6(gpu_ctx, case):
    fx = load_fixture(case)
    nf = min(len(fx["freq_au"]), 24)
    s, freq, aw = setup(gpu_ctx, fx, nf)
    tol = 1e-12 * np.sqrt(s.n_var)
    gpu_ctx.set_option("batch", 16)
    got = gpu_ctx.solve_spectrum(s.n_var, freq, fx["e0"], method=_lib.ITERATIVE_OTF, tol_abs=tol, krylov_dim=200,
                                 precision=_lib.MIXED)
    gpu_ctx.set_option("batch", 1)
    ref = pxo.solve_spectrum(s, freq, fx["e0"], alpha_w=aw, method=0)
    assert np.all(got["true_residual"] < tol)
    for f in range(nf):
        err = np.linalg.norm(got["polar"][f] - ref["polar"][f]) / np.linalg.norm(ref["polar"][f])
        assert err < 1e-6, (f, err)


def test_batched_result_independent_of_batch(gpu_ctx):
    fx = load_fixture("wfqfmu_ag_inversion")
    freq = np.linspace(2.5, 3.5, 9) / 27.211386245981
    s = system_from_fixture(fx)
    gpu_ctx.set_option("kbar_cutoff", 0)
    gpu_ctx.set_atoms(s)
    gpu_ctx.set_alpha(pm.silver_etchegoin(freq, polar=True)[None, :])
    tol = 1e-11 * np.sqrt(s.n_var)
    out = []
    for b in (1, 3, 27):
        gpu_ctx.set_option("batch", b)
        out.append(gpu_ctx.solve_spectrum(s.n_var, freq, fx["e0"], method=_lib.ITERATIVE_OTF, tol_abs=tol,
                                          krylov_dim=200, precision=_lib.MIXED)["polar"])
    gpu_ctx.set_option("batch", 1)
    assert relerr(out[1], out[0]) < 1e-8
    assert relerr(out[2], out[0]) < 1e-8


@pytest.mark.parametrize("case", all_cases("bem"))
def test_bem_matches_oracle_and_log(gpu_ctx, case):
    fx = load_fixture(case)
    m = mesh_from_fixture(fx)
    freq = np.array(fx["freq_au"])
    eps = np.array([complex(a, b) for a, b in fx["eps_w"]]) if "eps_w" in fx else pm.silver_etchegoin(freq)
    gpu_ctx.set_mesh(m)
    gpu_ctx.set_permittivity(eps)
    D, Di, SA = pxo.bem_d(m)
    assert relerr(gpu_ctx.constant_matrix(m.n), D) < 1e-11
    meth = method_id(fx["method"])
    tol = fx["tol"] * np.sqrt(m.n)
    gpu_ctx.set_option("batch", 1)
    got = gpu_ctx.solve_spectrum(m.n, freq, fx["e0"], method=meth, tol_abs=tol, krylov_dim=fx["gmres_dim"])
    ref = pxo.bem_solve_spectrum(m, freq, eps, fx["e0"], method=0)
    assert relerr(got["polar"], ref["polar"]) < (1e-10 if meth == 0 else 1e-5)
    exp = np.array(fx["expected"]["results"])
    assert rel_close(got["results"][:, :12], exp[:, :12], 1e-5 if meth == 0 else 1e-4, 1e-6).all()


@pytest.mark.parametrize("case", ["bem_dpcm_accurate_inversion", "bem_ief_accurate_inversion",
                                  "bem_ief_approximate_inversion"])
@pytest.mark.parametrize("drop", [0, 1])
@pytest.mark.parametrize("nvec", [1, 17, 40])
def test_bem_matvec_matches_oracle(gpu_ctx, case, drop, nvec):
    """The BEM operator products (matrix-core FP64 GEMM) against the oracle's
    D and S A^-1: DPCM y = (D + lambda I) x, IEF y = -(D + lambda I) S A^-1 x,
    with lambda = 2 pi (eps_s + eps) / (eps_s - eps).  284 tesserae (a
    partial row tile) or 283 (odd), 1..40 vectors (one, two and three column
    tiles); the FP32-matrix product within 1e-6."""
    import dataclasses
    fx = load_fixture(case)
    m = mesh_from_fixture(fx)
    if drop:
        m = dataclasses.replace(m, verts_bohr=m.verts_bohr[:m.n - drop])
    freq = np.array([3.2, 3.6]) / 27.211386245981
    eps = pm.silver_etchegoin(freq)
    gpu_ctx.set_mesh(m)
    gpu_ctx.set_permittivity(eps)
    D, Di, SA = pxo.bem_d(m)
    lam = 2 * np.pi * (m.eps_solvent + eps[1]) / (m.eps_solvent - eps[1])
    A = D + lam * np.eye(m.n)
    if m.variant == "iefpcm":
        A = -A @ SA
    rng = np.random.default_rng(nvec + 100 * drop)
    x = rng.normal(size=(nvec, m.n)) + 1j * rng.normal(size=(nvec, m.n))
    ref = x @ A.T
    got = gpu_ctx.matvec(freq[1], 1, x)
    for k in range(nvec):
        assert relerr(got[k], ref[k]) < 1e-12, k
    gpu_ctx.set_option("matvec_mixed", 1)
    got = gpu_ctx.matvec(freq[1], 1, x)
    gpu_ctx.set_option("matvec_mixed", 0)
    for k in range(nvec):
        assert relerr(got[k], ref[k]) < 1e-6, k


@pytest.mark.parametrize("variant", ["dpcm", "iefpcm"])
@pytest.mark.parametrize("nvec", [3, 40])
def test_bem_matvec_split_k(gpu_ctx, variant, nvec):
    """1280 tesserae: the product columns are split over several workgroups
    per tile (partials summed by a second kernel); against the oracle's
    matrices as above."""
    from plasmonx_amd import geometry
    from plasmonx_amd.model import Mesh
    m = Mesh(verts_bohr=geometry.icosphere(60.0, 3, seed=1) / 0.529177210903, variant=variant,
             green="approximate", sphere_r_bohr=60.0 / 0.529177210903)
    freq = np.array([3.2, 3.6]) / 27.211386245981
    eps = pm.silver_etchegoin(freq)
    gpu_ctx.set_mesh(m)
    gpu_ctx.set_permittivity(eps)
    D, Di, SA = pxo.bem_d(m)
    lam = 2 * np.pi * (m.eps_solvent + eps[1]) / (m.eps_solvent - eps[1])
    A = D + lam * np.eye(m.n)
    if variant == "iefpcm":
        A = -A @ SA
    rng = np.random.default_rng(nvec)
    x = rng.normal(size=(nvec, m.n)) + 1j * rng.normal(size=(nvec, m.n))
    ref = x @ A.T
    got = gpu_ctx.matvec(freq[1], 1, x)
    for k in range(nvec):
        assert relerr(got[k], ref[k]) < 1e-12, k


@pytest.mark.parametrize("mfma", [0, 1])
@pytest.mark.parametrize("nvec", [1, 5, 20, 50])
def test_mixed_product_accuracy(gpu_ctx, mfma, nvec):
    """The low-precision product (FP32 VALU tiles, or fp16 hi/lo MFMA) against
    the FP64 product on a 675-atom Ag sphere: relative error <= 2e-5 per vector
    (the inner solves need ~1e-5; FP64 refinement fixes the final accuracy)."""
    from plasmonx_amd.workloads import c2_silver_sphere
    s, freq = c2_silver_sphere(radius=14.0, n_freq=3)
    gpu_ctx.set_option("kbar_cutoff", 1)
    gpu_ctx.set_atoms(s)
    gpu_ctx.set_alpha(s.alpha_table(freq))
    rng = np.random.default_rng(nvec)
    x = rng.normal(size=(nvec, s.n_var)) + 1j * rng.normal(size=(nvec, s.n_var))
    x[-1] *= 1e-3  # columns of different magnitude
    ref = gpu_ctx.matvec(freq[1], 1, x)
    gpu_ctx.set_option("mfma", mfma)
    gpu_ctx.set_option("matvec_mixed", 1)
    got = gpu_ctx.matvec(freq[1], 1, x)
    gpu_ctx.set_option("matvec_mixed", 0)
    gpu_ctx.set_option("mfma", 1)
    for k in range(nvec):
        err = np.linalg.norm(got[k] - ref[k]) / np.linalg.norm(ref[k])
        assert err < 2e-5, (k, err)


@pytest.mark.parametrize("alloy", [False, True])
def test_four_row_block_layout_is_bit_identical(gpu_ctx, alloy):
    """One-column-block MFMA passes (<= 16 systems) run four row blocks per
    wave (each pair's elements built once per workgroup): the products equal
    those of the two-row-block layout bit for bit (same elements, same
    accumulation order), for one and for two atom types."""
    from plasmonx_amd.workloads import c2_silver_sphere, core_shell_sphere
    if alloy:
        s = core_shell_sphere(7.0, 12.0)
        freq = np.linspace(2.8, 4.2, 3) / 27.211386245981
    else:
        s, freq = c2_silver_sphere(radius=14.0, n_freq=3)
    gpu_ctx.set_option("kbar_cutoff", 1)
    gpu_ctx.set_atoms(s)
    gpu_ctx.set_alpha(s.alpha_table(freq))
    rng = np.random.default_rng(16)
    x = rng.normal(size=(16, s.n_var)) + 1j * rng.normal(size=(16, s.n_var))
    gpu_ctx.set_option("matvec_mixed", 1)
    got = []
    for r4 in (1, 0):
        gpu_ctx.set_option("mfma_r4", r4)
        got.append(gpu_ctx.matvec(freq[1], 1, x))
    gpu_ctx.set_option("mfma_r4", 1)
    gpu_ctx.set_option("matvec_mixed", 0)
    assert np.array_equal(got[0], got[1])
    ref = gpu_ctx.matvec(freq[1], 1, x)
    assert max(np.linalg.norm(got[0][k] - ref[k]) / np.linalg.norm(ref[k]) for k in range(16)) < 2e-5


def test_autotuned_mfma_product(gpu_ctx):
    """3.4k-atom Ag sphere, 48 vectors: the first full MFMA pass of this atom
    count times the candidate source splits and keeps the fastest; the
    product with the tuned splits and with the fixed heuristic both stay
    within the low-precision bar (2e-5 per vector) of the FP64 product."""
    from plasmonx_amd.workloads import c2_silver_sphere
    s, freq = c2_silver_sphere(radius=24.0, n_freq=2)
    assert s.n_atoms >= 2048
    gpu_ctx.set_option("kbar_cutoff", 1)
    gpu_ctx.set_atoms(s)
    gpu_ctx.set_alpha(s.alpha_table(freq))
    rng = np.random.default_rng(3)
    x = rng.normal(size=(48, s.n_var)) + 1j * rng.normal(size=(48, s.n_var))
    ref = gpu_ctx.matvec(freq[1], 1, x)
    gpu_ctx.set_option("matvec_mixed", 1)
    outs = []
    for tune in (1, 0, 1):
        gpu_ctx.set_option("autotune", tune)
        outs.append(gpu_ctx.matvec(freq[1], 1, x))
    gpu_ctx.set_option("autotune", 1)
    gpu_ctx.set_option("matvec_mixed", 0)
    for got in outs:
        for k in range(48):
            assert relerr(got[k], ref[k]) < 2e-5, k
    assert np.array_equal(outs[0], outs[2])  # the tuned choice is kept


@pytest.mark.parametrize("case", ["bem_ief_accurate_inversion", "bem_ief_approximate_inversion"])
def test_bem_ief_iterative_agrees_with_inversion(gpu_ctx, case):
    """The iterative IEF applies the full operator -(D + lambda I) S A^-1, so
    GMRES reaches the inversion's answer (the reference's iterative IEF omits
    S A^-1, algorithm_module.f90:356-358)."""
    fx = load_fixture(case)
    m = mesh_from_fixture(fx)
    freq = np.array(fx["freq_au"])[:4]
    eps = np.array([complex(a, b) for a, b in fx["eps_w"]])[:4] if "eps_w" in fx else pm.silver_etchegoin(freq)
    gpu_ctx.set_mesh(m)
    gpu_ctx.set_permittivity(eps)
    gpu_ctx.set_option("batch", 1)
    tol = 1e-11 * np.sqrt(m.n)
    got = gpu_ctx.solve_spectrum(m.n, freq, fx["e0"], method=_lib.ITERATIVE, tol_abs=tol, krylov_dim=600)
    ref = pxo.bem_solve_spectrum(m, freq, eps, fx["e0"], method=0)
    assert np.all(got["true_residual"] < tol)
    assert relerr(got["polar"], ref["polar"]) < 1e-6


@pytest.mark.parametrize("case", ["wfqfmu_ag_inversion", "wfqfmu_hetero_agau_inversion"])
def test_yaml_input_to_log_results(gpu_ctx, case):
    """A reference input file read by plasmonx_amd.inputs, solved by libpxh:
    the log's result columns (E14.6) to 1e-5."""
    import os
    from plasmonx_amd.inputs import read_input, system_from_case
    c = read_input(os.path.join(os.path.dirname(__file__), "golden", "inputs", case, case + ".yaml"))
    s = system_from_case(c)
    freq = np.array(c["freq_au"])
    gpu_ctx.set_atoms(s)
    gpu_ctx.set_alpha(s.alpha_table(freq))
    got = gpu_ctx.solve_spectrum(s.n_var, freq, c["e0"], method=_lib.INVERSION)
    exp = np.array(load_fixture(case)["expected"]["results"])
    assert rel_close(got["results"][:, :12], exp[:, :12], 1e-5, 1e-6).all()


@pytest.mark.parametrize("case", ["wfq_sodium_minimal_input", "wfq_graphene_minimal_input", "wfqfmu_ag_minimal_input",
                                  "wfqfmu_au_minimal_input", "wfqfmu_hetero_agau_minimal_input",
                                  "wfqfmu_au_read_xyz"])
def test_minimal_deck_to_log_results(gpu_ctx, case):
    """The reference's minimal decks, whose atom types and parameters come
    from the reference's defaults, read and solved by inversion at up to 12
    of the logged frequencies: the result columns (E14.6) to 1e-5."""
    import json
    import os
    from plasmonx_amd.inputs import read_input, system_from_case
    d = os.path.join(os.path.dirname(__file__), "golden", "inputs", case)
    c = read_input(os.path.join(d, case + ".yaml"))
    with open(os.path.join(d, "log.json")) as f:
        log = json.load(f)
    ev = np.array(c["freq_au"]) * 27.211386245981
    idx = [int(np.argmin(np.abs(ev - e))) for e in log["results_ev"]]
    freq = np.array(c["freq_au"])[idx]
    s = system_from_case(c)
    for t in s.types:  # the etchegoin fits at the picked frequencies only
        if t.alpha_w is not None:
            t.alpha_w = np.asarray(t.alpha_w)[idx]
    gpu_ctx.set_option("kbar_cutoff", 0)
    gpu_ctx.set_atoms(s)
    if s.model == "wfqfmu":
        gpu_ctx.set_alpha(s.alpha_table(freq))
    got = gpu_ctx.solve_spectrum(s.n_var, freq, c["e0"], method=_lib.INVERSION)
    exp = np.array(log["results"])
    assert rel_close(got["results"][:, :12], exp[:, :12], 1e-5, 1e-6).all()


@pytest.mark.parametrize("case", ["wfqfmu_ag_inversion", "wfqfmu_hetero_agau_inversion", "sphere20"])
@pytest.mark.parametrize("nvec", [5, 8, 11])
def test_fp64_matrix_core_product_matches_vector_kernel(gpu_ctx, case, nvec):
    """FP64 omegaFQFmu products of >= 5 vectors run on the FP64 matrix cores
    (k_tsweep_f64mfma); the same sums as the FP64 vector kernel in another
    order: 1e-13 between the two, 1e-12 against the oracle."""
    if case == "sphere20":
        from plasmonx_amd import geometry
        from plasmonx_amd.model import TOBOHR, AtomSystem
        from plasmonx_amd.workloads import silver_type
        coord = geometry.fcc_sphere(4.09, 20.0) * TOBOHR
        s = AtomSystem(model="wfqfmu", kernel="gaussian", symbols=["Ag"] * len(coord), coord_bohr=coord,
                       types=[silver_type()], fermi_d=np.array([[12.0]]), fermi_s=np.array([[1.1]]),
                       r0=np.array([[2.885 * TOBOHR]]))
        freq = np.array([3.4 / 27.211386245981])
        gpu_ctx.set_atoms(s)
        aw = s.alpha_table(freq)
        gpu_ctx.set_alpha(aw)
    else:
        s, freq, aw = setup(gpu_ctx, load_fixture(case), 1)
    print(f"{case}: {s.n_atoms} atoms, {nvec} vectors", flush=True)
    rng = np.random.default_rng(nvec)
    x = rng.normal(size=(nvec, s.n_var)) + 1j * rng.normal(size=(nvec, s.n_var))
    gpu_ctx.set_option("fp64_mfma", 1)
    got = gpu_ctx.matvec(freq[0], 0, x)
    gpu_ctx.set_option("fp64_mfma", 0)
    vec = gpu_ctx.matvec(freq[0], 0, x)
    gpu_ctx.set_option("fp64_mfma", 1)
    assert relerr(got, vec) < 1e-13
    if s.n_atoms < 1000:
        ref = pxo.matvec_otf(s, freq[0], aw[:, 0], x[nvec - 1])
        assert relerr(got[nvec - 1], ref) < 1e-12


@pytest.mark.parametrize("case", ["bem_dpcm_accurate_inversion", "bem_ief_accurate_inversion"])
@pytest.mark.parametrize("nvec", [1, 3, 4])
def test_bem_streaming_product_matches_mfma_tiles(gpu_ctx, case, nvec):
    """BEM FP64 products of <= 4 systems take the streaming pass
    (k_bem_gemv, the multi-shift spectra's product); the MFMA tiles form the
    same sums in another order: 1e-13 between the two (the oracle's D and
    S A^-1 are checked by test_bem_matvec_matches_oracle)."""
    fx = load_fixture(case)
    m = mesh_from_fixture(fx)
    freq = np.array([3.2, 3.6]) / 27.211386245981
    gpu_ctx.set_mesh(m)
    gpu_ctx.set_permittivity(pm.silver_etchegoin(freq))
    rng = np.random.default_rng(nvec)
    x = rng.normal(size=(nvec, m.n)) + 1j * rng.normal(size=(nvec, m.n))
    gpu_ctx.set_option("bem_gemv", 1)
    got = gpu_ctx.matvec(freq[1], 1, x)
    gpu_ctx.set_option("bem_gemv", 0)
    ref = gpu_ctx.matvec(freq[1], 1, x)
    gpu_ctx.set_option("bem_gemv", 1)
    assert relerr(got, ref) < 1e-13


def test_bem_streaming_product_row_pairs_on_the_rod():
    """10240 tesserae (the C3 rod): the streaming pass with row pairs
    (16-byte loads) against the MFMA tiles, 1e-13."""
    from plasmonx_amd import Context
    from plasmonx_amd.workloads import c3_meshes, c3_permittivity
    rod, freq = c3_meshes("dpcm")[1], c3_meshes("dpcm")[2][:2]
    ctx = Context(0)
    ctx.set_mesh(rod)
    ctx.set_permittivity(c3_permittivity(freq))
    print(f"rod: {rod.n} tesserae", flush=True)
    x = np.random.default_rng(5).normal(size=(3, rod.n)) + 1j * np.random.default_rng(6).normal(size=(3, rod.n))
    got = ctx.matvec(freq[1], 1, x)
    ctx.set_option("bem_gemv", 0)
    ref = ctx.matvec(freq[1], 1, x)
    ctx.close()
    assert relerr(got, ref) < 1e-13
