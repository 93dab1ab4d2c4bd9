"""pxh_induced_field and pxh_plasmon_density (HIP) against the oracle and
the reference's cubes.

Bars: the GPU field equals the oracle's to 1e-12 relative (FP64 on both
sides, summation order aside); |E|/E0 at the sampled grid points equals the
reference's printed cube values to their 6 digits; on a 15683-atom omegaFQFmu
sphere (the C2 target) with random variables the field at 64 points matches
the oracle, and the throughput of a full 100k-point grid is reported.  The
density is held to the same bars (real and imaginary cubes, parts summing to
the total).  The planes of the reference's plane tests are computed whole and
must give its CSV values (16 digits) to 1e-10, and the effective volume of
its field boxes the printed figures.
"""
import time

import numpy as np
import pytest

from oracle import pxo
from plasmonx_amd import Context
from plasmonx_amd.analysis import cubic_grid, field_enhancement, plasmon_density
from plasmonx_amd.model import TOBOHR
from test_analysis_golden import CASES, DENSITY_CASES, half_ulp, load, system_from_analysis_fixture
from test_analysis_planes import (PLANE_CASES, VOLUME_CASES, check_plane, check_volume, plane_setup, plane_values,
                                  volume_setup)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CASES)
def test_field_matches_oracle_and_cube(case):
    fx = load(case)
    s = system_from_analysis_fixture(fx)
    _, _, pts = cubic_grid(fx["coord_ang"], fx["grid"]["n"], fx["grid_offset_ang"])
    with Context(0) as ctx:
        ctx.set_atoms(s)
        for d, cube in fx["cube"].items():
            i_pol = "xyz".index(d)
            var = np.array([complex(a, b) for a, b in fx["variables"][d]])
            # the whole 30^3 grid against the oracle
            got = ctx.induced_field(pts * TOBOHR, var, i_pol, fx["e0"])
            ref = pxo.induced_field(s, pts * TOBOHR, var, i_pol, fx["e0"])
            err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
            print(f"{case} E{d}: {pts.shape[0]} points, max relative difference to the oracle {err:.1e}", flush=True)
            assert err < 1e-12
            # the sampled cube values, to the printed digits
            idx = np.array(cube["idx"])
            enh = field_enhancement(ctx, pts[idx], var, i_pol, fx["e0"])
            val = np.array(cube["values"])
            half_ulp = 0.5 * 10.0 ** (np.floor(np.log10(np.abs(val))) + 1 - 6)
            assert np.all(np.abs(enh - val) <= 1.01 * half_ulp)


def test_field_large_target():
    from plasmonx_amd.workloads import c2_silver_sphere
    s, _ = c2_silver_sphere(n_freq=2)
    rng = np.random.default_rng(5)
    var = rng.normal(size=s.n_var) + 1j * rng.normal(size=s.n_var)
    var[: s.n_atoms] -= var[: s.n_atoms].mean()  # neutral
    lo = s.coord_bohr.min(axis=0) - 10.0
    hi = s.coord_bohr.max(axis=0) + 10.0
    pts = lo + rng.random((100_000, 3)) * (hi - lo)
    with Context(0) as ctx:
        ctx.set_atoms(s)
        ctx.induced_field(pts[:256], var, 0, 1e-4)  # warm-up
        t = time.perf_counter()
        got = ctx.induced_field(pts, var, 0, 1e-4)
        dt = time.perf_counter() - t
    pairs = pts.shape[0] * s.n_atoms
    print(f"15683-atom omegaFQFmu target, {pts.shape[0]} points: {dt * 1e3:.1f} ms with the copies, "
          f"{pairs / dt / 1e9:.1f} G point-atom pairs/s", flush=True)
    sel = np.arange(0, pts.shape[0], pts.shape[0] // 64)
    ref = pxo.induced_field(s, pts[sel], var, 0, 1e-4)
    err = np.max(np.abs(got[sel] - ref)) / np.max(np.abs(ref))
    assert err < 1e-11


@pytest.mark.parametrize("case", DENSITY_CASES)
def test_density_matches_oracle_and_cube(case):
    fx = load(case)
    s = system_from_analysis_fixture(fx)
    _, _, pts = cubic_grid(fx["coord_ang"], fx["grid"]["n"], fx["grid_offset_ang"])
    with Context(0) as ctx:
        ctx.set_atoms(s)
        for d, cube in fx["cube"].items():
            var = np.array([complex(a, b) for a, b in fx["variables"][d]])
            got, gq, gm = ctx.plasmon_density(pts * TOBOHR, var, parts=True)
            ref, rq, rm = pxo.plasmon_density(s, pts * TOBOHR, var)
            scale = np.max(np.abs(ref))
            err = max(np.max(np.abs(a - b)) for a, b in ((got, ref), (gq, rq), (gm, rm))) / scale
            print(f"{case} E{d}: {pts.shape[0]} points, density max relative difference to the oracle {err:.1e}",
                  flush=True)
            assert err < 1e-12
            assert np.array_equal(ctx.plasmon_density(pts * TOBOHR, var), got)
            idx = np.array(cube["idx"])
            rho = plasmon_density(ctx, pts[idx], var)
            for part, v in (("re", rho.real), ("im", rho.imag)):
                val = np.array(cube[part])
                assert np.all(np.abs(v - val) <= 1.01 * half_ulp(val))


def test_density_large_target():
    from plasmonx_amd.workloads import c2_silver_sphere
    s, _ = c2_silver_sphere(n_freq=2)
    rng = np.random.default_rng(6)
    var = rng.normal(size=s.n_var) + 1j * rng.normal(size=s.n_var)
    lo = s.coord_bohr.min(axis=0) - 10.0
    hi = s.coord_bohr.max(axis=0) + 10.0
    pts = lo + rng.random((100_000, 3)) * (hi - lo)
    with Context(0) as ctx:
        ctx.set_atoms(s)
        ctx.plasmon_density(pts[:256], var)  # warm-up
        t = time.perf_counter()
        got = ctx.plasmon_density(pts, var)
        dt = time.perf_counter() - t
    print(f"15683-atom omegaFQFmu target, {pts.shape[0]} points: density {dt * 1e3:.1f} ms with the copies, "
          f"{pts.shape[0] * s.n_atoms / dt / 1e9:.1f} G point-atom pairs/s", flush=True)
    sel = np.arange(0, pts.shape[0], pts.shape[0] // 64)
    ref = pxo.plasmon_density(s, pts[sel], var)[0]
    assert np.max(np.abs(got[sel] - ref)) / np.max(np.abs(ref)) < 1e-11


@pytest.mark.parametrize("case", PLANE_CASES)
def test_planes_match_reference_csv(case):
    """Every plane of the reference's plane tests, whole, on the GPU: the
    sampled CSV values to 1e-10 and the whole plane to the oracle."""
    fx = load(case)
    s = system_from_analysis_fixture(fx)
    with Context(0) as ctx:
        ctx.set_atoms(s)
        for pl in fx["planes"]:
            pts = plane_setup(fx, pl)
            got = plane_values(fx, pl, pts, ctx.induced_field, ctx.plasmon_density)
            ref = plane_values(fx, pl, pts, lambda p, v, i, e0: pxo.induced_field(s, p, v, i, e0),
                               lambda p, v: pxo.plasmon_density(s, p, v)[0])
            for part in got:
                assert np.max(np.abs(got[part] - ref[part])) <= 1e-12 * np.max(np.abs(ref[part]))
            idx = np.array(pl["idx"])
            check_plane(fx, pl, {k: v[idx] for k, v in got.items()})


@pytest.mark.parametrize("case", VOLUME_CASES)
def test_effective_volume_gpu(case):
    fx = load(case)
    s = system_from_analysis_fixture(fx)
    lo, hi, step, pts = volume_setup(fx)
    d = fx["field_dir"]
    var = np.array([complex(a, b) for a, b in fx["variables"][d]])
    with Context(0) as ctx:
        ctx.set_atoms(s)
        f = field_enhancement(ctx, pts, var, "xyz".index(d), fx["e0"])
    check_volume(fx, lo, hi, step, pts, f)


def test_bem_field_cube_and_plane():
    """pxh_bem_induced_field on the whole BEM cube and plane: equal to the
    oracle (zeros inside included) and to the reference's printed values."""
    from test_analysis_bem import bem_case, bem_cube_points, bem_plane_points
    with Context(0) as ctx:
        fx, c, q, i_pol = bem_case("analysis_field_3d_y_bem")
        _, _, pts = bem_cube_points(fx)
        got = ctx.bem_induced_field(c, q, pts * TOBOHR, i_pol, fx["e0"])
        ref = pxo.bem_induced_field(c, q, pts * TOBOHR, i_pol, fx["e0"])
        assert np.array_equal(got == 0, ref == 0)
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
        cube = next(iter(fx["cube"].values()))
        idx = np.array(cube["idx"])
        f = np.sqrt((np.abs(got[idx]) ** 2).sum(axis=1)) / fx["e0"]
        val = np.array(cube["values"])
        assert np.all(np.abs(f - val) <= 1.01 * half_ulp(val))

        fx, c, q, i_pol = bem_case("analysis_field_2d_y_xy_bem")
        pl = fx["planes"][0]
        pts = bem_plane_points(fx, pl)
        got = ctx.bem_induced_field(c, q, pts * TOBOHR, i_pol, fx["e0"])
        f = np.sqrt((np.abs(got[np.array(pl["idx"])]) ** 2).sum(axis=1)) / fx["e0"]
        assert np.allclose(f, pl["values"], rtol=1e-10, atol=0)


@pytest.mark.parametrize("case", ["analysis_field_3d_x_y_z_wfqfmu_ag_plt", "analysis_density_3d_y_wfqfmu_au_plt"])
def test_plt_gpu(case, tmp_path):
    from test_analysis_planes import check_plt_case
    fx = load(case)
    s = system_from_analysis_fixture(fx)
    with Context(0) as ctx:
        ctx.set_atoms(s)
        check_plt_case(fx, tmp_path, ctx.induced_field, ctx.plasmon_density)
