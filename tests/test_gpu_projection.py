"""Projected starting iterates in the batched solver: systems admitted coarse
to fine over the frequency grid start from the x in the span of the nearest
converged frequencies of the same field that minimises the residual.  Same
answers as cold starts (both stop on the FP64 true residual), fewer
iterations, and the ABI option turns it off."""
import numpy as np
import pytest

from plasmonx_amd import _lib
from plasmonx_amd.workloads import c2_silver_sphere

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("precision", [_lib.MIXED, _lib.FP64])
def test_projection_same_answer_fewer_iterations(gpu_ctx, precision):
    s, freq = c2_silver_sphere(radius=12.0, n_freq=33)
    gpu_ctx.set_option("kbar_cutoff", 1)
    gpu_ctx.set_atoms(s)
    gpu_ctx.set_alpha(s.alpha_table(freq))
    gpu_ctx.set_option("batch", 12)
    tol = 1e-11 * np.sqrt(s.n_var)
    kw = dict(method=_lib.ITERATIVE_OTF, tol_abs=tol, krylov_dim=120, precision=precision)
    gpu_ctx.set_option("x0_projection", 0)
    cold = gpu_ctx.solve_spectrum(s.n_var, freq, 1e-4, **kw)
    gpu_ctx.set_option("x0_projection", 4)
    warm = gpu_ctx.solve_spectrum(s.n_var, freq, 1e-4, **kw)
    gpu_ctx.set_option("batch", 1)
    assert np.all(warm["true_residual"] < tol) and np.all(cold["true_residual"] < tol)
    for f in range(len(freq)):
        assert relerr(warm["polar"][f], cold["polar"][f]) < 1e-7, f
    assert warm["iters"].sum() < 0.9 * cold["iters"].sum(), (warm["iters"].sum(), cold["iters"].sum())
