"""Dense inversion at sizes the FP64 oracle cannot factorise in seconds.

The LU path chooses its panel kernel by the rows below the panel: panels held
in registers at up to 1, 2, 4, 8, 12, 16, 20 and 24 rows per work-item (up to
12288 rows), and the global-memory panel above that.  These rods take their
first panels through the 12-, 16-, 20- and 24-row kernels and the
global-memory one.  The check is
size-independent: the residual of A(w) x = b, with A(w) x from the FP64
on-the-fly product (itself checked against the oracle elsewhere) and b the
field right-hand side.  A backward-stable FP64 solve leaves a relative
residual near 1e-13 here; the bar is 1e-10.
"""
import numpy as np
import pytest

from plasmonx_amd import Context, _lib
from plasmonx_amd.workloads import ev, sodium_rod

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("length,radius", [(64.0, 25.0), (85.0, 25.0), (92.0, 26.0), (110.0, 27.0), (118.0, 28.0)],
                         ids=["4891-atoms", "7081-atoms", "8453-atoms", "11095-atoms", "12959-atoms"])
def test_large_inversion_residual(length, radius):
    s = sodium_rod(length, radius)
    freq = ev(np.array([3.0, 3.6]))
    e0 = 1e-4
    ctx = Context(0)
    try:
        ctx.set_atoms(s)
        out = ctx.solve_spectrum(s.n_var, freq, e0, pol_mask=3, method=_lib.INVERSION, want_variables=True)
        for f in range(len(freq)):
            b = ctx.dynamic_rhs(s.n_var, freq[f], e0)
            ax = ctx.matvec(freq[f], f, out["variables"][f][:2])
            for k in range(2):
                res = np.linalg.norm(ax[k] - b[k]) / np.linalg.norm(b[k])
                print(f"n {s.n_var} w {f} field {k}: relative residual {res:.2e}", flush=True)
                assert res < 1e-10
    finally:
        ctx.close()
