"""pxh_solve_group: the dense inversion of many targets at once (targets of
similar size padded to a common n and factorised in the same LU launches),
against each target solved alone by pxh_solve_spectrum and against the
oracle's LU.  Bars: 1e-12 against the lone solve (same kernels, padded with
an identity block), 1e-10 against the oracle."""
import numpy as np
import pytest

from oracle import pxo
from plasmonx_amd import Context, _lib, geometry
from plasmonx_amd.model import TOBOHR, AtomSystem
from plasmonx_amd.workloads import c4_flakes, silver_type

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def ag_sphere(radius):
    coord = geometry.fcc_sphere(4.09, radius) * TOBOHR
    return AtomSystem(model="wfqfmu", kernel="gaussian", symbols=["Ag"] * len(coord), coord_bohr=coord,
                      types=[silver_type()], fermi_d=np.array([[12.0]]), fermi_s=np.array([[1.1]]),
                      r0=np.array([[2.885 * TOBOHR]]))


def test_group_matches_lone_solves_and_oracle():
    flakes, freq = c4_flakes(8, hi=1500)
    targets = flakes + [ag_sphere(8.0), ag_sphere(9.0)]  # two omegaFQFmu spheres: another group
    pick = freq[[5, 40, 90, 127]]
    print("group sizes:", [t.n_var for t in targets], flush=True)
    ctxs = []
    for t in targets:
        c = Context(0)
        c.set_atoms(t)
        if t.model == "wfqfmu":
            c.set_alpha(t.alpha_table(pick))
        ctxs.append(c)
    out = _lib.solve_group(ctxs, pick, 1e-4, pol_mask=3)
    for t, c, o in zip(targets, ctxs, out):
        alone = c.solve_spectrum(t.n_var, pick, 1e-4, pol_mask=3, method=_lib.INVERSION)
        assert relerr(o["polar"], alone["polar"]) < 1e-12
        assert relerr(o["results"], alone["results"]) < 1e-12
    for t, o in list(zip(targets, out))[::3]:
        ref = pxo.solve_spectrum(t, pick, 1e-4, alpha_w=t.alpha_table(pick) if t.model == "wfqfmu" else None,
                                 pol_mask=3, method=0)
        assert relerr(o["polar"], ref["polar"]) < 1e-10
    for c in ctxs:
        c.close()


def test_group_rejects_mixed_inputs():
    flakes, freq = c4_flakes(2, hi=800)
    a, b = Context(0), Context(0)
    a.set_atoms(flakes[0])
    with pytest.raises(_lib.PxhError):
        _lib.solve_group([a, b], freq[:2], 1e-4)  # b has no target
    with pytest.raises(_lib.PxhError):
        _lib.solve_group([a, a], freq[:2], 1e-4)  # the same context twice
    a.close()
    b.close()
