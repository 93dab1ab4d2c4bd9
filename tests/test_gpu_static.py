"""libpxh's static FQ / FQFMu solve (pxh_set_static / pxh_solve_static)
against the oracle and the reference's fq_* / fqfmu_* logs, through the C
ABI.  Bars: charges, dipoles, multipliers and energy 1e-10 relative to the
oracle's solve; the FQ polarizability 1e-10 against the oracle's formula
(explicit J^-1); the logs to their printed digits."""
import json
import os

import numpy as np
import pytest

from oracle import pxo
from plasmonx_amd import _lib, geometry
from plasmonx_amd.inputs import read_static_input, static_system_from_case
from plasmonx_amd.model import TOBOHR

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
STATIC = ["fq_energy", "fq_polarizability", "fq_principal_axis_energy", "fqfmu_energy", "fqfmu_rq_rmu_energy"]


def case(name):
    d = os.path.join(HERE, "golden", "inputs", name)
    c = read_static_input(os.path.join(d, name + ".yaml"))
    with open(os.path.join(d, "log.json")) as f:
        log = json.load(f)
    return c, static_system_from_case(c), log


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("name", STATIC)
def test_static_cases_match_oracle_and_log(gpu_ctx, name):
    c, s, log = case(name)
    gpu_ctx.set_atoms(s)
    gpu_ctx.set_static(s, c["mol_count"])
    gs = gpu_ctx.solve_static(_lib.ENERGY)
    var, _, energy = pxo.static_solve(s, c["mol_count"], 0)
    assert relerr(gs["variables"], var) < 1e-10
    assert abs(gs["energy"] - energy) < 1e-12 * abs(energy)
    if "energy" in log:
        assert abs(gs["energy"] - log["energy"]) <= 0.5e-8 * 1.0001
    if s.model == "fq":
        sf = gpu_ctx.solve_static(_lib.STATIC_FIELD, e0=c["e0"], rhs_form=_lib.RHS_POTENTIAL)
        _, pol, _ = pxo.static_solve(s, c["mol_count"], 1, e0=c["e0"], rhs_form=2)
        assert relerr(sf["polar"], pol) < 1e-10
        if "polar_iso" in log:
            assert abs(np.trace(sf["polar"]) / 3 - log["polar_iso"]) <= 0.5e-10 * 1.0001
    else:
        with pytest.raises(_lib.PxhError) as e:
            gpu_ctx.solve_static(_lib.STATIC_FIELD, e0=1e-3)
        assert e.value.code == _lib.PXH_EINVAL
        sf = gpu_ctx.solve_static(_lib.STATIC_FIELD, e0=1e-3, want_polar=False)
        var, _, _ = pxo.static_solve(s, c["mol_count"], 1, e0=1e-3)
        assert relerr(sf["variables"], var) < 1e-10


@pytest.mark.parametrize("model,n_mol", [("fq", 250), ("fqfmu", 100)])
def test_static_water_cluster_matches_oracle(gpu_ctx, model, n_mol):
    """A seeded cluster of water-like molecules (FQ: 250 molecules, 1000
    unknowns; FQFMu: 100 molecules, 1300 unknowns: several LU panels), one
    multiplier per molecule, against the oracle."""
    c, s0, _ = case("fq_energy" if model == "fq" else "fqfmu_energy")
    rng = np.random.default_rng(250)
    w = np.array(c["coord_ang"][:3])
    centres = []
    while len(centres) < n_mol:
        p = rng.uniform(-20, 20, size=3)
        if all(np.linalg.norm(p - q) > 3.1 for q in centres):
            centres.append(p)
    xyz, sym = [], []
    for p in centres:
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        for k in range(3):
            xyz.append(p + (w[k] - w[0]) @ q)
            sym.append(c["symbols"][k])
    big = dict(c, symbols=sym, coord_ang=[list(v) for v in xyz], mol_count=[3] * n_mol)
    s = static_system_from_case(big)
    print(f"{model}: {s.n_atoms} atoms", flush=True)
    gpu_ctx.set_atoms(s)
    gpu_ctx.set_static(s, big["mol_count"])
    gs = gpu_ctx.solve_static(_lib.ENERGY)
    var, _, energy = pxo.static_solve(s, big["mol_count"], 0)
    assert relerr(gs["variables"], var) < 1e-9
    assert abs(gs["energy"] - energy) < 1e-10 * abs(energy)
    if model == "fq":
        sf = gpu_ctx.solve_static(_lib.STATIC_FIELD)
        _, pol, _ = pxo.static_solve(s, big["mol_count"], 1)
        assert relerr(sf["polar"], pol) < 1e-9


def test_static_target_refuses_dynamic_calls(gpu_ctx):
    c, s, _ = case("fq_energy")
    gpu_ctx.set_atoms(s)
    with pytest.raises(_lib.PxhError):
        gpu_ctx.solve_spectrum(s.n_var, np.array([0.1]), 1e-4, method=_lib.INVERSION)
    with pytest.raises(_lib.PxhError):
        gpu_ctx.solve_static(_lib.ENERGY)  # pxh_set_static not called yet
