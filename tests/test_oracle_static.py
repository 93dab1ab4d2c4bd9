"""The oracle's static FQ / FQFMu restatement (oracle/pxo_static.c) against
the reference's fq_* / fqfmu_* logs (copied into tests/golden/inputs by
tests/golden/copy_inputs.py) and against the closed form of the constrained
solve.  The YAML inputs are read by plasmonx_amd.inputs.read_static_input.

Printed precision: matrices E11.4 (half a unit of the 4th digit), charges,
dipoles and right-hand sides E14.6, energies F14.8, the isotropic
polarizability F19.10."""
import json
import os

import numpy as np
import pytest

from oracle import pxo
from plasmonx_amd.inputs import read_static_input, static_system_from_case

HERE = os.path.dirname(os.path.abspath(__file__))
STATIC = ["fq_energy", "fq_polarizability", "fq_principal_axis_energy", "fqfmu_energy", "fqfmu_rq_rmu_energy"]


def case(name):
    d = os.path.join(HERE, "golden", "inputs", name)
    c = read_static_input(os.path.join(d, name + ".yaml"))
    with open(os.path.join(d, "log.json")) as f:
        log = json.load(f)
    return c, static_system_from_case(c), log


def close_e(got, printed, digits, noise=1e-12):
    """|got - printed| within half a unit of the last of `digits` printed
    significant digits (E format); printed values below `noise` are rounding
    noise of a zero (the reference prints e.g. 0.1776E-14) and need |got| <
    noise."""
    got, printed = np.asarray(got, dtype=float), np.asarray(printed, dtype=float)
    safe = np.where(printed != 0, np.abs(printed), 1.0)
    tol = 0.5 * 10.0 ** (np.floor(np.log10(safe)) + 1 - digits) * 1.0001
    small = np.abs(printed) < noise
    return np.where(small, np.abs(got) < noise, np.abs(got - printed) <= tol)


@pytest.mark.parametrize("name", [n for n in STATIC if n != "fq_polarizability"])
def test_ground_state_matches_log(name):
    c, s, log = case(name)
    var, _, energy = pxo.static_solve(s, c["mol_count"], 0)
    n = s.n_atoms
    assert abs(energy - log["energy"]) <= 0.5e-8 * 1.0001
    q = var[:n, 0]
    assert close_e(q, [r[0] for r in log["Charges"]], 6).all()
    if "Dipoles" in log:
        mu = var[n + len(c["mol_count"]):, 0]
        assert close_e(mu, [r[0] for r in log["Dipoles"]], 6).all()
    for title in ("FQ Matrix", "FQFMu Matrix"):
        if title in log:
            A = pxo.static_matrix(s, c["mol_count"])
            for r, col, v in log[title]:
                assert close_e(A[r - 1, col - 1], v, 4), (r, col)
    if "RHS" in log:
        b = np.zeros(s.n_atoms + len(c["mol_count"]) + (3 * n if s.model == "fqfmu" else 0))
        b[:n] = -s.arr("chi")[s.type_idx]
        assert close_e(b, [r[0] for r in log["RHS"]], 6).all()


def test_fq_polarizability_matches_log():
    c, s, log = case("fq_polarizability")
    A = pxo.static_matrix(s, c["mol_count"])
    for r, col, v in log["FQ Matrix"]:
        assert close_e(A[r - 1, col - 1], v, 4), (r, col)
    n = s.n_atoms
    Jinv = np.linalg.inv(A[:n, :n])
    for r, col, v in log["JMatrix-1"]:
        assert close_e(Jinv[r - 1, col - 1], v, 4), (r, col)
    var, pol, _ = pxo.static_solve(s, c["mol_count"], 1, e0=c["e0"], rhs_form=2)
    assert c["e0"] == 0.0 and np.all(var == 0.0)  # the static field's intensity is 0 for 'static response'
    printed = np.zeros((3, 3))
    for r, col, v in log["polarizability tensor (a.u.)"]:
        printed[r - 1, col - 1] = v
    assert close_e(pol, printed, 4).all()
    assert abs(np.trace(pol) / 3 - log["polar_iso"]) <= 0.5e-10 * 1.0001


@pytest.mark.parametrize("n_mol_split", [1, 2])
def test_fq_polarizability_formula_equals_the_constrained_solve(n_mol_split):
    """The reference's formula alpha = R J^-1 R^T - (L^T J^-1 R)^T (L^T J^-1
    L)^-1 (L^T J^-1 R) (explicit J^-1) equals -R^T q of the saddle-point solve
    at unit field, and each molecule's charges sum to zero."""
    c, s, log = case("fq_polarizability")
    mc = c["mol_count"] if n_mol_split == 2 else [s.n_atoms]
    _, pol, _ = pxo.static_solve(s, mc, 1, e0=0.0)
    var, _, _ = pxo.static_solve(s, mc, 1, e0=1.0, rhs_form=1)
    q = var[:s.n_atoms]
    alpha = -(s.coord_bohr.T @ q)
    assert np.abs(alpha - pol).max() < 1e-10 * np.abs(pol).max()
    first = 0
    for k in mc:
        assert np.abs(q[first:first + k].sum(axis=0)).max() < 1e-12
        first += k


def test_potential_form_changes_only_the_multipliers():
    """'potential' shifts the charge rows' potential by a constant per
    direction, which the Lagrange multipliers absorb: same charges."""
    c, s, _ = case("fq_polarizability")
    vf, _, _ = pxo.static_solve(s, c["mol_count"], 1, e0=1e-3, rhs_form=1)
    vp, _, _ = pxo.static_solve(s, c["mol_count"], 1, e0=1e-3, rhs_form=2)
    n = s.n_atoms
    assert np.abs(vf[:n] - vp[:n]).max() < 1e-12 * np.abs(vf[:n]).max()
    assert np.abs(vf[n:] - vp[n:]).max() > 1e-6
