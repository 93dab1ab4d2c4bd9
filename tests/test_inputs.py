"""The YAML reader against the fixtures: reading the reference's own input
files (tests/golden/inputs, copied by tests/golden/copy_inputs.py) gives the
same target, spectrum, field and method as the fixtures hold, and therefore
drives libpxh on exactly what the parity tests check against the logs."""
import os

import numpy as np
import pytest

from conftest import load_fixture
from plasmonx_amd.inputs import freq_to_au, param_to_au, read_input, system_from_case
from plasmonx_amd.model import system_from_fixture

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["wfq_sodium_inversion", "wfqfmu_ag_inversion", "wfqfmu_hetero_agau_inversion", "wfqfmu_ag_diff_units"]
KEYS = ["model", "kernel", "rhs_form", "method", "tol", "gmres_dim", "e0", "freq_au", "types", "fermi_d", "fermi_s",
        "r0", "symbols", "coord_ang"]


def input_path(case):
    return os.path.join(HERE, "golden", "inputs", case, case + ".yaml")


@pytest.mark.parametrize("case", CASES)
def test_reader_matches_fixture(case):
    got = read_input(input_path(case))
    fx = load_fixture(case)
    for k in KEYS:
        assert got[k] == fx[k], k
    a, b = system_from_case(got), system_from_fixture(fx)
    assert np.array_equal(a.coord_bohr, b.coord_bohr) and a.n_var == b.n_var


def test_units():
    """The unit rules against hand values: 1 eV, 500 nm, 1 fs, Angstrom radii."""
    assert abs(freq_to_au("27.211386245981 eV") - 1.0) < 1e-15
    assert abs(freq_to_au("500 nm") - 0.0911267050) < 1e-9  # 2.4797 eV
    assert abs(param_to_au("1 fs", "tau") - 41.3413733351702) < 1e-9
    assert abs(param_to_au("0.529177210544 ang", "rq") - 1.0) < 1e-15


def test_unit_rules_by_hand():
    """Each unit branch of the reference's conversions (validation.py:8-140)
    against values worked out by hand, not through any other reader."""
    assert abs(param_to_au("1 fs-1", "gamma") - 1 / 41.3413733351702) < 1e-15
    assert abs(param_to_au("2e14 s-1", "gamma") - 2e14 / 41.3413733351702e15) < 1e-15
    assert abs(param_to_au("2.0 eV", "eta") - 2.0 / 27.211386245981) < 1e-15
    assert abs(param_to_au("1 ang3", "alpha") - 6.748334) < 1e-6        # (1 / 0.529177)^3
    assert abs(param_to_au("1 nm3", "alpha") - 6748.334) < 1e-3
    assert abs(param_to_au("1 nm", "rq") - 18.897261) < 1e-6
    assert abs(param_to_au("1e7 S/m", "sigma0") - 2.1739) < 1e-12
    assert abs(param_to_au("1 ang2", "a_ij") - 3.571064) < 1e-6
    assert abs(freq_to_au("1 um") - 10000.0 / 219474.6313632) < 1e-15
    assert abs(freq_to_au("219474.6313632 cm-1") - 1.0) < 1e-12
    assert abs(freq_to_au("6579.683920502 THz") - 1.0) < 1e-12
    for bad in (("1 ps", "tau"), ("1 kcal", "eta"), ("1 bohr3", "alpha"), ("1 ms-1", "gamma")):
        with pytest.raises(ValueError):
            param_to_au(*bad)


def test_gamma_gives_tau_as_its_inverse():
    from plasmonx_amd.inputs import drude_parameters
    tau, sigma0 = drude_parameters({"gamma": "0.5 fs-1", "w_p": "9.0 eV"}, "Ag")
    assert abs(tau - 2 * 41.3413733351702) < 1e-12
    assert abs(sigma0 - (9.0 / 27.211386245981) ** 2 * tau / (4 * np.pi)) < 1e-15


@pytest.mark.parametrize("field,n,first,last", [
    ({"min freq": 0.005, "max freq": 1.0, "nfreq": 200}, 200, 0.005, 1.0),
    ({"min freq": 2.5, "max freq": 4.0, "step freq": 0.025}, 61, 2.5, 4.0),
    ({"min freq": 2.5, "step freq": 0.5, "nfreq": 4}, 4, 2.5, 4.0),
    ({"max freq": 4.0, "step freq": 0.5, "nfreq": 4}, 4, 2.5, 4.0),
    ({"min freq": 1.0, "max freq": 2.05, "step freq": 0.5}, 4, 1.0, 2.5),  # ceil, not round
])
def test_frequency_combinations(field, n, first, last):
    from plasmonx_amd.inputs import frequencies
    f = np.array(frequencies(field)) * 27.211386245981
    assert len(f) == n
    assert abs(f[0] - first) < 1e-12 and abs(f[-1] - last) < 1e-12


MINIMAL = ["wfq_sodium_minimal_input", "wfq_graphene_minimal_input", "wfqfmu_ag_minimal_input",
           "wfqfmu_au_minimal_input", "wfqfmu_hetero_agau_minimal_input", "wfqfmu_au_read_xyz"]


def minimal_log(case):
    import json
    with open(os.path.join(HERE, "golden", "inputs", case, "log.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", MINIMAL)
def test_minimal_decks_give_the_logged_parameters(case):
    """The reference's minimal decks (its defaults fill in the atom types and
    parameters): the spectrum and the per-type parameters the reader builds
    equal what the reference's log prints about its input, to the printed
    digits."""
    c = read_input(input_path(case))
    log = minimal_log(case)
    head = log["header"]
    assert len(c["freq_au"]) == int(head[0][1])
    printed = {}
    cur = None
    for k, v in head:
        if k == "AtomType":
            cur = v
            printed[cur] = {}
        elif cur is not None and k not in printed[cur]:
            printed[cur][k] = v.split()[0]
    names = {"Eta": "eta", "Alpha": "alpha", "R_q": "rq", "R_mu": "rmu", "tau": "tau", "sigma_0": "sigma0",
             "A_ij": "a_ij", "fermi_d": "fermi_d", "fermi_s": "fermi_s", "density": "density"}
    for t in c["types"]:
        for k, v in printed[t["name"]].items():
            if k in names:
                got = t[names[k]]
                # half a unit of the last printed digit (F10.3 or E10.3)
                mant, _, ex = v.upper().partition("E")
                half = 0.5 * 10.0 ** (int(ex or 0) - len(mant.split(".")[1]))
                assert abs(got - float(v)) <= half * (1 + 1e-9), (t["name"], k, got, v)
    ev = np.array(c["freq_au"]) * 27.211386245981
    for e in log["results_ev"]:
        assert np.min(np.abs(ev - e)) < 1e-6
