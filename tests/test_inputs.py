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
