"""The reference's result formats (plasmonx_amd.output) against the
reference's own files, byte for byte where the reference's text is kept as a
fixture: the per-frequency backup (tests/golden/inputs/wfq_sodium_restart/
*.plasmonX.bk), the result blocks and the maxima analysis of the
wfq_sodium_inversion and wfq_sodium_restart logs (log_tail.txt, verbatim
excerpts copied by tests/golden/copy_inputs.py)."""
import json
import os
import shutil

import numpy as np

from conftest import load_fixture
from plasmonx_amd import output

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "inputs")


def test_fortran_e_format():
    assert output.fort_e(-18646.1, 14, 6) == " -0.186461E+05"
    assert output.fort_e(0.0, 14, 6) == "  0.000000E+00"
    assert output.fort_e(-0.0, 14, 6) == " -0.000000E+00"
    assert output.fort_e(9.99999951, 14, 6) == "  0.100000E+02"  # rounding carries into the exponent
    assert output.fort_e(0.0735, 9, 3) == "0.735E-01"
    assert output.fort_e(1.5e-12, 12, 5) == " 0.15000E-11"


def test_backup_file_round_trip_is_byte_exact(tmp_path):
    """Reading the reference's backup and writing it again gives the same
    bytes (header E9.3 triple, 20 results E14.6, sticks)."""
    src = os.path.join(INPUTS, "wfq_sodium_restart", "wfq_sodium_restart-2.00000.plasmonX.bk")
    (f_au, f_nm, f_ev), row = output.read_bk(src)
    assert f_ev == 2.0
    log = str(tmp_path / "wfq_sodium_restart.log")
    path = output.write_bk(log, 2.0 / output.AU_TO_EV, row)
    assert os.path.basename(path) == "wfq_sodium_restart-2.00000.plasmonX.bk"
    assert open(path).read() == open(src).read()


def test_restart_reads_existing_backups(tmp_path):
    log = str(tmp_path / "wfq_sodium_restart.log")
    shutil.copy(os.path.join(INPUTS, "wfq_sodium_restart", "wfq_sodium_restart-2.00000.plasmonX.bk"), tmp_path)
    freq = np.linspace(2.0, 3.0, 11) / output.AU_TO_EV
    todo, rec = output.restart_split(log, freq)
    assert todo == list(range(1, 11)) and list(rec) == [0]
    assert rec[0][0] == -0.186461e5


def test_result_blocks_and_maxima_match_the_log_text():
    """wfq_sodium_inversion: the clean-up message, the first three result
    blocks and the maxima analysis, from the log's own printed values, equal
    the log's text line for line."""
    fx = load_fixture("wfq_sodium_inversion")
    freq = np.array(fx["freq_au"])
    res = np.array(fx["expected"]["results"])
    ref = open(os.path.join(INPUTS, "wfq_sodium_inversion", "log_tail.txt")).read().splitlines()
    got = output.log_tail(freq, res).splitlines()
    head = 5 + 3 * 23
    assert got[:head] == ref[:head]
    m = next(i for i, ln in enumerate(got) if ln.startswith(" Maxima Analysis"))
    assert got[m:] == ref[head:]


def test_csv_layout(tmp_path):
    fx = load_fixture("wfq_sodium_inversion")
    freq = np.array(fx["freq_au"])[:5]
    res = np.array(fx["expected"]["results"])[:5]
    path = output.write_csv(str(tmp_path / "case.log"), freq, res)
    lines = open(path).read().splitlines()
    assert lines[3] == "#      1        Frequency (eV)             "
    assert lines[12] == "#      10       Iso. Abs. Cross. Sect.     "
    assert lines[24] == "#------------------------------------------"
    data = np.loadtxt(path, comments="#")
    assert data.shape == (5, 21)
    assert np.allclose(data[:, 0], freq * output.AU_TO_EV, rtol=1e-5)
    assert np.allclose(data[:, 1:], res, rtol=1e-6)
    assert all(len(ln) == 1 + 12 + 20 * 17 for ln in lines[25:])


def test_freq_file_layout(tmp_path):
    v = np.array([[1 + 2j, -3e-5 + 0j], [0j, 1j], [2.5 + 0j, 0j]])
    path = output.write_freq(str(tmp_path / "case.log"), 2.0 / output.AU_TO_EV, v, "wfq")
    assert os.path.basename(path) == "case-2.0000000000.freq"
    lines = open(path).read().splitlines()
    assert lines[0] == "complex Charges: x. freq = " + f"{2.0:25.16f}"
    assert lines[1] == f"{1:8d}     {1.0:25.16f}   {2.0:25.16f}   "
    assert len(lines) == 3 * 3
