"""Reference inputs driven end to end through libpxh and written in the
reference's formats (plasmonx_amd.output), compared with the reference's
logs the way its regression tests compare them: the numbers of the lines
from ' I am cleaning' to 'Maxima Analysis' within a relative 1e-3
(tests/wfq_sodium_inversion/test), and, for a restart, the 'Recovering from
backup frequency' line exactly and the maxima analysis within an absolute
1e-8 (tests/wfq_sodium_restart/test)."""
import json
import os
import re
import shutil

import numpy as np
import pytest

from conftest import load_fixture
from plasmonx_amd import _lib, output
from plasmonx_amd.inputs import read_input, system_from_case

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "inputs")
NUM = re.compile(r"[-+]?\d+\.\d*(?:[EeDd][-+]?\d+)?")


def numbers(lines):
    return [float(x.replace("D", "E")) for ln in lines for x in NUM.findall(ln)]


def region(text, start, stop=None, n_lines=None):
    lines = text.splitlines()
    k = next(i for i, ln in enumerate(lines) if start in ln)
    if n_lines is not None:
        return lines[k:k + n_lines]
    e = next(i for i in range(k, len(lines)) if stop in lines[i])
    return lines[k:e]


def close(got, ref, rel=None, abs_=None):
    got, ref = np.array(got), np.array(ref)
    assert got.shape == ref.shape
    if rel is not None:
        return bool(np.all(np.abs(got - ref) <= rel * np.abs(ref)))
    return bool(np.all(np.abs(got - ref) <= abs_))


def solve(ctx, c, freq):
    s = system_from_case(c)
    ctx.set_option("kbar_cutoff", 0)
    ctx.set_atoms(s)
    return ctx.solve_spectrum(s.n_var, freq, c["e0"], method=_lib.INVERSION)["results"]


def test_inversion_log_tail_matches_reference(gpu_ctx, tmp_path):
    c = read_input(os.path.join(INPUTS, "wfq_sodium_inversion", "wfq_sodium_inversion.yaml"))
    freq = np.array(c["freq_au"])
    res = solve(gpu_ctx, c, freq)
    text = output.log_tail(freq, res)
    # the whole region against every result the log prints (fixture), and
    # the verbatim excerpt (first blocks and the maxima block)
    fx = load_fixture("wfq_sodium_inversion")
    ref_all = output.log_tail(freq, np.array(fx["expected"]["results"]))
    assert close(numbers(region(text, " I am cleaning", "Maxima Analysis")),
                 numbers(region(ref_all, " I am cleaning", "Maxima Analysis")), rel=1e-3)
    excerpt = open(os.path.join(INPUTS, "wfq_sodium_inversion", "log_tail.txt")).read()
    head = excerpt.splitlines()[:5 + 3 * 23]
    assert close(numbers(text.splitlines()[:len(head)]), numbers(head), rel=1e-3)
    assert close(numbers(region(text, "Maxima Analysis", n_lines=8)),
                 numbers(region(excerpt, "Maxima Analysis", n_lines=8)), rel=1e-3)
    csv = output.write_csv(str(tmp_path / "wfq_sodium_inversion.log"), freq, res)
    assert np.loadtxt(csv, comments="#").shape == (len(freq), 21)


def test_restart_recovers_backups_and_matches_reference(gpu_ctx, tmp_path):
    d = os.path.join(INPUTS, "wfq_sodium_restart")
    c = read_input(os.path.join(d, "wfq_sodium_restart.yaml"))
    freq = np.array(c["freq_au"])
    assert len(freq) == 11
    log = str(tmp_path / "wfq_sodium_restart.log")
    for name in os.listdir(d):
        if name.endswith(".plasmonX.bk"):
            shutil.copy(os.path.join(d, name), tmp_path)
    todo, rec = output.restart_split(log, freq)
    assert list(rec) == [0]
    res = np.zeros((len(freq), 20))
    res[todo] = solve(gpu_ctx, c, freq[todo])
    for i, row in rec.items():
        res[i] = row
    for i in todo:  # each solved frequency leaves its backup, as the reference does
        output.write_bk(log, freq[i], res[i])
    text = output.log_tail(freq, res, recovered=sorted(rec))
    excerpt = open(os.path.join(d, "log_tail.txt")).read()
    assert " Recovering from backup frequency: 2.00000 eV." in text.splitlines()
    assert close(numbers(region(text, " Maxima Analysis", n_lines=5)),
                 numbers(region(excerpt, " Maxima Analysis", n_lines=5)), abs_=1e-8)
    with open(os.path.join(d, "log.json")) as f:
        expected = np.array(json.load(f)["results"])
    assert close(res, expected, rel=1e-3)
    todo2, rec2 = output.restart_split(log, freq)  # a second run recovers everything
    assert todo2 == [] and len(rec2) == 11
