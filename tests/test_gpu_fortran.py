"""The Fortran host path: plasmonx_amd/fortran/pxh_driver (module pxh over
iso_c_binding, the call sequence a reference-side maintainer would add) gives
the same polarizabilities as the ctypes binding and the oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_fixture
from oracle import pxo
from plasmonx_amd import _lib
from plasmonx_amd.model import system_from_fixture

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(ROOT, "plasmonx_amd", "fortran", "pxh_driver")


def write_input(path, s, freq, aw, e0, method, tol, kdim, precision, pol_mask=7):
    f = lambda a: " ".join(repr(float(v)) for v in np.ravel(a, order="F"))
    with open(path, "w") as o:
        o.write(f"{s.model_id} {s.kernel_id} {s.n_atoms} {s.n_types}\n")
        o.write(f(s.coord_bohr.T) + "\n")
        o.write(" ".join(str(int(t) + 1) for t in s.type_idx) + "\n")
        for k in ("eta", "r_q", "r_mu", "a_ij", "tau", "density"):
            o.write(f(getattr(s, k)) + "\n")
        for m in (s.fermi_d, s.fermi_s, s.r0):
            o.write(f(np.asarray(m)) + "\n")
        o.write(" ".join(str(int(v)) for v in np.ravel(np.asarray(s.nb_counts).T, order="F")) + "\n")
        o.write(" ".join(str(int(v)) for v in s.nb_total) + "\n")
        o.write(f"{len(freq)}\n{f(freq)}\n")
        if s.model == "wfqfmu":
            o.write(" ".join(f"{float(a.real)!r} {float(a.imag)!r}" for a in np.ravel(aw, order="F")) + "\n")
        o.write(f"{float(e0)!r} {method} {float(tol)!r} {kdim} {precision} {pol_mask}\n")


def read_output(path, nf):
    lines = open(path).read().split("\n")
    st = int(lines[0])
    assert st == 0, lines[1]
    pol = np.zeros((nf, 3, 3), dtype=np.complex128)
    for k in range(nf):
        v = np.array([float(x) for x in lines[1 + k].split()])
        c = v[0::2] + 1j * v[1::2]  # (a, b) column-major: a fastest
        pol[k] = c.reshape(3, 3, order="F")
    return pol


@pytest.mark.parametrize("case,method", [("wfqfmu_ag_inversion", _lib.INVERSION),
                                         ("wfq_sodium_iterative_on_the_fly", _lib.ITERATIVE_OTF)])
def test_fortran_driver_matches_ctypes_and_oracle(tmp_path, case, method):
    if not os.path.exists(DRIVER):
        pytest.skip("pxh_driver not built")
    fx = load_fixture(case)
    s = system_from_fixture(fx)
    freq = np.array(fx["freq_au"])[:4]
    aw = s.alpha_table(freq) if s.model == "wfqfmu" else None
    tol = fx["tol"] * np.sqrt(s.n_var)
    write_input(tmp_path / "in.txt", s, freq, aw, fx["e0"], method, tol, fx["gmres_dim"], _lib.FP64)
    subprocess.run([DRIVER, str(tmp_path / "in.txt"), str(tmp_path / "out.txt")], check=True, timeout=120)
    pol_f = read_output(tmp_path / "out.txt", len(freq))
    with _lib.Context(0) as c:
        c.set_atoms(s)
        if aw is not None:
            c.set_alpha(aw)
        c.set_option("batch", 1)
        got = c.solve_spectrum(s.n_var, freq, fx["e0"], method=method, tol_abs=tol, krylov_dim=fx["gmres_dim"])
    # pol_f[f][a][b] is the reference's polar_w(a, b), as the ctypes binding returns it
    ref = pxo.solve_spectrum(s, freq, fx["e0"], alpha_w=aw, method=0)
    scale = np.abs(ref["polar"]).max()
    assert np.abs(pol_f - got["polar"]).max() <= 1e-12 * scale
    bar = 1e-10 if method == _lib.INVERSION else 1e-5
    assert np.abs(pol_f - ref["polar"]).max() <= bar * scale
