"""Frequency-dependent material inputs: Etchegoin permittivity / interband
polarizability of Ag and Au (src/parameters/parameters_module.f90:107-194,
295-375, with the published Etchegoin-Le Ru-Meyer fit parameters), the
permittivity CSV reader (:384-462) and writer, the Lorentz-Drude silver
interband polarizability of the omegaFQFmu spectrum and a Brendel-Bormann
permittivity (not implemented by the reference, :223-225).
Host-side preparation of the per-frequency tables passed to libpxh.
"""
import math

import numpy as np

TOBOHR = 1.8897261254578281
AU_TO_EV = 27.211386245981
AU_TO_CM = 219474.63136320


def _freq3(freq_au):
    nm = 10000000.0 / (np.asarray(freq_au, dtype=np.float64) * AU_TO_CM)
    return 1.0 / nm


def _cp(aj, lj, gj, phi, f3):
    return aj * lj * (np.exp(1j * phi) / ((lj - f3) - 1j * gj) + np.exp(-1j * phi) / ((lj + f3) + 1j * gj))


def silver_etchegoin(freq_au, polar=False):
    f3 = _freq3(freq_au)
    A = (0.1632212, 0.9672775, 0.6150938)
    PH = (-1.709221, 0.1546664, -1.782307)
    L = (1 / 307.9550, 1 / 260.7301, 1 / 236.2128)
    G = (1 / 4277.721, 1 / 1051.474, 1 / 1008.622)
    s = sum(_cp(A[k], L[k], G[k], PH[k], f3) for k in range(3))
    n0 = 10.49 / 107.8682 * 6.022e23 / (1.0e8 * TOBOHR) ** 3
    wp2 = 4.0 * math.pi * n0
    if polar:
        return (1j * 0.2450325 + s) / wp2
    lam_p, gam_p = 137.9643, 1 / 74484.21
    return (1.469322 + 0.2450325j) - 1.0 / (lam_p ** 2 * f3 * (f3 + 1j * gam_p)) + s


def gold_etchegoin(freq_au, polar=False):
    f3 = _freq3(freq_au)
    A = (1.27, 1.10)
    PH = (-math.pi / 4, -math.pi / 4)
    L = (1 / 470.0, 1 / 325.0)
    G = (1 / 1900.0, 1 / 1060.0)
    s = sum(_cp(A[k], L[k], G[k], PH[k], f3) for k in range(2))
    n0 = 19.320 / 196.96657 * 6.022e23 / (1.0e8 * TOBOHR) ** 3
    wp2 = 4.0 * math.pi * n0
    if polar:
        v = s / wp2
        return np.where(v.imag <= 0.0, v.real + 1e-10j, v)
    lam_p, gam_p = 143.0, 1 / 14500.0
    return 1.54 - 1.0 / (lam_p ** 2 * f3 * (f3 + 1j * gam_p)) + s


def read_table(path, freq_au):
    """CSV 'freq(eV) Re Im' with leading '#' comments; the value at each
    requested frequency (|dE| <= 1e-8 eV), else an error."""
    rows = []
    for line in open(path):
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        p = s.replace(",", " ").split()
        rows.append((float(p[0]), float(p[1]), float(p[2])))
    rows = np.array(rows)
    out = np.empty(len(freq_au), dtype=np.complex128)
    for k, f in enumerate(freq_au):
        ev = f * AU_TO_EV
        hit = np.nonzero(np.abs(ev - rows[:, 0]) <= 1.0e-8)[0]
        if not len(hit):
            raise ValueError(f"no data for freq = {ev:.6f} eV in {path}")
        out[k] = rows[hit[0], 1] + 1j * rows[hit[0], 2]
    return out


# Lorentz-Drude silver (A.D. Rakic et al., Appl. Opt. 37 (1998) 5271,
# Table 1): plasma energy and the interband oscillators (f_j, w_j, Gamma_j)
# in eV; the intraband (Drude) term f0 = 0.845, Gamma0 = 0.048 eV is the
# omegaFQ part of the model and not part of the interband table.
LD_SILVER = dict(wp_ev=9.01, osc=[(0.065, 0.816, 3.886), (0.124, 4.481, 0.452), (0.011, 8.185, 0.065),
                                  (0.840, 9.083, 0.916), (5.646, 20.29, 2.419)])
# conduction-electron (atomic) density of silver: 10.49 g/cm3, 107.8682 g/mol, in bohr^-3
AG_DENSITY = 10.49 / 107.8682 * 6.022e23 / (1.0e8 * TOBOHR) ** 3


def lorentz_drude_interband(freq_au, wp_ev=LD_SILVER["wp_ev"], osc=LD_SILVER["osc"], density=AG_DENSITY):
    """Interband polarizability per atom of a Lorentz-Drude metal:
    alpha(w) = eps_ib(w) / (4 pi n0), eps_ib = sum_j f_j wp^2 / (w_j^2 - w^2
    - i Gamma_j w), the form of the omegaFQFmu dipole term that the
    Etchegoin 'polar' variant takes (parameters_module.f90:107-194)."""
    w = np.asarray(freq_au, dtype=np.float64) * AU_TO_EV
    eib = np.zeros(w.shape, dtype=np.complex128)
    for f, w0, g in osc:
        eib = eib + f * wp_ev ** 2 / (w0 ** 2 - w ** 2 - 1j * g * w)
    return eib / (4.0 * math.pi * density)


def write_table(path, freq_au, values):
    """A 'freq (eV)  Re  Im' table in the layout the reference reads for
    'wfqfmu file' / 'permittivity file' (parameters_module.f90:384-462): one
    '#' comment line, then list-directed rows with the frequency written to
    full precision so that the 1e-8 eV match finds every row."""
    with open(path, "w") as fh:
        fh.write("# freq (eV)   Re   Im\n")
        for f, v in zip(np.asarray(freq_au) * AU_TO_EV, np.asarray(values)):
            fh.write(f"{float(f)!r:>28}  {float(v.real)!r:>28}  {float(v.imag)!r:>28}\n")


def faddeeva(z, N=32):
    """w(z) = exp(-z^2) erfc(-i z) for Im z >= 0 by Weideman's rational
    expansion (J.A.C. Weideman, SIAM J. Numer. Anal. 31 (1994) 1497), N terms."""
    z = np.asarray(z, dtype=np.complex128)
    M = 2 * N
    M2 = 2 * M
    k = np.arange(-M + 1, M)
    L = math.sqrt(N / math.sqrt(2.0))
    theta = k * math.pi / M
    t = L * np.tan(theta / 2.0)
    f = np.exp(-t ** 2) * (L ** 2 + t ** 2)
    f = np.concatenate(([0.0], f))
    a = np.real(np.fft.fft(np.fft.fftshift(f))) / M2
    a = np.flipud(a[1:N + 1])
    Z = (L + 1j * z) / (L - 1j * z)
    p = np.polyval(a, Z)
    return 2.0 * p / (L - 1j * z) ** 2 + (1.0 / math.sqrt(math.pi)) / (L - 1j * z)


def brendel_bormann(freq_au, eps_inf, wp_ev, f0, g0_ev, osc):
    """Brendel-Bormann permittivity (A.D. Rakic et al., Appl. Opt. 37 (1998)
    5271): Drude term plus Gaussian-broadened Lorentz oscillators
    osc = [(f_j, w_j, g_j, s_j)] in eV, evaluated with the Faddeeva function."""
    w = np.asarray(freq_au, dtype=np.float64) * AU_TO_EV
    wp = wp_ev
    eps = eps_inf - f0 * wp ** 2 / (w * (w + 1j * g0_ev))
    for fj, wj, gj, sj in osc:
        a = np.sqrt(w * (w + 1j * gj))
        a = np.where(a.imag < 0, -a, a)
        zp = (a - wj) / (math.sqrt(2.0) * sj)
        zm = (a + wj) / (math.sqrt(2.0) * sj)
        chi = 1j * math.sqrt(math.pi) * fj * wp ** 2 / (2.0 * math.sqrt(2.0) * a * sj) * (faddeeva(zp) + faddeeva(zm))
        eps = eps + chi
    return eps


# Brendel-Bormann parameters of silver (Rakic et al. 1998, Table 3).
BB_SILVER = dict(eps_inf=1.0, wp_ev=9.01, f0=0.821, g0_ev=0.049,
                 osc=[(0.050, 2.025, 0.189, 1.894), (0.133, 5.185, 0.067, 0.665), (0.051, 4.343, 0.019, 0.189),
                      (0.467, 9.809, 0.117, 0.499), (4.000, 18.56, 0.052, 0.001)])
