"""The reference's result files, written and read in its formats.

- ``results_block`` / ``results_text``: the "Results for w" blocks that
  print_dynamic_results writes into the log (target_module.f90:1661-1734) and
  that each per-frequency backup holds;
- ``write_bk`` / ``read_bk`` / ``bk_path``: the per-frequency backup files
  ``<name>-<eV, f7.5>.plasmonX.bk`` (save_intermediate_results,
  target_module.f90:1765-1848; recover_from_backup_file,
  algorithm_module.f90:744-805) and ``restart_split``, the restart rule of
  the inversion and iterative loops (check_if_restart,
  algorithm_module.f90:716-741: a frequency whose backup exists is read back,
  not solved);
- ``write_csv``: the 21-column spectrum (save_csv_file,
  target_module.f90:1853-1911);
- ``write_freq``: the full-precision variables of one frequency
  (save_dynamic_variables, algorithm_module.f90:811-904);
- ``maxima_text``: the maxima analysis block (print_maxima,
  algorithm_module.f90:908-973);
- ``log_tail``: what the log holds from the backup clean-up to the maxima
  analysis (output_module.f90:174-186, then the blocks above), the region
  the reference's regression tests compare.

Fortran edit descriptors are reproduced by ``fort_e`` (Ew.d, correctly
rounded: the mantissa 0.d1..dd and a two-digit exponent) and ``fort_f``.
"""
import math
import os

import numpy as np

AU_TO_EV = 27.211386245981
AU_TO_CM = 219474.63136320
STICKS = " " + "-" * 80   # out_%sticks
RESULT_LABELS = [
    "Isotr. Real Polar.       = ", "Isotr. Imag Polar.       = ",
    "Long. Real Polar. X      = ", "Long. Real Polar. Y      = ", "Long. Real Polar. Z      = ",
    "Long. Imag Polar. X      = ", "Long. Imag Polar. Y      = ", "Long. Imag Polar. Z      = ",
    "Iso. Abs. Cross. Sect.   = ", "Long Abs. Cross. Sect. X = ", "Long Abs. Cross. Sect. Y = ",
    "Long Abs. Cross. Sect. Z = ", "Iso. Sca. Cross. Sect.   = ", "Long Sca. Cross. Sect. X = ",
    "Long Sca. Cross. Sect. Y = ", "Long Sca. Cross. Sect. Z = ", "Iso. Ext. Cross. Sect.   = ",
    "Long Ext. Cross. Sect. X = ", "Long Ext. Cross. Sect. Y = ", "Long Ext. Cross. Sect. Z = ",
]
CSV_HEADER = ["Frequency (eV)"] + [l.rstrip(" =").strip() for l in RESULT_LABELS]


def fort_e(x, w, d):
    """Fortran Ew.d: [-]0.d1d2..ddE+ee right-justified in w characters."""
    x = float(x)
    neg = math.copysign(1.0, x) < 0
    if x == 0.0 or not math.isfinite(x):
        digits, e = "0" * d, 0
        if not math.isfinite(x):
            return ("*" * w)
    else:
        s = f"{abs(x):.{d - 1}e}"          # correctly rounded d.ddd e+XX
        mant, ex = s.split("e")
        digits = mant.replace(".", "")
        e = int(ex) + 1
    body = f"0.{digits}E{'+' if e >= 0 else '-'}{abs(e):02d}"
    if neg:
        body = "-" + body
    return body.rjust(w)


def fort_f(x, w, d):
    """Fortran Fw.d."""
    return f"{float(x):{w}.{d}f}"


def freq_nm(freq_au):
    """freqautonm (parameters_module.f90): wavelength in nm."""
    return 1.0e7 / (freq_au * AU_TO_CM)


def results_block(freq_au, row):
    """One "Results for w" block of 20 result columns (lines, no trailing
    newline): format (1x,'Results for w = ',E9.3,' a.u.',3x,E9.3,' nm',3x,
    E9.3,' eV',/) then (1x,a,6x,E14.6,' a.u.') per column, then sticks."""
    ev = freq_au * AU_TO_EV
    lines = [f" Results for w = {fort_e(freq_au, 9, 3)} a.u.   {fort_e(freq_nm(freq_au), 9, 3)} nm   "
             f"{fort_e(ev, 9, 3)} eV", ""]
    for label, v in zip(RESULT_LABELS, row):
        lines.append(f" {label}      {fort_e(v, 14, 6)} a.u.")
    lines.append(STICKS)
    return lines


def results_text(freq_au, results):
    """print_dynamic_results: the blocks of every frequency, in order."""
    out = []
    for f, row in zip(freq_au, results):
        out += results_block(f, row)
    return "\n".join(out) + "\n"


def bk_path(log_path, freq_au):
    """<log name without .log>-<eV as f7.5>.plasmonX.bk"""
    base = log_path[:-4] if log_path.endswith(".log") else log_path
    return f"{base}-{fort_f(freq_au * AU_TO_EV, 7, 5).strip()}.plasmonX.bk"


def write_bk(log_path, freq_au, row):
    """save_intermediate_results: sticks, the block, sticks."""
    path = bk_path(log_path, freq_au)
    with open(path, "w") as f:
        f.write("\n".join([STICKS] + results_block(freq_au, row)) + "\n")
    return path


def read_bk(path):
    """recover_from_backup_file: (freq_au, freq_nm, freq_ev) from the header
    (17x,e9.3,8x,e9.3,6x,e9.3) and the 20 results (34x,e14.6)."""
    with open(path) as f:
        lines = f.read().splitlines()
    head = lines[1]
    freqs = (float(head[17:26]), float(head[34:43]), float(head[49:58]))
    row = [float(lines[3 + k][34:48]) for k in range(20)]
    return freqs, np.array(row)


def restart_split(log_path, freq_au):
    """The frequencies to solve and those recovered from existing backups
    (check_if_restart): (indices to solve, {index: results row})."""
    todo, recovered = [], {}
    for i, f in enumerate(freq_au):
        p = bk_path(log_path, f)
        if os.path.exists(p):
            recovered[i] = read_bk(p)[1]
        else:
            todo.append(i)
    return todo, recovered


def write_csv(log_path, freq_au, results):
    """save_csv_file: header comment lines, then (1x,E12.5,20(3X,E14.6))."""
    base = log_path[:-4] if log_path.endswith(".log") else log_path
    path = base + ".csv"
    head = ["#------------------------------------------", "#    Column          Quantity",
            "#------------------------------------------"]
    for k, name in enumerate(CSV_HEADER):
        head.append(f"#      {k + 1:<9d}{name:<27s}")
    head.append("#------------------------------------------")
    with open(path, "w") as fh:
        fh.write("\n".join(head) + "\n")
        for f, row in zip(freq_au, results):
            fh.write(" " + fort_e(f * AU_TO_EV, 12, 5) + "".join("   " + fort_e(v, 14, 6) for v in row) + "\n")
    return path


def write_freq(log_path, freq_au, variables, model, area=None):
    """save_dynamic_variables: variables (3, n_var) complex of one frequency,
    one block per field direction, (i8,5x,2(f25.16,3x)); BEM DPCM writes
    sigma * area (the tessera charge)."""
    base = log_path[:-4] if log_path.endswith(".log") else log_path
    ev = freq_au * AU_TO_EV
    path = f"{base}-{ev:12.10f}.freq".replace(" ", "")
    name = "Charges/Dipoles" if model == "wfqfmu" else "Charges"
    with open(path, "w") as fh:
        for k, comp in enumerate("xyz"):
            fh.write(f"complex {name}: {comp}. freq = {ev:25.16f}\n")
            v = np.asarray(variables[k])
            if area is not None:
                v = v * np.asarray(area)
            for i, z in enumerate(v):
                fh.write(f"{i + 1:8d}     {z.real:25.16f}   {z.imag:25.16f}   \n")
    return path


def maxima_text(freq_au, results, kind="absorption"):
    """print_maxima: interior local maxima of |isotropic cross section|
    (column 9 absorption, 13 scattering, 17 extinction)."""
    col = {"absorption": 8, "scattering": 12, "exctinction": 16, "extinction": 16}[kind]
    title = {8: "Isotr. Abs. Cross. Sec. ", 12: "Isotr. Sca. Cross. Sec. ", 16: "Isotr. Ext. Cross. Sec. "}[col]
    r = np.asarray(results)
    peaks = [i for i in range(1, len(freq_au) - 1)
             if abs(r[i - 1, col]) < abs(r[i, col]) and abs(r[i + 1, col]) < abs(r[i, col])]
    if not peaks:
        return ""
    lines = [f" Maxima Analysis: NumExFreq = {len(freq_au):6d}", "",
             f"          NState      Freq(eV)     {title} (a.u.)"]
    for k, i in enumerate(peaks):
        lines.append(f"          {k + 1:3d}     {fort_f(freq_au[i] * AU_TO_EV, 10, 3)}             "
                     f"{fort_e(r[i, col], 12, 5)}")
    lines.append(STICKS)
    return "\n".join(lines) + "\n"


def log_tail(freq_au, results, kind="absorption", recovered=()):
    """The log from the backup clean-up (output_module.f90:174-186) through
    print_dynamic_results and print_maxima (algorithm_module.f90:173-175);
    a "Recovering from backup frequency" line (algorithm_module.f90:770-775)
    precedes it for every frequency read from a backup."""
    head = [f" Recovering from backup frequency: {fort_f(freq_au[i] * AU_TO_EV, 7, 5)} eV." for i in recovered]
    text = "\n".join(head + [STICKS, "", " I am cleaning up the backup (*.plasmonX.bk)", "", STICKS]) + "\n"
    text += results_text(freq_au, results)
    if len(freq_au) > 1:
        text += maxima_text(freq_au, results, kind)
    return text
