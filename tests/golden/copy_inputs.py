#!/usr/bin/env python3
"""Copy some of the reference's test inputs into tests/golden/inputs.

Run in a checkout that has the reference tree (default /root/reference):

    python tests/golden/copy_inputs.py [--ref /root/reference]

Each case keeps its YAML input as it is, the xyz file it names and, of each
alpha(omega) table it names, the rows at the case's frequencies (the reader
matches rows by frequency, so the rest is not read).  For the static decks
(fq_*, fqfmu_*) it writes log.json with what the log prints: the model
matrix (E11.4), right-hand side, charges, dipoles, energy, J^-1 and the
polarizability tensor.  For the minimal decks
(the reference's defaults fill in the atom types and parameters) it also
writes log.json: what the case's reference/*.log prints about the input
(spectrum and per-type parameters, as printed) and the 20 result columns at
up to 12 frequencies spread over the spectrum.  Data only; nothing is run.
"""
import argparse
import json
import os
import shutil
import sys

import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from plasmonx_amd.inputs import AU_TO_EV, frequencies  # noqa: E402

CASES = ["wfq_sodium_inversion", "wfqfmu_ag_inversion", "wfqfmu_hetero_agau_inversion", "wfqfmu_ag_diff_units"]
MINIMAL = ["wfq_sodium_minimal_input", "wfq_graphene_minimal_input", "wfqfmu_ag_minimal_input",
           "wfqfmu_au_minimal_input", "wfqfmu_hetero_agau_minimal_input", "wfqfmu_au_read_xyz"]
# cases whose log tail (backup clean-up, result blocks, maxima analysis) is
# kept as text: the first three result blocks and the maxima block, verbatim
TAIL = ["wfq_sodium_inversion", "wfq_sodium_restart"]
STATIC = ["fq_energy", "fq_polarizability", "fq_principal_axis_energy", "fqfmu_energy", "fqfmu_rq_rmu_energy"]
HEADER_KEYS = ["Num. Freq.", "Min. Freq.", "Max. Freq.", "Step Freq.", "AtomType", "Eta", "Alpha", "R_q", "R_mu",
               "tau", "sigma_0", "scaling sigma0-tau", "A_ij", "fermi_d", "fermi_s", "density"]


def log_summary(path, keep=12):
    """Header lines 'key : value' of the input echo, and the result rows."""
    lines = open(path).read().splitlines()
    header = []
    for ln in lines:
        if ln.strip().startswith("Results for w ="):
            break
        if ":" in ln:
            k, v = ln.split(":", 1)
            if k.strip() in HEADER_KEYS:
                header.append([k.strip(), v.strip()])
    rows = {}
    for k, ln in enumerate(lines):
        if ln.strip().startswith("Results for w =") and k + 2 < len(lines) and "Isotr. Real Polar." in lines[k + 2]:
            ev = float(ln.split()[-2].replace("D", "E"))
            rows[ev] = [float(lines[j].split("=")[1].split()[0].replace("D", "E")) for j in range(k + 2, k + 22)]
    evs = sorted(rows)
    if len(evs) <= keep:
        pick = evs
    else:
        pick = sorted(set(evs[int(round(i * (len(evs) - 1) / max(1, keep - 1)))] for i in range(keep)))
    return {"header": header, "results_ev": pick, "results": [rows[e] for e in pick]}


def printed_matrix(lines, title):
    """A matrix printed by out_%print_matrix (blocks of 5 columns, E11.4):
    {(row, col): value} with 1-based indices, or {} when absent."""
    start = next((k for k, ln in enumerate(lines) if ln.strip() == title), None)
    out = {}
    if start is None:
        return out
    k = start + 2
    cols = []
    while k < len(lines):
        ln = lines[k].strip()
        if ln.startswith("---"):
            break
        parts = ln.split()
        if parts and all(p.isdigit() for p in parts):
            cols = [int(p) for p in parts]
        elif parts and parts[0].isdigit() and cols:
            row = int(parts[0])
            for c, v in zip(cols, parts[1:]):
                out[(row, c)] = float(v.replace("D", "E"))
        k += 1
    return out


def printed_column_block(lines, title, after=0):
    """Rows 'i v [v v]' following a title line (charges, dipoles, RHS)."""
    start = next((k for k in range(after, len(lines)) if lines[k].strip() == title), None)
    rows = []
    if start is None:
        return rows
    for ln in lines[start + 1:]:
        parts = ln.split()
        if not parts:
            if rows:
                break
            continue
        if parts[0].isdigit():
            rows.append([float(v.replace("D", "E")) for v in parts[1:]])
        elif rows or parts[0].startswith("---") or parts[0] == "ChErr":
            if rows:
                break
    return rows


def static_log_summary(path):
    """What a static (energy / static response) log prints: the model matrix,
    the right-hand side, charges, dipoles, energy, J^-1 and the polarizability."""
    lines = open(path).read().splitlines()
    out = {}
    for title in ("FQ Matrix", "FQFMu Matrix", "JMatrix-1", "polarizability tensor (a.u.)"):
        m = printed_matrix(lines, title)
        if m:
            out[title] = [[r, c, v] for (r, c), v in sorted(m.items())]
    for title in ("RHS", "Charges", "Dipoles"):
        rows = printed_column_block(lines, title)
        if rows:
            out[title] = rows
    for ln in lines:
        t = ln.strip()
        if t.startswith("Energy ="):
            out["energy"] = float(t.split("=")[1].split()[0])
        if t.startswith("Polar Iso  ="):
            out["polar_iso"] = float(t.split("=")[1].split()[0])
    return out


def log_tail_excerpt(path, blocks=3):
    """Verbatim lines of a dynamic log: any 'Recovering from backup' lines,
    the clean-up message with its sticks, the first result blocks and the
    maxima analysis block."""
    lines = open(path).read().splitlines()
    rec = [ln for ln in lines if ln.startswith(" Recovering from backup frequency")]
    k = next(i for i, ln in enumerate(lines) if "I am cleaning up the backup" in ln)
    out = rec + lines[k - 2:k + 3] + lines[k + 3:k + 3 + 23 * blocks]
    m = next(i for i, ln in enumerate(lines) if ln.startswith(" Maxima Analysis"))
    e = next(i for i in range(m + 1, len(lines)) if lines[i].startswith(" ----"))
    return "\n".join(out + lines[m:e + 1]) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    for case in STATIC:
        dst = os.path.join(HERE, "inputs", case)
        os.makedirs(dst, exist_ok=True)
        shutil.copy(os.path.join(a.ref, "tests", case, case + ".yaml"), os.path.join(dst, case + ".yaml"))
        with open(os.path.join(dst, "log.json"), "w") as f:
            json.dump(static_log_summary(os.path.join(a.ref, "tests", case, "reference", case + ".log")), f)
    for case in TAIL:
        dst = os.path.join(HERE, "inputs", case)
        os.makedirs(dst, exist_ok=True)
        shutil.copy(os.path.join(a.ref, "tests", case, case + ".yaml"), os.path.join(dst, case + ".yaml"))
        with open(os.path.join(dst, "log_tail.txt"), "w") as f:
            f.write(log_tail_excerpt(os.path.join(a.ref, "tests", case, "reference", case + ".log")))
        bk = os.path.join(a.ref, "tests", case, "bk")
        if os.path.isdir(bk):
            for name in sorted(os.listdir(bk)):
                shutil.copy(os.path.join(bk, name), os.path.join(dst, name))
        if case not in CASES:
            with open(os.path.join(dst, "log.json"), "w") as f:
                json.dump(log_summary(os.path.join(a.ref, "tests", case, "reference", case + ".log"), keep=100), f,
                          indent=0)
    for case in CASES + MINIMAL:
        src = os.path.join(a.ref, "tests", case, case + ".yaml")
        dst = os.path.join(HERE, "inputs", case)
        os.makedirs(dst, exist_ok=True)
        shutil.copy(src, os.path.join(dst, case + ".yaml"))
        with open(src) as f:
            y = yaml.safe_load(f)
        ev = [w * AU_TO_EV for w in frequencies(y.get("field", {}))]
        geo = y.get("input_geometry")
        if isinstance(geo, dict) and "external xyz file" in geo:
            shutil.copy(os.path.join(a.ref, "tests", case, geo["external xyz file"]),
                        os.path.join(dst, geo["external xyz file"]))
        if case in MINIMAL:
            log = os.path.join(a.ref, "tests", case, "reference", case + ".log")
            with open(os.path.join(dst, "log.json"), "w") as f:
                json.dump(log_summary(log), f, indent=0)
        for p in (y.get("parameters") or {}).values():
            if isinstance(p, dict) and p.get("wfqfmu file", "none") != "none":
                with open(os.path.join(a.ref, "tests", case, p["wfqfmu file"])) as f:
                    rows = [ln for ln in f if ln.strip() and not ln.startswith("#") and
                            any(abs(float(ln.replace(",", " ").split()[0]) - e) <= 1e-8 for e in ev)]
                with open(os.path.join(dst, p["wfqfmu file"]), "w") as f:
                    f.writelines(rows)


if __name__ == "__main__":
    main()
