#!/usr/bin/env python3
"""Turn the reference's regression inputs and logs into small JSON fixtures.

Run in a checkout that has the reference tree (default /root/reference):

    python tests/golden/make_fixtures.py [--ref /root/reference]

Only data is read: the YAML inputs, xyz / msh / csv data files and the
printed ``reference/*.log`` outputs of each test case.  Nothing from the
reference is imported or executed.  The YAML -> atomic-unit conversion rules
restated here are those of python/input_/validation.py:8-140 (units),
:1065-1090 (R_q, R_mu from eta / alpha), :1113-1320 (parameters) and
src/input_output/input_module.f90:238,251 (Angstrom -> bohr, threshold).

Each fixture holds the inputs (geometry, per-type parameters, frequencies,
field, method) and what the log prints about the solution: the 20 result
columns per frequency (E14.6), and at verbose 3 the RHS, charges, dipoles
and polarizability tensors, a deterministic sample of the printed matrices
(E11.4), and the GMRES residual histories (E13.5).
"""
import argparse
import csv
import json
import math
import os
import re

import yaml

TOBOHR = 1.8897261254578281
AU_TO_EV = 27.211386245981
AU_TO_CM = 219474.63136320
R0_BOHR = {  # src/parameters/parameters_module.f90:50-57
    "Ag": 2.885 * TOBOHR, "Au": 2.885 * TOBOHR, "Na": 6.9226100144600462,
    "C": 1.418 * TOBOHR, "Al": 2.861 * TOBOHR, "Cu": 2.543463 * TOBOHR,
}

ATOM_CASES = [
    "wfq_sodium_inversion", "wfq_sodium_iterative_on_memory", "wfq_sodium_iterative_on_the_fly",
    "wfq_graphene_disk_inversion", "wfq_graphene_ribbon_inversion", "wfq_graphene_triangle_inversion",
    "wfqfmu_ag_inversion", "wfqfmu_ag_iterative_on_memory", "wfqfmu_ag_iterative_on_the_fly",
    "wfqfmu_ag_internal_permittivity", "wfqfmu_ag_diff_units",
    "wfqfmu_hetero_agau_inversion", "wfqfmu_hetero_agau_iterative_on_memory",
    "wfqfmu_hetero_agau_iterative_on_the_fly", "wfqfmu_hetero_agau_read_1param_inversion",
]
BEM_CASES = [
    "bem_dpcm_accurate_inversion", "bem_ief_accurate_inversion", "bem_ief_approximate_inversion",
    "bem_dpcm_accurate_iterative_on_memory", "bem_dpcm_external_brendel_bormann_ag",
]


def value_unit(v):
    if isinstance(v, (int, float)):
        return float(v), None
    m = re.match(r"\s*([-+0-9.eEdD]+)\s*(.*)$", str(v))
    return float(m.group(1).replace("d", "e").replace("D", "e")), (m.group(2).strip() or None)


def freq_to_au(v):
    x, u = value_unit(v)
    u = (u or "ev").lower()
    if u == "ev":
        return x / AU_TO_EV
    if u == "nm":
        return (10000000.0 / AU_TO_CM) / x
    if u in ("au", "a.u."):
        return x
    raise ValueError(u)


def param_to_au(v, kind):
    x, u = value_unit(v)
    u = None if u is None else u.lower()
    if kind == "tau":
        return x if u in (None, "au", "a.u.") else x * 41.3413733351702 * (1e15 if u == "s" else 1.0)
    if kind == "sigma0":
        return x * 0.00000021739 if u in (None, "s/m") else x
    if kind == "a_ij":
        if u in (None, "au", "a.u."):
            return x
        if u in ("ang2", "ang^2", "angstrom2", "angstrom^2"):
            return x / 0.529177210544 ** 2
        return x / 0.0529177210544 ** 2
    if kind == "w_p":
        return freq_to_au(v)
    if kind in ("rq", "rmu"):
        if u in (None, "au", "a.u."):
            return x
        return x / 0.529177210544 if u in ("ang", "angstrom") else x / 0.0529177210544
    if kind in ("eta", "alpha", "chi"):
        return x
    raise ValueError(kind)


def frequencies(field):
    if "external freq" in field:
        ef = field["external freq"]
        ef = ef if isinstance(ef, list) else [ef]
        return [freq_to_au(f) for f in ef]
    n = int(field.get("nfreq", 0))
    lo, st = field["min freq"], field["step freq"]
    lo_au, st_au = freq_to_au(lo), freq_to_au(st)
    if n == 0:
        n = int(round((freq_to_au(field["max freq"]) - lo_au) / st_au)) + 1
    # src/field/field_module.f90:157-176: freq(i) = min + step (i-1)
    return [lo_au + st_au * i for i in range(n)]


def read_alpha_csv(path, freqs_au):
    rows = []
    for line in open(path):
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        p = s.replace(",", " ").split()
        rows.append((float(p[0]), float(p[1]), float(p[2])))
    out = []
    for f in freqs_au:
        ev = f * AU_TO_EV
        hit = [r for r in rows if abs(ev - r[0]) <= 1.0e-8]
        out.append([hit[0][1], hit[0][2]])
    return out


def parse_geometry(text):
    sym, xyz = [], []
    for line in text.strip().splitlines():
        p = line.split()
        if len(p) < 4:
            continue
        sym.append(p[0])
        xyz.append([float(p[1]), float(p[2]), float(p[3])])
    return sym, xyz


def num(s):
    return float(s.replace("D", "E"))


def parse_results(log):
    """Final 'Results for w =' blocks: 20 numbers each, keyed by eV."""
    out = {}
    lines = log.splitlines()
    for k, line in enumerate(lines):
        if line.strip().startswith("Results for w =") and k + 2 < len(lines) and "Isotr. Real Polar." in lines[k + 2]:
            ev = num(line.split()[-2])
            vals = []
            for j in range(k + 2, k + 22):
                vals.append(num(lines[j].split("=")[1].split()[0]))
            out[ev] = vals
    return [out[k] for k in sorted(out)]


def parse_block(lines, start, ncols):
    """Rows '  i  a b | c d | e f' after a header; returns list of rows."""
    rows = []
    k = start
    while k < len(lines):
        s = lines[k].replace("|", " ").split()
        if not s:
            if rows:
                break
            k += 1
            continue
        if s[0].startswith("---") or s[0] in ("ChErr",):
            break
        try:
            int(s[0])
            vals = [num(x) for x in s[1:1 + ncols]]
            if len(vals) == ncols:
                rows.append(vals)
            elif rows:
                break
        except ValueError:
            if rows:
                break
        k += 1
    return rows


def find(lines, text, after=0):
    for k in range(after, len(lines)):
        if text in lines[k]:
            return k
    return -1


def parse_matrix(lines, title, n):
    """Matrix printed in column blocks of 5 (E11.4): returns n x n list."""
    k = find(lines, title)
    if k < 0:
        return None
    M = [[None] * n for _ in range(n)]
    k += 1
    filled = 0
    while k < len(lines) and filled < n * n:
        s = lines[k].split()
        if s and all(re.fullmatch(r"\d+", t) for t in s):
            cols = [int(t) - 1 for t in s]
            for r in range(n):
                k += 1
                while not lines[k].strip():
                    k += 1
                p = lines[k].split()
                row = int(p[0]) - 1
                for c, v in zip(cols, p[1:]):
                    M[row][c] = num(v)
                    filled += 1
        k += 1
    return M


def sample_matrix(M, n, count=600):
    idx = []
    rng = 12345
    for _ in range(count):
        rng = (1103515245 * rng + 12345) % (2 ** 31)
        i = rng % n
        rng = (1103515245 * rng + 12345) % (2 ** 31)
        j = rng % n
        idx.append([i, j, M[i][j]])
    for i in range(min(n, 40)):
        idx.append([i, i, M[i][i]])
    return idx


def parse_gmres(lines):
    out = []
    k = 0
    while True:
        k = find(lines, "GMRES Iterative Algorithm. Freq:", k)
        if k < 0:
            break
        ev = float(re.search(r"Freq:\s*([0-9.]+)", lines[k]).group(1))
        thr = num(lines[k].split("=")[-1])
        j = k + 1
        while j < len(lines) and j < k + 6:
            m = re.search(r"- Polarization: ([XYZ])", lines[j])
            if m:
                break
            j += 1
        k = j
        while True:
            m = re.search(r"- Polarization: ([XYZ])", lines[k]) if k < len(lines) else None
            if not m:
                break
            pol = m.group(1)
            hist = []
            k += 2
            while k < len(lines):
                p = lines[k].split()
                if len(p) == 2 and re.fullmatch(r"\d+", p[0]):
                    hist.append(num(p[1]))
                    k += 1
                elif len(p) == 2 and p[0] == "Iteration":
                    k += 1
                else:
                    break
            out.append({"ev": ev, "threshold": thr, "pol": pol, "hist": hist})
            while k < len(lines) and not lines[k].strip():
                k += 1
    return out


def parse_verbose(lines, n_atoms, model):
    ex = {}
    # alloys print the static (rhs_0, 0) first and the per-frequency RHS last
    hits = [i for i, l in enumerate(lines) if "RHS Complex Response" in l]
    k = hits[-1] if hits else -1
    if k >= 0:
        ex["rhs"] = parse_block(lines, k + 1, 6)
    k = find(lines, "wFQ Charges")
    if k >= 0:
        ex["charges"] = parse_block(lines, k + 1, 6)
        c = find(lines, "ChErr", k)
        ex["cherr"] = [num(x) for x in lines[c].replace("|", " ").split()[1:7]]
    k = find(lines, "wFMu Dipoles")
    if k >= 0:
        ex["dipoles"] = parse_block(lines, k + 1, 6)
    k = find(lines, "BEM Charges")
    if k < 0:
        k = find(lines, "BEM charges")
    if k >= 0:
        ex["charges"] = parse_block(lines, k + 1, 6)
    for tag, title in (("polar_re", "Real Polarizability Tensor"), ("polar_im", "Imaginary Polarizability Tensor")):
        k = find(lines, title)
        if k >= 0:
            rows = []
            j = k + 1
            while len(rows) < 3 and j < len(lines):
                p = lines[j].split()
                if len(p) == 4 and p[0] in ("1", "2", "3"):
                    rows.append([num(x) for x in p[1:]])
                j += 1
            ex[tag] = rows
    return ex


def load_case(ref, case):
    d = os.path.join(ref, "tests", case)
    y = yaml.safe_load(open(os.path.join(d, case + ".yaml")))
    log = open(os.path.join(d, "reference", case + ".log")).read()
    return d, y, log


def atom_fixture(ref, case):
    d, y, log = load_case(ref, case)
    lines = log.splitlines()
    ff = y.get("forcefield", {})
    model = ff["dynamic"].lower()
    field = y.get("field", {})
    alg = y.get("algorithm", {})
    freqs = frequencies(field)
    if "input_geometry" in y:
        sym, xyz = parse_geometry(y["input_geometry"])
    else:
        xyzfile = [f for f in os.listdir(d) if f.endswith(".xyz")][0]
        body = open(os.path.join(d, xyzfile)).read().splitlines()[2:]
        sym, xyz = parse_geometry("\n".join(body))
    names = []
    for s in sym:
        if s not in names:
            names.append(s)
    # validation.py:1190-1194: element keys are normalised to 'Ag', 'Na', ...
    params = {(k if k.lower().startswith("interaction") else (k.lower().capitalize() if len(k) <= 2 else k.upper())): v
              for k, v in y["parameters"].items()}
    atom_types = {k.lower().capitalize() if len(k) <= 2 else k.upper(): v for k, v in y["atom_types"].items()}
    types = []
    for nm in names:
        p = params[nm]
        at = atom_types[nm]
        t = {"name": nm}
        t["eta"] = param_to_au(at["eta"], "eta")
        t["alpha"] = param_to_au(at.get("alpha", 0.0), "alpha")
        rq = param_to_au(at.get("rq", 0.0), "rq")
        rmu = param_to_au(at.get("rmu", 0.0), "rmu")
        if rq == 0.0:
            rq = math.sqrt(2.0 / math.pi) / t["eta"]
        if model == "wfqfmu" and rmu == 0.0:
            rmu = (t["alpha"] / (3.0 * math.sqrt(math.pi / 2.0))) ** (1.0 / 3.0)
        t["rq"], t["rmu"] = rq, rmu
        t["a_ij"] = param_to_au(p["A_ij"], "a_ij")
        tau = param_to_au(p["tau"], "tau") if "tau" in p else None
        if "gamma" in p and tau is None:
            tau = 1.0 / param_to_au(p["gamma"], "tau")
        if "sigma0" in p:
            sigma0 = param_to_au(p["sigma0"], "sigma0")
        elif "w_p" in p:
            sigma0 = param_to_au(p["w_p"], "w_p") ** 2 * tau / (4.0 * math.pi)
        else:
            sigma0 = 0.0
        scaling = float(p.get("scaling sigma0-tau", 1.0)) or 1.0
        if nm == "C":
            density = float(p["fermi energy"]) / (math.pi * AU_TO_EV)
        else:
            density = sigma0 / tau
        # wfq_module.f90:96-127: the density uses unscaled sigma0/tau, then both are scaled
        t["density"] = density
        t["tau"] = tau / scaling
        t["sigma0"] = sigma0 / scaling
        t["fermi_d"] = float(p["fermi_function"]["d"])
        t["fermi_s"] = float(p["fermi_function"]["s"])
        t["r0"] = R0_BOHR[nm]
        if model == "wfqfmu":
            if "wfqfmu file" in p and p["wfqfmu file"] != "none":
                t["alpha_w"] = read_alpha_csv(os.path.join(d, p["wfqfmu file"]), freqs)
            else:
                t["permittivity"] = p["permittivity"]
        types.append(t)
    nt = len(names)
    fd = [[0.0] * nt for _ in range(nt)]
    fs = [[0.0] * nt for _ in range(nt)]
    r0 = [[0.0] * nt for _ in range(nt)]
    for a, ta in enumerate(types):
        for b, tb in enumerate(types):
            if a == b:
                fd[a][b], fs[a][b] = ta["fermi_d"], ta["fermi_s"]
            r0[a][b] = R0_BOHR[ta["name"]] if ta["name"] == tb["name"] else 0.5 * (R0_BOHR[ta["name"]] + R0_BOHR[tb["name"]])
    given = set()
    for key, v in params.items():
        if key.startswith("interaction"):
            a_nm, b_nm = [s.strip() for s in key.split(" ", 1)[1].split("->")]
            a, b = names.index(a_nm), names.index(b_nm)
            fd[a][b] = float(v["fermi_function"]["d"])
            fs[a][b] = float(v["fermi_function"]["s"])
            given.add((a, b))
    # validation.py:1408-1520: a missing direction copies the given one; both
    # missing -> Ag/Au defaults (d 12, s 0.95)
    for a in range(nt):
        for b in range(nt):
            if a == b or (a, b) in given:
                continue
            if (b, a) in given:
                fd[a][b], fs[a][b] = fd[b][a], fs[b][a]
            else:
                fd[a][b], fs[a][b] = 12.0, 0.95
    method = alg.get("method", "inversion")
    fx = {
        "case": case, "kind": "atoms", "model": model, "kernel": ff.get("kernel", "gaussian").lower(),
        "rhs_form": field.get("rhs type", "field"), "method": method,
        "tol": float(str(alg.get("tolerance", 1.0e-9)).replace("d", "e")),
        "gmres_dim": int(alg.get("gmres dimension", 1000)),
        "e0": float(field.get("field intensity", 1.0e-4)),
        "freq_au": freqs, "types": types, "fermi_d": fd, "fermi_s": fs, "r0": r0,
        "symbols": sym, "coord_ang": xyz,
    }
    ex = {"results": parse_results(log)}
    ex.update(parse_verbose(lines, len(sym), model))
    nv = len(sym) * (4 if model == "wfqfmu" else 1)
    for tag, title in (("constant", "wfqfmu Constant Matrix"), ("dyn_re", "Matrix: Real Part"),
                       ("dyn_im", "Matrix: IMag Part")):
        M = parse_matrix(lines, title, nv)
        if M is not None and all(v is not None for row in M for v in row):
            ex[tag] = sample_matrix(M, nv)
    g = parse_gmres(lines)
    if g:
        ex["gmres"] = g
    fx["expected"] = ex
    return fx


def read_msh(path):
    lines = open(path).read().splitlines()
    k = [i for i, l in enumerate(lines) if l.strip().lower() == "$nodes"][0]
    nn = int(lines[k + 1])
    nodes = {}
    for l in lines[k + 2:k + 2 + nn]:
        p = l.split()
        nodes[int(p[0])] = [float(p[1]), float(p[2]), float(p[3])]
    k = [i for i, l in enumerate(lines) if l.strip().lower() == "$elements"][0]
    ne = int(lines[k + 1])
    faces = []
    for l in lines[k + 2:k + 2 + ne]:
        p = l.split()
        if len(p) >= 8:  # bem_module.f90:384-410: the last three numbers
            faces.append([int(p[5]), int(p[6]), int(p[7])])
    return nodes, faces


def bem_fixture(ref, case):
    d, y, log = load_case(ref, case)
    lines = log.splitlines()
    b = y["bem"]
    field = y.get("field", {})
    alg = y.get("algorithm", {})
    freqs = frequencies(field)
    nodes, faces = read_msh(os.path.join(d, b["mesh file"]))
    verts = [[nodes[f[0]], nodes[f[1]], nodes[f[2]]] for f in faces]
    fx = {
        "case": case, "kind": "bem", "variant": b["variant"], "green": b["green function"],
        "normal_factor": float(b.get("normal scalar factor", 1.0)),
        "sphere_r_bohr": (value_unit(b["sphere radius"])[0] / 0.529177210544) if "sphere radius" in b else 0.0,
        "eps_solvent": 1.0, "method": alg.get("method", "inversion"),
        "tol": float(str(alg.get("tolerance", 1.0e-9)).replace("d", "e")),
        "gmres_dim": int(alg.get("gmres dimension", 1000)),
        "e0": float(field.get("field intensity", 1.0e-4)), "freq_au": freqs,
        "verts_ang": verts,
    }
    if "permittivity file" in b:
        fx["eps_w"] = read_alpha_csv(os.path.join(d, b["permittivity file"]), freqs)
    else:
        fx["permittivity"] = b["permittivity"]
    ex = {"results": parse_results(log)}
    ex.update(parse_verbose(lines, len(faces), "bem"))
    n = len(faces)
    for tag, title in (("dyn_re", "BEM Matrix: Real Part"), ("dyn_im", "BEM Matrix: IMag Part")):
        M = parse_matrix(lines, title, n)
        if M is not None and all(v is not None for row in M for v in row):
            ex[tag] = sample_matrix(M, n)
    g = parse_gmres(lines)
    if g:
        ex["gmres"] = g
    fx["expected"] = ex
    return fx


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--out", default=os.path.dirname(os.path.abspath(__file__)))
    a = ap.parse_args()
    for case in ATOM_CASES + BEM_CASES:
        fx = atom_fixture(a.ref, case) if case in ATOM_CASES else bem_fixture(a.ref, case)
        path = os.path.join(a.out, case + ".json")
        with open(path, "w") as f:
            json.dump(fx, f, separators=(",", ":"))
        ex = fx["expected"]
        print(f"{case}: n_freq={len(fx['freq_au'])} results={len(ex['results'])} "
              f"keys={sorted(k for k in ex if k != 'results')} {os.path.getsize(path) // 1024} KB")


if __name__ == "__main__":
    main()
