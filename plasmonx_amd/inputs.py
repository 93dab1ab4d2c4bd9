"""Reader of the reference's YAML inputs for the frequency-response path.

Turns a plasmonX input file (the ``what: dynamic response`` cases with an
explicit geometry or an external xyz file) into the target, spectrum, field
and method that libpxh takes, with the reference's unit rules:
python/input_/validation.py:8-140 (units), :1065-1090 (R_q, R_mu from eta
and alpha), :1113-1320 (parameters), :1408-1520 (interaction defaults) and
src/input_output/input_module.f90:238,251 (Angstrom -> bohr, threshold);
frequencies as src/field/field_module.f90:157-176; the electronic density
from unscaled sigma0 / tau (src/wfq/wfq_module.f90:96-127).

``read_input(path)`` returns a dict whose keys are those ``system_from_case``
and the solver take: model, kernel, rhs_form, method, tol, gmres_dim, e0,
freq_au, types (per-type parameters in atomic units), fermi_d, fermi_s, r0
(per type pair), symbols, coord_ang.
"""
import math
import os
import re

import numpy as np
import yaml

from .model import TOBOHR, AtomSystem, AtomType

AU_TO_EV = 27.211386245981
AU_TO_CM = 219474.63136320
R0_BOHR = {  # src/parameters/parameters_module.f90:50-57
    "Ag": 2.885 * TOBOHR, "Au": 2.885 * TOBOHR, "Na": 6.9226100144600462,
    "C": 1.418 * TOBOHR, "Al": 2.861 * TOBOHR, "Cu": 2.543463 * TOBOHR,
}


def value_unit(v):
    """(number, unit or None) of a YAML value such as ``0.142E+02 a.u.``."""
    if isinstance(v, (int, float)):
        return float(v), None
    m = re.match(r"\s*([-+0-9.eEdD]+)\s*(.*)$", str(v))
    return float(m.group(1).replace("d", "e").replace("D", "e")), (m.group(2).strip() or None)


FS_TO_AU = 41.3413733351702      # atomic units of time per femtosecond
BOHR_ANG = 0.529177210544        # Angstrom per bohr
BOHR_NM = 0.0529177210544        # nm per bohr
SIGMA0_SI_TO_AU = 0.00000021739  # S/m -> a.u.


def freq_to_au(v):
    """Frequency in a.u. from eV (the default), nm, micron, cm-1, THz or a.u.
    (validation.py:8-33)."""
    x, u = value_unit(v)
    u = (u or "ev").lower()
    if u == "ev":
        return x / AU_TO_EV
    if u == "nm":
        return (10000000.0 / AU_TO_CM) / x
    if u in ("micron", "um"):
        return (10000.0 / AU_TO_CM) / x
    if u in ("cm-1", "cm^-1"):
        return x / AU_TO_CM
    if u == "thz":
        return x / 6579.683920502
    if u in ("au", "a.u."):
        return x
    raise ValueError(f"unknown frequency unit {u!r}")


# accepted units per parameter: unit -> conversion to a.u. (None stands for
# no unit written); the reference divides by the size of the atomic unit
# (eV, Angstrom, nm, s^-1) and multiplies by the atomic units per fs and per S/m
def _div(s):
    return lambda x: x / s


def _mul(s):
    return lambda x: x * s


_ONE = lambda x: x  # noqa: E731
_AU = {None: _ONE, "au": _ONE, "a.u.": _ONE}
_UNITS = {
    # atom_types (validation.py:35-77)
    "eta": {**_AU, "ev": _div(AU_TO_EV)},
    "chi": {**_AU, "ev": _div(AU_TO_EV)},
    "alpha": {**_AU, **{u: _div(BOHR_NM ** 3) for u in ("nm^3", "nm3")},
              **{u: _div(BOHR_ANG ** 3) for u in ("angstrom^3", "ang^3", "angstrom3", "ang3")}},
    "rq": {**_AU, "angstrom": _div(BOHR_ANG), "ang": _div(BOHR_ANG), "nm": _div(BOHR_NM)},
    # parameters (validation.py:79-140)
    "tau": {**_AU, "fs": _mul(FS_TO_AU), "s": _mul(FS_TO_AU * 1e15)},
    "gamma": {**_AU, "fs^-1": _div(FS_TO_AU), "fs-1": _div(FS_TO_AU), "s^-1": _div(FS_TO_AU * 1e15),
              "s-1": _div(FS_TO_AU * 1e15)},
    "sigma0": {None: _mul(SIGMA0_SI_TO_AU), "s/m": _mul(SIGMA0_SI_TO_AU), "au": _ONE, "a.u.": _ONE},
    "a_ij": {**_AU, **{u: _div(BOHR_NM ** 2) for u in ("nm^2", "nm2")},
             **{u: _div(BOHR_ANG ** 2) for u in ("angstrom^2", "ang2", "ang^2", "angstrom2")}},
}
_UNITS["rmu"] = _UNITS["rq"]


def param_to_au(v, kind):
    """A per-type parameter in atomic units, with the units the reference
    accepts for it (validation.py:35-140): eta and chi in eV; alpha in nm3 or
    Angstrom3; R_q, R_mu in Angstrom or nm; tau in fs or s; gamma in fs-1 or
    s-1; sigma0 in S/m (the default); A_ij in nm2 or Angstrom2; w_p in any
    frequency unit.  Other units raise ValueError."""
    if kind == "w_p":
        return freq_to_au(v)
    if kind not in _UNITS:
        raise ValueError(f"unknown parameter {kind!r}")
    x, u = value_unit(v)
    key = None if u is None else u.lower()
    if key not in _UNITS[kind]:
        raise ValueError(f"unsupported unit {u!r} for {kind!r}")
    return _UNITS[kind][key](x)


def frequencies(field):
    """The spectrum in a.u. (field_module.f90:157-176): the external list, or
    n = ceil((max - min) / step) + 1 points from min by step, with any three
    of min, max, step and nfreq given (validation.py:823-850)."""
    if "external freq" in field:
        ef = field["external freq"]
        if isinstance(ef, str):  # a comma-separated list, each value with its unit
            ef = [e for e in ef.split(",") if e.strip()]
        ef = ef if isinstance(ef, list) else [ef]
        return [freq_to_au(f) for f in ef]
    keys = ("min freq", "max freq", "step freq")
    have = [k for k in keys if k in field]
    if len(have) + ("nfreq" in field) < 3:
        raise ValueError("at least three of 'min freq', 'max freq', 'step freq', 'nfreq' are needed")
    # the missing one of the three follows from the raw numbers as written,
    # then min and step are converted (the count uses the written values, as
    # the reference's validation does before field_module converts them)
    raw = {k: value_unit(field[k]) for k in have}
    unit = next((u for _, u in raw.values() if u), None)
    x = {k: v[0] for k, v in raw.items()}
    n = int(field["nfreq"]) if "nfreq" in field else None
    if len(have) == 3:
        n_calc = int(math.ceil((x["max freq"] - x["min freq"]) / x["step freq"])) + 1
        if n is not None and n != n_calc:
            raise ValueError("'nfreq' is not coherent with min, max and step")
        n = n_calc
    elif "max freq" not in x:
        x["max freq"] = x["min freq"] + (n - 1) * x["step freq"]
    elif "min freq" not in x:
        x["min freq"] = x["max freq"] - (n - 1) * x["step freq"]
    else:
        x["step freq"] = (x["max freq"] - x["min freq"]) / (n - 1) if n > 1 else 0.0
    u = "" if unit is None else " " + unit
    lo, st = freq_to_au(f"{x['min freq']!r}{u}"), freq_to_au(f"{x['step freq']!r}{u}")
    return [lo + st * i for i in range(n)]


def read_alpha_csv(path, freqs_au):
    """[re, im] of alpha(omega) per frequency from a ``eV, re, im`` table
    (rows matched to 1e-8 eV)."""
    rows = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            p = s.replace(",", " ").split()
            rows.append((float(p[0]), float(p[1]), float(p[2])))
    out = []
    for fr in freqs_au:
        ev = fr * AU_TO_EV
        hit = [r for r in rows if abs(ev - r[0]) <= 1.0e-8]
        if not hit:
            raise ValueError(f"{path}: no row at {ev:.10f} eV")
        out.append([hit[0][1], hit[0][2]])
    return out


def parse_geometry(text):
    """Symbols and Angstrom coordinates of ``Sym x y z`` lines."""
    sym, xyz = [], []
    for line in text.strip().splitlines():
        p = line.split()
        if len(p) < 4:
            continue
        sym.append(p[0])
        xyz.append([float(p[1]), float(p[2]), float(p[3])])
    return sym, xyz


# atom types the reference fills in when the input's atom_types section does
# not name them (python/input_/validation.py:952-979)
PREDEFINED_ATOM_TYPES = {
    "Au": {"chi": 0.0, "eta": 0.52685, "alpha": 39.5297, "rq": 3.4996572071713126, "rmu": 3.0710227258450864},
    "Ag": {"chi": 0.0, "eta": 0.4829838, "alpha": 49.9843, "rq": 3.3775793621738925, "rmu": 2.368922250085755},
    "C": {"chi": 0.0, "eta": 0.372124, "rq": 0.0},
    "Na": {"chi": 0.0, "eta": 0.292000, "rq": 0.0},
}
# Drude and connectivity parameters of a type the input's parameters section
# does not name (validation.py:1134-1163); tau in a.u., sigma0 in S/m
PREDEFINED_PARAMETERS = {
    "Au": {"tau": 318.018, "sigma0": 11834849.81, "a_ij": 9.61, "fermi_function": {"d": 12.0, "s": 1.1},
           "permittivity": "gold etchegoin"},
    "Ag": {"tau": 1633.608, "sigma0": 6.5313e7, "a_ij": 9.61, "fermi_function": {"d": 12.0, "s": 1.1},
           "permittivity": "silver etchegoin"},
    "C": {"tau": 170.0, "a_ij": 1.7424, "fermi_function": {"d": 100.0, "s": 1.2}, "fermi energy": 0.4},
    "Na": {"tau": 1323.0, "sigma0": 2.39494681e7, "scaling sigma0-tau": 10.0, "a_ij": 12.07910025,
           "fermi_function": {"d": 12.0, "s": 1.1}},
}


def drude_parameters(p, name):
    """(tau, sigma0) in a.u. from the two of tau, sigma0, w_p, gamma a type
    gives (carbon: tau or gamma alone), as validation.py:1268-1292:
    tau = 1/gamma, sigma0 = w_p^2 tau / (4 pi)."""
    tau = param_to_au(p["tau"], "tau") if "tau" in p else None
    if tau is None and "gamma" in p:
        tau = 1.0 / param_to_au(p["gamma"], "gamma")
    if tau is None:
        raise ValueError(f"type {name!r}: tau or gamma is needed")
    if name == "C":
        return tau, 0.0
    if "sigma0" in p:
        return tau, param_to_au(p["sigma0"], "sigma0")
    if "w_p" in p:
        return tau, param_to_au(p["w_p"], "w_p") ** 2 * tau / (4.0 * math.pi)
    raise ValueError(f"type {name!r}: give (sigma0, tau), (w_p, tau), (w_p, gamma) or (sigma0, gamma)")


def _type_key(k):
    return k.lower().capitalize() if len(k) <= 2 else k.upper()


def read_input(path):
    """The case of a reference YAML input (atomistic dynamic response)."""
    d = os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        y = yaml.safe_load(f)
    ff = y.get("forcefield", {})
    if "dynamic" not in ff:
        raise ValueError(f"{path}: not a dynamic-response atomistic input")
    model = ff["dynamic"].lower()
    field = y.get("field", {})
    alg = y.get("algorithm", {})
    freqs = frequencies(field)
    geo = y.get("input_geometry")
    if isinstance(geo, str):
        sym, xyz = parse_geometry(geo)
    elif isinstance(geo, dict) and "external xyz file" in geo:
        with open(os.path.join(d, geo["external xyz file"])) as f:
            body = f.read().splitlines()[2:]
        sym, xyz = parse_geometry("\n".join(body))
    else:
        raise ValueError(f"{path}: generated geometries are built with plasmonx_amd.geometry")
    names = []
    for s in sym:
        if s not in names:
            names.append(s)
    given_params = y.get("parameters") or {}
    params = {(k if k.lower().startswith("interaction") else _type_key(k)): v for k, v in given_params.items()}
    atom_types = {_type_key(k): v for k, v in (y.get("atom_types") or {}).items()}
    types = []
    for nm in names:
        if nm not in atom_types and nm not in PREDEFINED_ATOM_TYPES:
            raise ValueError(f"{path}: no atom_types entry for {nm!r} and no predefined values")
        if nm not in params and nm not in PREDEFINED_PARAMETERS:
            raise ValueError(f"{path}: no parameters entry for {nm!r} and no predefined values")
        at = atom_types.get(nm, PREDEFINED_ATOM_TYPES.get(nm))
        p = params.get(nm, PREDEFINED_PARAMETERS.get(nm))
        t = {"name": nm, "eta": param_to_au(at["eta"], "eta"), "alpha": param_to_au(at.get("alpha", 0.0), "alpha")}
        rq = param_to_au(at.get("rq", 0.0), "rq")
        rmu = param_to_au(at.get("rmu", 0.0), "rmu")
        if rq == 0.0:
            rq = math.sqrt(2.0 / math.pi) / t["eta"]
        if model == "wfqfmu" and rmu == 0.0:
            rmu = (t["alpha"] / (3.0 * math.sqrt(math.pi / 2.0))) ** (1.0 / 3.0)
        t["rq"], t["rmu"] = rq, rmu
        t["a_ij"] = param_to_au(p["a_ij"] if "a_ij" in p else p["A_ij"], "a_ij")
        tau, sigma0 = drude_parameters(p, nm)
        scaling = float(p.get("scaling sigma0-tau", 1.0)) or 1.0
        t["density"] = float(p["fermi energy"]) / (math.pi * AU_TO_EV) if nm == "C" else sigma0 / tau
        t["tau"] = tau / scaling
        t["sigma0"] = sigma0 / scaling
        t["fermi_d"] = float(p["fermi_function"]["d"])
        t["fermi_s"] = float(p["fermi_function"]["s"])
        t["r0"] = R0_BOHR[nm]
        if model == "wfqfmu":
            if "wfqfmu file" in p and p["wfqfmu file"] != "none":
                t["alpha_w"] = read_alpha_csv(os.path.join(d, p["wfqfmu file"]), freqs)
            else:
                t["permittivity"] = p.get("permittivity", PREDEFINED_PARAMETERS.get(nm, {}).get("permittivity"))
        types.append(t)
    nt = len(names)
    fd = [[0.0] * nt for _ in range(nt)]
    fs = [[0.0] * nt for _ in range(nt)]
    r0 = [[0.0] * nt for _ in range(nt)]
    for a, ta in enumerate(types):
        for b, tb in enumerate(types):
            if a == b:
                fd[a][b], fs[a][b] = ta["fermi_d"], ta["fermi_s"]
            r0[a][b] = (R0_BOHR[ta["name"]] if ta["name"] == tb["name"]
                        else 0.5 * (R0_BOHR[ta["name"]] + R0_BOHR[tb["name"]]))
    given = set()
    for key, v in params.items():
        if key.startswith("interaction"):
            a_nm, b_nm = [s.strip() for s in key.split(" ", 1)[1].split("->")]
            a, b = names.index(a_nm), names.index(b_nm)
            fd[a][b] = float(v["fermi_function"]["d"])
            fs[a][b] = float(v["fermi_function"]["s"])
            given.add((a, b))
    # a missing direction copies the given one; both missing: d 12, s 0.95
    for a in range(nt):
        for b in range(nt):
            if a == b or (a, b) in given:
                continue
            if (b, a) in given:
                fd[a][b], fs[a][b] = fd[b][a], fs[b][a]
            else:
                fd[a][b], fs[a][b] = 12.0, 0.95
    return {
        "model": model, "kernel": ff.get("kernel", "gaussian").lower(),
        "rhs_form": field.get("rhs type", "field"), "method": alg.get("method", "inversion"),
        "tol": float(str(alg.get("tolerance", 1.0e-9)).replace("d", "e")),
        "gmres_dim": int(alg.get("gmres dimension", 1000)),
        "e0": float(field.get("field intensity", 1.0e-4)),
        "freq_au": freqs, "types": types, "fermi_d": fd, "fermi_s": fs, "r0": r0,
        "symbols": sym, "coord_ang": xyz,
    }


def system_from_case(case):
    """AtomSystem of a case from read_input."""
    types = []
    for t in case["types"]:
        aw = None if t.get("alpha_w") is None else np.array([complex(a, b) for a, b in t["alpha_w"]])
        types.append(AtomType(name=t["name"], eta=t["eta"], a_ij=t["a_ij"], tau=t["tau"], density=t["density"],
                              fermi_d=t["fermi_d"], fermi_s=t["fermi_s"], r_q=t["rq"], r_mu=t["rmu"],
                              alpha=t.get("alpha", 0.0), alpha_w=aw, permittivity=t.get("permittivity", "")))
    return AtomSystem(model=case["model"], kernel=case["kernel"], symbols=list(case["symbols"]),
                      coord_bohr=np.array(case["coord_ang"], dtype=np.float64) * TOBOHR, types=types,
                      fermi_d=np.array(case["fermi_d"]), fermi_s=np.array(case["fermi_s"]), r0=np.array(case["r0"]))


_IMOL = re.compile(r"^(.*?)\s*\[\s*IMol\s*=\s*(\d+)\s*\]\s*(.*)$", re.IGNORECASE)


def parse_geometry_molecules(text):
    """Symbols, Angstrom coordinates and molecule index of ``Sym[IMol= k] x y
    z`` lines (python/input_/io_operations.py:123-190; a line without the
    tag is molecule 1)."""
    sym, xyz, mol = [], [], []
    for line in text.strip().splitlines():
        m = _IMOL.match(line.strip())
        if m:
            name, k, rest = m.group(1).strip(), int(m.group(2)), m.group(3).split()
        else:
            p = line.split()
            if len(p) < 4:
                continue
            name, k, rest = p[0], 1, p[1:]
        sym.append(name)
        mol.append(k)
        xyz.append([float(v) for v in rest[:3]])
    return sym, xyz, mol


def molecule_runs(mol):
    """Atoms per molecule as runs of consecutive equal molecule indices
    (create_n_atoms_per_molecule, input_module.f90:1007-1031)."""
    counts = []
    for k, m in enumerate(mol):
        if k == 0 or m != mol[k - 1]:
            counts.append(0)
        counts[-1] += 1
    return counts


def read_static_input(path):
    """The case of a reference YAML input for the static models ('what:
    energy' or 'static response', 'forcefield static: fq | fqfmu'): kernel
    defaults (validation.py:881-903), per-type chi / eta / alpha / R_q / R_mu
    with the reference's units and the Gaussian widths from eta and alpha
    (:1065-1090), the static field (intensity 0, :752-776), molecules from the
    IMol tags."""
    d = os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        y = yaml.safe_load(f)
    what = y.get("what")
    what = (what[0] if isinstance(what, list) else what).lower()
    if what not in ("energy", "static response"):
        raise ValueError(f"{path}: not a static (energy / static response) input")
    ff = y.get("forcefield", {})
    model = ff["static"].lower()
    kernel = ff.get("kernel", {"fq": "ohno", "fqfmu": "gaussian"}.get(model, "ohno")).lower()
    geo = y.get("input_geometry")
    if isinstance(geo, dict) and "external xyz file" in geo:
        with open(os.path.join(d, geo["external xyz file"])) as f:
            geo = "\n".join(f.read().splitlines()[2:])
    sym, xyz, mol = parse_geometry_molecules(geo)
    names = []
    for s in sym:
        if s not in names:
            names.append(s)
    atom_types = {_type_key(k): v for k, v in (y.get("atom_types") or {}).items()}
    types = []
    for nm in names:
        at = atom_types[_type_key(nm)]
        t = {"name": nm, "chi": param_to_au(at.get("chi", 0.0), "chi"), "eta": param_to_au(at["eta"], "eta"),
             "alpha": param_to_au(at.get("alpha", 0.0), "alpha"), "rq": param_to_au(at.get("rq", 0.0), "rq"),
             "rmu": param_to_au(at.get("rmu", 0.0), "rmu")}
        if kernel == "gaussian":
            if t["rq"] == 0.0:
                t["rq"] = math.sqrt(2.0 / math.pi) / t["eta"]
            if model == "fqfmu" and t["rmu"] == 0.0:
                t["rmu"] = (t["alpha"] / (3.0 * math.sqrt(math.pi / 2.0))) ** (1.0 / 3.0)
        types.append(t)
    field = y.get("field") or {}
    control = [str(c).lower() for c in (y.get("control") or [])]
    return {"model": model, "kernel": kernel, "what": what, "rhs_form": field.get("rhs type", "field"),
            "e0": 0.0 if what == "static response" else float(field.get("field intensity", 0.0)),
            "types": types, "symbols": sym, "coord_ang": xyz, "mol_count": molecule_runs(mol),
            "principal_axes": "principal axes" in control}


def static_system_from_case(case):
    """AtomSystem of a static case from read_static_input (the Drude and
    connectivity fields are unused by the static models)."""
    types = [AtomType(name=t["name"], eta=t["eta"], a_ij=0.0, tau=0.0, density=0.0, fermi_d=0.0, fermi_s=0.0,
                      r_q=t["rq"], r_mu=t["rmu"], alpha=t["alpha"], chi=t["chi"]) for t in case["types"]]
    nt = len(types)
    z = np.zeros((nt, nt))
    return AtomSystem(model=case["model"], kernel=case["kernel"], symbols=list(case["symbols"]),
                      coord_bohr=np.array(case["coord_ang"], dtype=np.float64) * TOBOHR, types=types,
                      fermi_d=z, fermi_s=z, r0=z)
