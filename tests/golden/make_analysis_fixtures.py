#!/usr/bin/env python3
"""Fixtures for the induced-field grid from the reference's analysis tests.

Run in a checkout that has the reference tree (default /root/reference):

    python tests/golden/make_analysis_fixtures.py [--ref /root/reference]

Only data is read: each test's input archive (the solver's ``.info`` target
description and the ``.freq`` complex variables of one frequency) and the
``reference/*/*.cube`` grids that plasmonX_analysis printed for it.  Nothing
from the reference is imported or executed.

The cube holds |E(r)| / E0 on the cubic grid of
src/plasmonX_analysis/grid_module.f90:64-88 (atoms' extent +- the offset,
10 Angstrom here, n points per axis), x outermost and z innermost
(plot_module.f90:544-558, E14.6), with E the incident field plus the field
of the Gaussian charges (matrix_module.f90:1140-1172) and, for omegaFQFmu,
of the Gaussian dipoles (:1174-1246), summed over atoms
(wfq_module.f90:592-617, wfqfmu_module.f90:973-1003).  The density cubes
hold the real and imaginary parts of the plasmon density on the same grid:
the normalised Gaussian of each charge (matrix_module.f90:1258-1284) and,
for omegaFQFmu, minus the dipole's derivative of the Gaussian of R_mu
(:1295-1324), summed over atoms (wfq_module.f90:625-648,
wfqfmu_module.f90:1012-1040).

Each fixture keeps the target (coordinates, per-type R_q / R_mu and the
dynamic parameters of the .info), E0, the variables of each field direction,
the grid (origin, step, points per axis, from the cube header) and a
deterministic sample of the cube's values with their flat indices.  Plane
fixtures keep the CSV rows (two in-plane coordinates in Angstrom and the
value, E25.16) of calculate_field_2d / calculate_density_2d
(grid_module.f90:310-453, plot_module.f90:568-611) on the planes of
create_plane_grid (grid_module.f90:93-154); volume fixtures keep the
effective volume and area printed by print_effective_volume_area
(plot_module.f90:680-735) and the maximum of the field.
"""
import argparse
import json
import os
import re
import tarfile

TOBOHR = 1.8897261254578281
R0_BOHR = {"Ag": 2.885 * TOBOHR, "Au": 2.885 * TOBOHR, "Na": 6.9226100144600462}
SYMBOL = {11: "Na", 47: "Ag", 79: "Au"}

CASES = {
    # case: (frequency tag, field directions with a cube, quantity)
    "analysis_field_3d_x_wfq_sodium": ("1.5800000000", ["x"], "field"),
    "analysis_field_3d_x_y_z_wfqfmu_ag": ("3.0000000000", ["x", "y", "z"], "field"),
    "analysis_density_3d_x_wfq_sodium": ("1.5800000000", ["x"], "density"),
    "analysis_density_3d_y_wfqfmu_au": ("3.0000000000", ["y"], "density"),
    "analysis_field_3d_y_bem": ("1.9100000000", ["y"], "field"),
}
SAMPLE = 600  # cube values kept per direction


def read_info(text):
    lines = text.splitlines()
    head = {}
    for ln in lines[:12]:
        if ":" in ln:
            k, v = ln.split(":", 1)
            head[k.strip()] = v.strip()
    if head["Model"] == "bem":  # tessera centroids only
        n = int(head["NVar"])
        i0 = next(i for i, ln in enumerate(lines) if ln.startswith("Input Geometry"))
        coord = [[float(v) for v in ln.split()[1:4]] for ln in lines[i0 + 1:i0 + 1 + n]]
        return {"model": "bem", "kernel": head["Kernel"], "e0": float(head["Input |E|"]), "symbols": [],
                "coord_ang": coord, "types": [], "dynamic": [], "n_var": n}
    n = int(head["NAtoms"])
    i0 = next(i for i, ln in enumerate(lines) if ln.startswith("Input Geometry"))
    symbols, coord, tidx = [], [], []
    for ln in lines[i0 + 1:i0 + 1 + n]:
        z, t, x, y, w = ln.split()
        symbols.append(SYMBOL[int(float(z))])
        tidx.append(int(t) - 1)
        coord.append([float(x), float(y), float(w)])
    j = next(i for i, ln in enumerate(lines) if ln.startswith("Parameters:"))
    nt = int(lines[j + 1].split(":")[1])
    types = []
    for ln in lines[j + 2:j + 2 + nt]:
        f = ln.split()
        # fq: name chi eta r_q ; fqfmu: name chi eta alpha r_q r_mu
        if len(f) == 4:
            types.append({"name": f[0], "chi": float(f[1]), "eta": float(f[2]), "rq": float(f[3]), "rmu": 0.0})
        else:
            types.append({"name": f[0], "chi": float(f[1]), "eta": float(f[2]), "alpha": float(f[3]),
                          "rq": float(f[4]), "rmu": float(f[5])})
    k = next(i for i, ln in enumerate(lines) if ln.startswith("Dynamic parameters"))
    ndyn = int(lines[k].split(":")[1])
    dyn = [float(lines[k + 1 + q].split()[1]) for q in range(ndyn)]
    return {"model": head["Model"], "kernel": head["Kernel"], "e0": float(head["Input |E|"]), "symbols": symbols,
            "coord_ang": coord, "type_index": tidx, "types": types, "dynamic": dyn}


def read_freq(text, n_var):
    out, cur = {}, None
    for ln in text.splitlines():
        m = re.match(r"complex .*: ([xyz])\. freq", ln)
        if m:
            cur = m.group(1)
            out[cur] = []
            continue
        f = ln.split()
        if cur and len(f) == 3:
            out[cur].append([float(f[1]), float(f[2])])
    for d in out:
        assert len(out[d]) == n_var, (d, len(out[d]), n_var)
    return out


def read_cube(text):
    lines = text.splitlines()
    n_at, *org = lines[2].split()
    axes = [lines[3 + a].split() for a in range(3)]
    npts = [int(a[0]) for a in axes]
    step = [float(axes[a][1 + a]) for a in range(3)]
    vals = " ".join(lines[6 + int(n_at):]).split()
    vals = [float(v) for v in vals]
    assert len(vals) == npts[0] * npts[1] * npts[2]
    return {"origin_bohr": [float(o) for o in org], "step_bohr": step, "n": npts}, vals


PLANE_CASES = {
    # case: (frequency tag, field direction, quantity, plane, third coordinates, user min/max or None)
    "analysis_field_2d_x_yz_wfq_sodium": ("1.5800000000", "x", "field", "yz", ["6.35"], None),
    "analysis_field_2d_z_xz_wfqfmu_ag": ("3.0000000000", "z", "field", "xz", ["0.00"], None),
    "analysis_field_2d_z_xz_wfqfmu_ag_reduced": ("3.0000000000", "z", "field", "xz", ["0.00"],
                                                 ([-5.0, 0.0, -5.0], [5.0, 0.0, 5.0])),
    "analysis_density_2d_y_xy_wfqfmu_ag": ("3.0000000000", "y", "field", "xy", ["0.00"], None),
    "analysis_density_2d_x_wfq_sodium": ("1.5800000000", "x", "density", "xy", ["6.35"], None),
    "analysis_density_2d_x_wfq_sodium_n_planes": ("1.5800000000", "x", "density", "xy", ["6.35", "6.85", "7.35"],
                                                  None),
    "analysis_field_2d_y_xy_bem": ("1.9100000000", "y", "field", "xy", ["0.00"], None),
}
VOLUME_CASES = {
    # case: (frequency tag, field direction, points per axis, user min, user max)
    "analysis_field_volume_x_wfq_sodium": ("1.5800000000", "x", 30, [53.0, 0.0, 0.0], [55.0, 10.0, 10.0]),
    "analysis_field_volume_z_wfqfmu_au": ("3.0000000000", "z", 30, [10.0, -4.0, -4.0], [12.0, 4.0, 4.0]),
}
PLT_CASES = {
    # case: (frequency tag, field directions, quantity)
    "analysis_field_3d_x_wfq_sodium_plt": ("1.5800000000", ["x"], "field"),
    "analysis_field_3d_x_y_z_wfqfmu_ag_plt": ("3.0000000000", ["x", "y", "z"], "field"),
    "analysis_density_3d_x_wfq_sodium_plt": ("1.5800000000", ["x"], "density"),
    "analysis_density_3d_y_wfqfmu_au_plt": ("3.0000000000", ["y"], "density"),
}
PLANE_KEEP = 1000  # planes with more points keep an even sample of SAMPLE


def read_target(ref, case, tag):
    tdir = os.path.join(ref, "tests", case)
    with tarfile.open(os.path.join(tdir, case + ".tar.gz")) as tf:
        info = read_info(tf.extractfile(case + ".info").read().decode())
        n = len(info["symbols"])
        n_var = info.get("n_var") or (4 * n if info["model"] == "wfqfmu" else n)
        var = read_freq(tf.extractfile(f"{case}-{tag}.freq").read().decode(), n_var)
    refdir = os.path.join(tdir, "reference")
    sub = [d for d in os.listdir(refdir) if os.path.isdir(os.path.join(refdir, d))][0]
    return info, var, os.path.join(refdir, sub)


def target_fields(case, info):
    return {"case": case, "model": info["model"], "kernel": info["kernel"], "e0": info["e0"],
            "symbols": info["symbols"], "coord_ang": info["coord_ang"], "types": info["types"],
            "dynamic": info["dynamic"], "r0_bohr": {s: R0_BOHR[s] for s in set(info["symbols"])},
            # BEM: coord_ang are the tessera centroids and the variables their charges
            "grid_offset_ang": 10.0}


def read_plt(data):
    """write_plt_file (plot_module.f90:452-495): int32 3, 200, nz, ny, nx;
    float32 z min/max, y min/max, x min/max (Angstrom); float32 values with x
    fastest."""
    import struct
    ints = struct.unpack("<5i", data[:20])
    box = struct.unpack("<6f", data[20:44])
    n = ints[2] * ints[3] * ints[4]
    vals = struct.unpack("<%df" % n, data[44:44 + 4 * n])
    assert len(data) == 44 + 4 * n
    return list(ints), list(box), list(vals)


def read_csv(text):
    """Rows of (coordinate 1, coordinate 2, value), E25.16; a blank line ends
    each run of the second coordinate."""
    rows, n2, run = [], None, 0
    for ln in text.splitlines():
        f = ln.split()
        if not f:
            n2 = n2 or run
            run = 0
            continue
        rows.append([float(v) for v in f])
        run += 1
    return rows, n2


def sample(total, *values):
    if total <= PLANE_KEEP:
        return list(range(total))
    extra = [max(range(total), key=v.__getitem__) for v in values] + [min(range(total), key=v.__getitem__)
                                                                      for v in values]
    return sorted(set(list(range(0, total, max(1, total // SAMPLE))) + extra))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    here = os.path.dirname(os.path.abspath(__file__))
    for case, (tag, dirs, what) in CASES.items():
        info, var, rdir = read_target(a.ref, case, tag)
        cubes, grid = {}, None
        for d in dirs:
            parts = {"values": ""} if what == "field" else {"re": "-densityRe", "im": "-densityIm"}
            read = {}
            for key, suffix in parts.items():
                with open(os.path.join(rdir, f"{case}-E{d}-{tag}{suffix}.cube")) as f:
                    g, read[key] = read_cube(f.read())
                grid = grid or g
            # deterministic sample: evenly spaced flat indices plus each part's extremes
            total = len(next(iter(read.values())))
            extra = [max(range(total), key=v.__getitem__) for v in read.values()]
            extra += [min(range(total), key=v.__getitem__) for v in read.values()]
            idx = sorted(set(list(range(0, total, max(1, total // SAMPLE))) + extra))
            cubes[d] = {"idx": idx, **{k: [v[i] for i in idx] for k, v in read.items()}}
        fx = {**target_fields(case, info), "quantity": what, "grid": grid,
              "variables": {d: var[d] for d in dirs}, "cube": cubes}
        with open(os.path.join(here, case + ".json"), "w") as f:
            json.dump(fx, f)
        print(case, len(info["symbols"]), "atoms", {d: len(cubes[d]["idx"]) for d in dirs}, "values")

    planes = {}
    for case, (tag, d, what, plane, thirds, user) in PLANE_CASES.items():
        info, var, rdir = read_target(a.ref, case, tag)
        out = []
        for third in thirds:
            parts = {"values": ""} if what == "field" else {"re": "-densityRe", "im": "-densityIm"}
            read, coords, n2 = {}, None, None
            for key, suffix in parts.items():
                with open(os.path.join(rdir, "planes", plane, f"{case}-E{d}-{tag}-p-{third}{suffix}.csv")) as f:
                    rows, n2 = read_csv(f.read())
                read[key] = [r[2] for r in rows]
                coords = coords or [r[:2] for r in rows]
            idx = sample(len(coords), *read.values())
            out.append({"third": float(third), "n": [len(coords) // n2, n2], "idx": idx,
                        "coords": [coords[i] for i in idx], **{k: [v[i] for i in idx] for k, v in read.items()}})
        fx = {**target_fields(case, info), "quantity": what, "plane": plane, "user_grid": user,
              "field_dir": d, "variables": {d: var[d]}, "planes": out}
        planes[case] = fx
        with open(os.path.join(here, case + ".json"), "w") as f:
            json.dump(fx, f)
        print(case, len(info["symbols"]), "atoms", [len(p["idx"]) for p in out], "plane values")

    for case, (tag, dirs, what) in PLT_CASES.items():
        info, var, rdir = read_target(a.ref, case, tag)
        out, head = {}, None
        for d in dirs:
            parts = {"values": ""} if what == "field" else {"re": "-densityRe", "im": "-densityIm"}
            read = {}
            for key, suffix in parts.items():
                with open(os.path.join(rdir, f"{case}-E{d}-{tag}{suffix}.plt"), "rb") as f:
                    ints, box, read[key] = read_plt(f.read())
                head = head or {"ints": ints, "box_ang": box}
            total = len(next(iter(read.values())))
            extra = [max(range(total), key=v.__getitem__) for v in read.values()]
            idx = sorted(set(list(range(0, total, max(1, total // SAMPLE))) + extra))
            out[d] = {"idx": idx, **{k: [v[i] for i in idx] for k, v in read.items()}}
        fx = {**target_fields(case, info), "quantity": what, "plt": head, "variables": {d: var[d] for d in dirs},
              "values": out}
        with open(os.path.join(here, case + ".json"), "w") as f:
            json.dump(fx, f)
        print(case, len(info["symbols"]), "atoms", head, {d: len(out[d]["idx"]) for d in dirs}, "plt values")

    for case, (tag, d, n, lo, hi) in VOLUME_CASES.items():
        info, var, rdir = read_target(a.ref, case, tag)
        with open(os.path.join(rdir, case + "_analysis.log")) as f:
            log = f.read().splitlines()
        vol = next(float(ln.split(":")[1].split()[0]) for ln in log if ln.strip().startswith("Effective Volume"))
        area = next(float(ln.split(":")[1].split()[0]) for ln in log if ln.strip().startswith("Effective Area"))
        i = next(k for k, ln in enumerate(log) if ln.split()[:2] == ["Field", "X"])
        f = log[i + 1].split()
        fx = {**target_fields(case, info), "quantity": "volume", "field_dir": d, "n": n, "min_ang": lo, "max_ang": hi,
              "variables": {d: var[d]}, "effective_volume_nm3": vol, "effective_area_nm2": area,
              "max_point_ang": [float(v) for v in f[1:4]], "max_value": float(f[4])}
        with open(os.path.join(here, case + ".json"), "w") as fo:
            json.dump(fx, fo)
        print(case, len(info["symbols"]), "atoms", vol, area, f)


if __name__ == "__main__":
    main()
