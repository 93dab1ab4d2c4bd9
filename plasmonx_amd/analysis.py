"""Induced field and plasmon density on a grid: the cubes of plasmonX_analysis.

The reference's post-processing evaluates, at the points of a cubic grid
around the target, the incident field plus the field of the solved Gaussian
charges (and, for omegaFQFmu, dipoles) of one frequency and one field
direction, and writes |E| / E0 (or the real and imaginary parts of the
plasmon density, grid_module.f90:248-301) as a Gaussian cube
(src/plasmonX_analysis/grid_module.f90:64-88 grid, :200-242 field,
src/plasmonX_analysis/plot_module.f90:496-562 cube).  The point sums run on
the GPU through pxh_induced_field and pxh_plasmon_density; this module builds
the grid and writes the cube.
"""
import numpy as np

from .model import TOBOHR


def cubic_grid(coord_ang, n=(30, 30, 30), offset_ang=10.0):
    """Grid of create_cubic_grid (grid_module.f90:64-88): the atoms' extent
    +- offset on each axis, n points per axis.  Returns (min_ang, step_ang,
    points_ang) with the points in cube order (x outermost, z innermost,
    plot_module.f90:544-558), shape (nx*ny*nz, 3)."""
    c = np.asarray(coord_ang, dtype=np.float64)
    lo = c.min(axis=0) - offset_ang
    hi = c.max(axis=0) + offset_ang
    n = np.asarray(n, dtype=np.int64)
    step = (hi - lo) / (n - 1)
    i, j, k = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    pts = np.stack([lo[0] + i * step[0], lo[1] + j * step[1], lo[2] + k * step[2]], axis=-1).reshape(-1, 3)
    return lo, step, pts


def field_enhancement(ctx, points_ang, variables, i_pol, e0):
    """|E(p)| / E0 at the points (Angstrom) for the variables (n_var complex)
    of the field direction i_pol (0, 1, 2), as calculate_field_3d
    (grid_module.f90:200-242)."""
    e = ctx.induced_field(np.asarray(points_ang, dtype=np.float64) * TOBOHR, variables, i_pol, e0)
    return np.sqrt((np.abs(e) ** 2).sum(axis=1)) / e0


def plasmon_density(ctx, points_ang, variables, parts=False):
    """Plasmon density at the points (Angstrom) for the variables (n_var
    complex), as calculate_density_3d (grid_module.f90:248-301); its real and
    imaginary parts are the densityRe / densityIm cubes.  parts=True returns
    (rho, rho_q, rho_mu) as the separate_q_mu option does."""
    return ctx.plasmon_density(np.asarray(points_ang, dtype=np.float64) * TOBOHR, variables, parts)


def fortran_e(x, w=14, d=6):
    """Fortran Ew.d edit of a real: 0.ddd...d E+ee (d significant digits,
    correctly rounded), right-aligned in w."""
    if x == 0.0 or not np.isfinite(x):
        digits, e, neg = "0" * d, 0, False
    else:
        mant, ex = ("%.*e" % (d - 1, abs(x))).split("e")
        digits, e, neg = mant.replace(".", ""), int(ex) + 1, x < 0
    return (("-" if neg else "") + "0." + digits + "E" + "%+03d" % e).rjust(w)


# plane name -> (first in-plane axis, second in-plane axis, normal axis)
PLANE_AXES = {"xy": (0, 1, 2), "xz": (0, 2, 1), "yz": (1, 2, 0)}


def plane_grid(coord_ang, plane, n, offset_ang=10.0, lo_ang=None, hi_ang=None):
    """Grid of create_plane_grid (grid_module.f90:93-154): along the plane's
    two axes the atoms' extent +- offset (or the given min / max), n = (n1,
    n2) points.  Returns (lo, step), length 3, zero on the normal axis."""
    a1, a2, _ = PLANE_AXES[plane]
    c = np.asarray(coord_ang, dtype=np.float64)
    lo = c.min(axis=0) - offset_ang if lo_ang is None else np.asarray(lo_ang, dtype=np.float64)
    hi = c.max(axis=0) + offset_ang if hi_ang is None else np.asarray(hi_ang, dtype=np.float64)
    lo3, step = np.zeros(3), np.zeros(3)
    for a, m in ((a1, n[0]), (a2, n[1])):
        lo3[a] = lo[a]
        step[a] = (hi[a] - lo[a]) / (m - 1)
    return lo3, step


def plane_points(lo, step, plane, n, third_ang):
    """Points (n1*n2, 3) Angstrom of one plane at third_ang on the normal
    axis, in the row order of the reference's CSV (first in-plane axis
    outermost, plot_module.f90:597-608)."""
    a1, a2, an = PLANE_AXES[plane]
    i, j = np.meshgrid(np.arange(n[0]), np.arange(n[1]), indexing="ij")
    pts = np.zeros((n[0] * n[1], 3))
    pts[:, a1] = lo[a1] + i.reshape(-1) * step[a1]
    pts[:, a2] = lo[a2] + j.reshape(-1) * step[a2]
    pts[:, an] = third_ang
    return pts


def write_csv(path, plane, points_ang, n, values):
    """Plane CSV in the layout of write_csv_file (plot_module.f90:568-611):
    two in-plane coordinates and the value, E25.16 three spaces apart (the
    trailing 3X of the format prints nothing), a blank line after each run of
    the second coordinate."""
    a1, a2, _ = PLANE_AXES[plane]
    v = np.asarray(values, dtype=np.float64).reshape(-1)
    with open(path, "w") as f:
        for k, (p, x) in enumerate(zip(points_ang, v)):
            x = x if abs(x) > 1e-99 else 0.0
            f.write("   ".join(fortran_e(y, 25, 16) for y in (p[a1], p[a2], x)) + "\n")
            if (k + 1) % n[1] == 0:
                f.write("\n")


def effective_volume(enhancement, lo_ang, hi_ang, step_ang):
    """Effective volume and area of the field in a box, as
    plot_field_or_density with print_effective_volume_area
    (plot_module.f90:202-215, 680-735): V = sum(f^2) dx dy dz / max(f)^2 with
    f = |E| / E0 on the box's grid, A = V / (the box's smallest side).
    Returns (V in nm^3, A in nm^2, smallest side in Angstrom, its axis)."""
    f = np.asarray(enhancement, dtype=np.float64)
    side = np.asarray(hi_ang, dtype=np.float64) - np.asarray(lo_ang, dtype=np.float64)
    k = int(np.argmin(side))
    if np.sum(side == side[k]) > 1 or side[k] < 1e-14:
        raise ValueError("the box has no single smallest side")
    vol = np.sum(f ** 2) * np.prod(step_ang) / np.max(f) ** 2
    return vol / 1e3, vol / side[k] / 1e2, side[k], "XYZ"[k]


def write_plt(path, lo_ang, hi_ang, n, values):
    """Binary grid in the layout of write_plt_file (plot_module.f90:452-495):
    int32 3, 200, nz, ny, nx; float32 z min, z max, y min, y max, x min,
    x max (Angstrom); then the values as float32 with x fastest.  values are
    in cube order (x outermost, as cubic_grid returns the points)."""
    v = np.asarray(values, dtype=np.float64).reshape(n[0], n[1], n[2])
    head = np.array([3, 200, n[2], n[1], n[0]], dtype="<i4")
    box = np.array([lo_ang[2], hi_ang[2], lo_ang[1], hi_ang[1], lo_ang[0], hi_ang[0]], dtype="<f4")
    with open(path, "wb") as f:
        f.write(head.tobytes())
        f.write(box.tobytes())
        f.write(np.ascontiguousarray(v.transpose(2, 1, 0)).astype("<f4").tobytes())


def write_cube(path, atomic_numbers, coord_ang, lo_ang, step_ang, n, values):
    """Gaussian cube in the layout of write_cube_file (plot_module.f90:496-562):
    header in bohr, atoms, then the values x outermost / z innermost, six
    per line (E14.6)."""
    v = np.asarray(values, dtype=np.float64).reshape(n[0], n[1], n[2])
    with open(path, "w") as f:
        f.write("Generated by plasmonx_amd analysis\n\n")
        f.write("%8d%12.6f%12.6f%12.6f\n" % (len(atomic_numbers), *(np.asarray(lo_ang) * TOBOHR)))
        for a in range(3):
            ax = [0.0, 0.0, 0.0]
            ax[a] = step_ang[a] * TOBOHR
            f.write("%8d%12.6f%12.6f%12.6f\n" % (n[a], *ax))
        for z, c in zip(atomic_numbers, coord_ang):
            f.write("%5d %10.6f%12.6f%12.6f%12.6f\n" % (int(z), 0.0, *(np.asarray(c) * TOBOHR)))
        for i in range(n[0]):
            for j in range(n[1]):
                row = v[i, j]
                for s in range(0, n[2], 6):
                    f.write("".join(" " + fortran_e(x if abs(x) > 1e-99 else 0.0) for x in row[s:s + 6]) + "\n")
