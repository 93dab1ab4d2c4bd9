"""Targets of the frequency-response solve: atomistic (omegaFQ, omegaFQFmu)
and BEM meshes, in the units and conventions the reference's Fortran core
holds after its input stage (atomic units, coordinates in bohr).

Mirrors the data of the reference's target_type / parameters_type
(src/target/target_module.f90:20-131, src/parameters/parameters_module.f90:
64-98) and the host-side preparation steps: Angstrom -> bohr scaling
(src/input_output/input_module.f90:238), Drude density and the
'scaling sigma0-tau' (src/wfq/wfq_module.f90:96-127), nearest-neighbour
counts (src/target/target_module.f90:417-464), tesserae from a GMSH mesh
(src/bem/bem_module.f90:224-491).
"""
from dataclasses import dataclass, field
import math

import numpy as np

from . import _lib

TOBOHR = 1.8897261254578281  # parameters_module.f90:40
AU_TO_EV = 27.211386245981
AU_TO_CM = 219474.63136320
# nearest-neighbour distances r0 in bohr (parameters_module.f90:50-57)
R0_BOHR = {"Ag": 2.885 * TOBOHR, "Au": 2.885 * TOBOHR, "Na": 6.9226100144600462, "C": 1.418 * TOBOHR,
           "Al": 2.861 * TOBOHR, "Cu": 2.543463 * TOBOHR, "O": 1.65, "H": 1.65}


@dataclass
class AtomType:
    name: str
    eta: float
    a_ij: float
    tau: float            # after scaling
    density: float        # n0 = sigma0/tau (or E_F/(pi au_to_ev) for carbon)
    fermi_d: float
    fermi_s: float
    r_q: float = 0.0
    r_mu: float = 0.0
    alpha: float = 0.0
    alpha_w: object = None   # complex array (n_freq,) of the interband polarizability, or None
    permittivity: str = ""   # 'silver etchegoin' / 'gold etchegoin' / 'silver lorentz-drude': alpha_w from a model
    chi: float = 0.0         # electronegativity (static FQ / FQFMu)


@dataclass
class AtomSystem:
    model: str                      # 'wfq' | 'wfqfmu' (dynamic) | 'fq' | 'fqfmu' (static)
    kernel: str                     # 'gaussian' | 'coulomb' | 'ohno'
    symbols: list
    coord_bohr: np.ndarray          # (n, 3)
    types: list                     # [AtomType]
    fermi_d: np.ndarray             # (nt, nt)
    fermi_s: np.ndarray
    r0: np.ndarray                  # (nt, nt) bohr
    type_idx: np.ndarray = None     # (n,) 0-based
    nb_counts: np.ndarray = None    # (n, nt)
    nb_total: np.ndarray = None     # (n,)

    def __post_init__(self):
        names = [t.name for t in self.types]
        if self.type_idx is None:
            self.type_idx = np.array([names.index(s) for s in self.symbols], dtype=np.int32)
        if self.nb_counts is None:
            self.nb_counts, self.nb_total = neighbours(self.coord_bohr, self.type_idx, self.r0)

    @property
    def n_atoms(self):
        return self.coord_bohr.shape[0]

    @property
    def n_types(self):
        return len(self.types)

    @property
    def n_var(self):
        """Unknowns of the interaction matrix T: [q] or [q, mu]."""
        return 4 * self.n_atoms if self.model in ("wfqfmu", "fqfmu") else self.n_atoms

    @property
    def model_id(self):
        return {"wfq": _lib.WFQ, "wfqfmu": _lib.WFQFMU, "fq": _lib.FQ, "fqfmu": _lib.FQFMU}[self.model]

    @property
    def kernel_id(self):
        return {"gaussian": _lib.GAUSSIAN, "coulomb": _lib.COULOMB, "ohno": _lib.OHNO}[self.kernel]

    @property
    def heterogeneous(self):
        return len(set(self.type_idx.tolist())) > 1

    def arr(self, key):
        return np.array([getattr(t, key) for t in self.types], dtype=np.float64)

    eta = property(lambda s: s.arr("eta"))
    r_q = property(lambda s: s.arr("r_q"))
    r_mu = property(lambda s: s.arr("r_mu"))
    a_ij = property(lambda s: s.arr("a_ij"))
    tau = property(lambda s: s.arr("tau"))
    density = property(lambda s: s.arr("density"))

    def alpha_table(self, freq_au):
        """Interband alpha_t(omega) per type and frequency: (nt, nf) complex."""
        from . import permittivity
        out = np.ones((self.n_types, len(freq_au)), dtype=np.complex128)
        for k, t in enumerate(self.types):
            if t.alpha_w is not None:
                out[k] = np.asarray(t.alpha_w, dtype=np.complex128)
            elif t.permittivity == "silver etchegoin":
                out[k] = permittivity.silver_etchegoin(freq_au, polar=True)
            elif t.permittivity == "gold etchegoin":
                out[k] = permittivity.gold_etchegoin(freq_au, polar=True)
            elif t.permittivity == "silver lorentz-drude":
                out[k] = permittivity.lorentz_drude_interband(freq_au)
        return out


def neighbours(coord_bohr, type_idx, r0):
    """Counts of neighbours within r0 by type, self included
    (target_module.f90:417-464)."""
    from scipy.spatial import cKDTree

    n = coord_bohr.shape[0]
    nt = int(np.max(type_idx)) + 1 if n else 0
    nt = max(nt, np.asarray(r0).shape[0])
    counts = np.zeros((n, nt), dtype=np.int32)
    counts[np.arange(n), type_idx] = 1
    rmax = float(np.max(r0)) * (1 + 1e-12)
    tree = cKDTree(coord_bohr)
    pairs = tree.query_pairs(rmax, output_type="ndarray")
    if len(pairs):
        i, j = pairs[:, 0], pairs[:, 1]
        d = np.linalg.norm(coord_bohr[i] - coord_bohr[j], axis=1)
        r0a = np.asarray(r0)
        for a, b in ((i, j), (j, i)):
            ok = d <= r0a[type_idx[a], type_idx[b]]
            np.add.at(counts, (a[ok], type_idx[b[ok]]), 1)
    total = counts.sum(axis=1).astype(np.int32)
    return counts, total


def system_from_fixture(fx):
    """AtomSystem from a JSON fixture of tests/golden (atom cases)."""
    types = []
    for t in fx["types"]:
        aw = None
        if t.get("alpha_w") is not None:
            aw = np.array([complex(a, b) for a, b in t["alpha_w"]])
        types.append(AtomType(name=t["name"], eta=t["eta"], a_ij=t["a_ij"], tau=t["tau"], density=t["density"],
                              fermi_d=t["fermi_d"], fermi_s=t["fermi_s"], r_q=t["rq"], r_mu=t["rmu"],
                              alpha=t.get("alpha", 0.0), alpha_w=aw, permittivity=t.get("permittivity", "")))
    coord = np.array(fx["coord_ang"], dtype=np.float64) * TOBOHR
    return AtomSystem(model=fx["model"], kernel=fx["kernel"], symbols=list(fx["symbols"]), coord_bohr=coord,
                      types=types, fermi_d=np.array(fx["fermi_d"]), fermi_s=np.array(fx["fermi_s"]),
                      r0=np.array(fx["r0"]))


@dataclass
class Mesh:
    """BEM surface: the three vertices of each triangle (bohr) and the keywords
    of the reference's 'bem' section."""
    verts_bohr: np.ndarray        # (n, 3, 3)
    variant: str = "dpcm"         # 'dpcm' | 'iefpcm'
    green: str = "accurate"       # 'accurate' | 'approximate'
    normal_factor: float = 1.0
    sphere_r_bohr: float = 0.0
    eps_solvent: float = 1.0
    charge_constraint: bool = False

    @property
    def n(self):
        return self.verts_bohr.shape[0]

    @property
    def variant_id(self):
        return _lib.IEFPCM if self.variant == "iefpcm" else _lib.DPCM

    @property
    def green_id(self):
        return _lib.ACCURATE if self.green == "accurate" else _lib.APPROXIMATE

    def tesserae(self):
        """Centroids, normals, tangents and areas (bem_module.f90:420-491)."""
        v = self.verts_bohr
        c = v.mean(axis=1)
        a = v[:, 1] - v[:, 0]
        b = v[:, 2] - v[:, 0]
        t1 = -a / np.linalg.norm(a, axis=1)[:, None]
        cr = np.cross(a, b)
        ncr = np.linalg.norm(cr, axis=1)
        area = ncr / 2.0
        normal = self.normal_factor * cr / ncr[:, None]
        t2 = np.cross(normal, t1)
        return c, normal, t1, t2, area


def mesh_from_fixture(fx):
    v = np.array(fx["verts_ang"], dtype=np.float64) * TOBOHR
    return Mesh(verts_bohr=v, variant=fx["variant"], green=fx["green"], normal_factor=fx["normal_factor"],
                sphere_r_bohr=fx["sphere_r_bohr"], eps_solvent=fx["eps_solvent"])


def read_gmsh(path):
    """Triangles of a GMSH 2.2 ASCII file (bem_module.f90:224-415): the
    elements with at least 8 numbers; their last three numbers are node
    labels.  Returns (n, 3, 3) vertex coordinates as written (Angstrom)."""
    lines = open(path).read().splitlines()
    low = [l.strip().lower() for l in lines]
    k = low.index("$nodes")
    nn = int(lines[k + 1])
    nodes = {}
    for l in lines[k + 2:k + 2 + nn]:
        p = l.replace("\t", " ").split()
        nodes[int(p[0])] = (float(p[1]), float(p[2]), float(p[3]))
    k = low.index("$elements")
    ne = int(lines[k + 1])
    tris = []
    for l in lines[k + 2:k + 2 + ne]:
        p = l.replace("\t", " ").split()
        if len(p) >= 8:
            tris.append([nodes[int(p[5])], nodes[int(p[6])], nodes[int(p[7])]])
    return np.array(tris, dtype=np.float64)


def write_gmsh(path, verts):
    """Write (n, 3, 3) triangles as GMSH 2.2 ASCII (2 tags per element), the
    format the reference reads."""
    pts, index, faces = [], {}, []
    for tri in np.asarray(verts):
        f = []
        for p in tri:
            key = tuple(np.round(p, 12))
            if key not in index:
                index[key] = len(pts) + 1
                pts.append(p)
            f.append(index[key])
        faces.append(f)
    with open(path, "w") as fh:
        fh.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(pts))
        for k, p in enumerate(pts):
            fh.write("%d %.16g %.16g %.16g\n" % (k + 1, p[0], p[1], p[2]))
        fh.write("$EndNodes\n$Elements\n%d\n" % len(faces))
        for k, f in enumerate(faces):
            fh.write("%d 2 2 0 1 %d %d %d\n" % (k + 1, f[0], f[1], f[2]))
        fh.write("$EndElements\n")
