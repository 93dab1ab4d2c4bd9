"""The benchmark workloads: deterministic synthetic targets with the
reference's default parameters (python/input_/validation.py:961-971 atom
types, :1134-1145 parameters; unit rules :8-140).

C1  omegaFQ Na fcc sphere a=4.23 A, r=2 nm (1745 atoms), 61 w in 2.5-4.0 eV, x field
C2  omegaFQFmu Ag fcc sphere a=4.09 A, r=4 nm (15683 atoms, 62732 unknowns),
    201 w in 2.8-4.2 eV, 3 fields, interband alpha(w) from the Lorentz-Drude
    silver permittivity (Rakic et al. 1998), passed as a 'wfqfmu file' table
C3  BEM DPCM / IEF on an icosphere R=10 nm (5120 triangles) and a capped rod
C4  96 seeded graphene flakes, 128 w in 0.1-1.0 eV
C5  a seeded Poisson stream of spectrum requests (omegaFQ Na rods, omegaFQFmu
    Au@Ag core-shell spheres, BEM rods)
"""
import math

import numpy as np

from . import geometry, permittivity
from .model import R0_BOHR, TOBOHR, AU_TO_EV, AtomSystem, AtomType, Mesh

SIGMA_SI_TO_AU = 0.00000021739
TAU_FS_TO_AU = 41.3413733351702


def ev(x):
    return np.asarray(x, dtype=np.float64) / AU_TO_EV


def silver_type():
    tau = 1633.608
    sigma0 = 6.5313e7 * SIGMA_SI_TO_AU
    return AtomType(name="Ag", eta=0.4829838, a_ij=9.61, tau=tau, density=sigma0 / tau, fermi_d=12.0, fermi_s=1.1,
                    r_q=3.3775793621738925, r_mu=2.368922250085755, alpha=49.9843, permittivity="silver etchegoin")


def gold_type():
    tau = 318.018
    sigma0 = 11834849.81 * SIGMA_SI_TO_AU
    return AtomType(name="Au", eta=0.52685, a_ij=9.61, tau=tau, density=sigma0 / tau, fermi_d=12.0, fermi_s=1.1,
                    r_q=3.4996572071713126, r_mu=3.0710227258450864, alpha=39.5297, permittivity="gold etchegoin")


def c1_sodium_sphere():
    """C1: tau = 1.5e-14 s, sigma0 from w_p = 5.9 eV, Gaussian R_q = 2.5 bohr,
    eta = sqrt(2/pi)/R_q, A_ij Na default, Fermi d = 12, s r0 = 1.1 l_nn with
    r0 = RNa0 (parameters_module.f90:53)."""
    xyz = geometry.fcc_sphere(4.23, 20.0)
    tau = 1.5e-14 * TAU_FS_TO_AU * 1e15
    wp = 5.9 / AU_TO_EV
    sigma0 = wp ** 2 * tau / (4.0 * math.pi)
    l_nn = 4.23 / math.sqrt(2.0) * TOBOHR
    s = 1.1 * l_nn / R0_BOHR["Na"]
    rq = 2.5
    t = AtomType(name="Na", eta=math.sqrt(2.0 / math.pi) / rq, a_ij=12.07910025, tau=tau, density=sigma0 / tau,
                 fermi_d=12.0, fermi_s=s, r_q=rq)
    sysm = AtomSystem(model="wfq", kernel="gaussian", symbols=["Na"] * len(xyz), coord_bohr=xyz * TOBOHR, types=[t],
                      fermi_d=np.array([[12.0]]), fermi_s=np.array([[s]]), r0=np.array([[R0_BOHR["Na"]]]))
    freq = ev(np.linspace(2.5, 4.0, 61))
    return sysm, freq


def c2_silver_sphere(radius=40.0, n_freq=201):
    """C2: the Ag defaults with the Lorentz-Drude interband polarizability
    (alpha_table evaluates it; c2_alpha_table writes and reads it as the
    reference's 'wfqfmu file' table)."""
    xyz = geometry.fcc_sphere(4.09, radius)
    t = silver_type()
    t.permittivity = "silver lorentz-drude"
    sysm = AtomSystem(model="wfqfmu", kernel="gaussian", symbols=["Ag"] * len(xyz), coord_bohr=xyz * TOBOHR,
                      types=[t], fermi_d=np.array([[12.0]]), fermi_s=np.array([[1.1]]), r0=np.array([[R0_BOHR["Ag"]]]))
    freq = ev(np.linspace(2.8, 4.2, n_freq))
    return sysm, freq


def c2_alpha_table(freq, path):
    """The C2 interband alpha(w) written as a 'freq (eV) Re Im' table at path
    and read back at the spectrum's frequencies, the route of the
    reference's 'wfqfmu file' (parameters_module.f90:384-462): (1, n_freq)."""
    permittivity.write_table(path, freq, permittivity.lorentz_drude_interband(freq))
    return permittivity.read_table(path, freq)[None, :]


def core_shell_sphere(r_core, r_shell, a=4.08):
    sym, xyz = geometry.core_shell(a, r_core, r_shell)
    types = [gold_type(), silver_type()]
    fd = np.full((2, 2), 12.0)
    fs = np.array([[1.1, 0.95], [0.95, 1.1]])
    r0 = np.full((2, 2), R0_BOHR["Ag"])
    return AtomSystem(model="wfqfmu", kernel="gaussian", symbols=sym, coord_bohr=xyz * TOBOHR, types=types,
                      fermi_d=fd, fermi_s=fs, r0=r0)


def graphene_flake(xyz, fermi_energy_ev, tau_s=1e-13):
    tau = tau_s * TAU_FS_TO_AU * 1e15
    t = AtomType(name="C", eta=0.372124, a_ij=1.7424, tau=tau, density=fermi_energy_ev / (math.pi * AU_TO_EV),
                 fermi_d=100.0, fermi_s=1.2, r_q=math.sqrt(2.0 / math.pi) / 0.372124)
    return AtomSystem(model="wfq", kernel="gaussian", symbols=["C"] * len(xyz), coord_bohr=np.asarray(xyz) * TOBOHR,
                      types=[t], fermi_d=np.array([[100.0]]), fermi_s=np.array([[1.2]]),
                      r0=np.array([[R0_BOHR["C"]]]))


def c3_meshes(variant="dpcm"):
    sphere = Mesh(verts_bohr=geometry.icosphere(100.0, 4, seed=0) * TOBOHR, variant=variant, green="accurate")
    rod = Mesh(verts_bohr=geometry.capped_rod(400.0, 100.0, n_ring=80, n_cap=20) * TOBOHR, variant=variant,
               green="accurate")
    freq = ev(np.linspace(3.0, 4.0, 400))
    return sphere, rod, freq


def c3_permittivity(freq):
    bb = permittivity.BB_SILVER
    return permittivity.brendel_bormann(freq, bb["eps_inf"], bb["wp_ev"], bb["f0"], bb["g0_ev"], bb["osc"])


def c4_flakes(n_flakes=96, seed=20261203, lo=500, hi=12000):
    """C4: the seeded flakes as omegaFQ systems (E_F per flake), 128 w in
    0.1-1.0 eV, in-plane fields x and y."""
    out = []
    for f in geometry.graphene_batch(n_flakes, seed, lo, hi):
        out.append(graphene_flake(f["xyz"], f["fermi_energy"]))
    return out, ev(np.linspace(0.1, 1.0, 128))


def sodium_rod(length_ang, radius_ang):
    """omegaFQ Na rod with the C1 parameters."""
    base, _ = c1_sodium_sphere()
    xyz = geometry.fcc_rod(4.23, length_ang, radius_ang)
    return AtomSystem(model="wfq", kernel="gaussian", symbols=["Na"] * len(xyz), coord_bohr=xyz * TOBOHR,
                      types=base.types, fermi_d=base.fermi_d, fermi_s=base.fermi_s, r0=base.r0)


def c5_requests(n_requests=24, seed=7, rate=200.0, atoms=(2000, 12000), tesserae=(2000, 8000), n_freq=(50, 100),
                rod_atoms=None):
    """C5: a seeded Poisson stream.  Each request is a dict with its arrival
    time (s), kind, a builder of its target and its frequency list (eV ranges
    as the C1-C3 configurations).  Sizes are drawn log-uniform in the given
    ranges (the full configuration is 2k-60k atoms for the core-shell
    spheres, 2k-20k tesserae, 50-400 frequencies; the bounds are parameters).
    rod_atoms bounds the omegaFQ Na rods separately (default: atoms)."""
    rod_atoms = rod_atoms or atoms
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n_requests):
        t += float(rng.exponential(1.0 / rate))
        kind = ["wfq_rod", "wfqfmu_coreshell", "bem_rod"][int(rng.integers(0, 3))]
        nf = int(rng.integers(n_freq[0], n_freq[1] + 1))
        if kind == "bem_rod":
            nt = float(np.exp(rng.uniform(np.log(tesserae[0]), np.log(tesserae[1]))))
            aspect = float(rng.uniform(2.0, 4.0))
            radius = 60.0
            # tesserae ~ 2 n_ring^2 (aspect - 1) / pi + 4 n_ring n_cap with n_cap = n_ring / 4
            n_ring = max(16, int(round(np.sqrt(nt / (2 * (aspect - 1) / np.pi + 1.0)))))
            n_cap = max(4, n_ring // 4)
            verts = geometry.capped_rod(aspect * 2 * radius, radius, n_ring=n_ring, n_cap=n_cap) * TOBOHR
            build = (lambda v=verts: Mesh(verts_bohr=v, variant="dpcm", green="accurate"))
            freq = ev(np.linspace(3.0, 4.0, nf))
        elif kind == "wfq_rod":
            na = float(np.exp(rng.uniform(np.log(rod_atoms[0]), np.log(rod_atoms[1]))))
            aspect = float(rng.uniform(2.0, 4.0))
            # fcc Na: 4 atoms per 4.23^3 A^3; volume of a capped rod pi r^2 (L - 2r/3)
            radius = (na * 4.23 ** 3 / 4 / (np.pi * (2 * aspect - 2.0 / 3.0))) ** (1.0 / 3.0)
            build = (lambda L=2 * aspect * radius, r=radius: sodium_rod(L, r))
            freq = ev(np.linspace(2.5, 4.0, nf))
        else:
            na = float(np.exp(rng.uniform(np.log(atoms[0]), np.log(atoms[1]))))
            r_shell = (na * 4.08 ** 3 / 4 / (4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)
            build = (lambda rs=r_shell: core_shell_sphere(0.6 * rs, rs))
            freq = ev(np.linspace(2.8, 4.2, nf))
        out.append({"arrival": t, "kind": kind, "build": build, "freq": freq})
    return out
