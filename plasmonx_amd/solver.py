"""Host-side mirror of the reference's frequency-response driver.

``solve_dynamic_field`` plays the role of algorithm%solve_dynamic_field
(src/solvers/algorithm_module.f90:150-179) with its keywords: the target
(omegaFQ / omegaFQFmu atoms or a BEM mesh), the field section (frequencies,
intensity E0, polarization, rhs type) and the algorithm section (method
'inversion' | 'iterative' | 'iterative on the fly', tolerance scaled by
sqrt(n_var) as src/input_output/input_module.f90:251 does, gmres
dimension).  All computation is done by libpxh on the GPU.
"""
from dataclasses import dataclass
import math

import numpy as np

from . import _lib
from .model import AtomSystem, Mesh

METHODS = {"inversion": _lib.INVERSION, "iterative": _lib.ITERATIVE, "iterative on the fly": _lib.ITERATIVE_OTF}
POL = {"x": 1, "y": 2, "z": 4, "all": 7}


@dataclass
class Field:
    freq_au: np.ndarray
    e0: float = 1.0e-4
    polarization: str = "all"
    rhs_form: str = "field"

    @property
    def pol_mask(self):
        return POL[self.polarization]


@dataclass
class Algorithm:
    method: str = "inversion"
    tolerance: float = 1.0e-9       # before the sqrt(n_var) scaling
    gmres_dimension: int = 1000
    precision: str = "fp64"         # 'fp64' | 'mixed'
    batch: int = 1                  # systems sharing each product (1 = one at a time)

    def threshold(self, n_var):
        return self.tolerance * math.sqrt(n_var)


def solve_dynamic_field(target, field, algorithm, ctx=None, device=0, want_variables=False, alpha_w=None,
                        eps_w=None):
    """Solve the frequency-dependent response; returns a dict with 'polar'
    (n_freq, 3, 3) complex, 'results' (n_freq, 20), 'iters', 'true_residual'
    and optionally 'variables' (n_freq, 3, n_var)."""
    own = ctx is None
    ctx = ctx or _lib.Context(device)
    try:
        freq = np.asarray(field.freq_au, dtype=np.float64)
        if isinstance(target, AtomSystem):
            ctx.set_atoms(target)
            if target.model == "wfqfmu":
                ctx.set_alpha(alpha_w if alpha_w is not None else target.alpha_table(freq))
            n_var = target.n_var
        elif isinstance(target, Mesh):
            ctx.set_mesh(target)
            ctx.set_permittivity(eps_w)
            n_var = target.n
        else:
            raise TypeError("target must be an AtomSystem or a Mesh")
        ctx.set_option("batch", algorithm.batch)
        rhs = _lib.RHS_POTENTIAL if field.rhs_form == "potential" else _lib.RHS_FIELD
        prec = _lib.MIXED if algorithm.precision == "mixed" else _lib.FP64
        return ctx.solve_spectrum(n_var, freq, field.e0, rhs_form=rhs, pol_mask=field.pol_mask,
                                  method=METHODS[algorithm.method], tol_abs=algorithm.threshold(n_var),
                                  krylov_dim=algorithm.gmres_dimension, precision=prec,
                                  want_variables=want_variables)
    finally:
        if own:
            ctx.close()


def solve_ground_state(target, mol_count, ctx=None, device=0):
    """algorithm%solve_ground_state (algorithm_module.f90:95-111) of a static
    FQ / FQFMu target: the charges (and dipoles) that minimise the energy
    with each molecule neutral; returns 'variables' (n_var, 1) in the order
    [q, lambda, mu] and 'energy' = -1/2 x^T b (a.u.)."""
    own = ctx is None
    ctx = ctx or _lib.Context(device)
    try:
        ctx.set_atoms(target)
        ctx.set_static(target, mol_count)
        return ctx.solve_static(_lib.ENERGY)
    finally:
        if own:
            ctx.close()


def solve_static_field(target, mol_count, e0=0.0, rhs_form="field", ctx=None, device=0):
    """algorithm%solve_static_field (algorithm_module.f90:114-131) of a static
    FQ / FQFMu target: 'variables' (n_var, 3) for the field e0 along x, y, z
    and, for FQ, 'polar' (3, 3), the static polarizability
    (fq_module.f90:385-597; FQFMu has none, as in the reference)."""
    own = ctx is None
    ctx = ctx or _lib.Context(device)
    try:
        ctx.set_atoms(target)
        ctx.set_static(target, mol_count)
        rhs = _lib.RHS_POTENTIAL if rhs_form == "potential" else _lib.RHS_FIELD
        return ctx.solve_static(_lib.STATIC_FIELD, e0=e0, rhs_form=rhs, want_polar=target.model == "fq")
    finally:
        if own:
            ctx.close()
