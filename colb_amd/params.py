"""Solver parameters and presets.

The reference keeps its parameters as module constants (abd.py:27-33, barrier.py:4-5,
contact.py:7-10, ortho.py:5, rbd_simple.py:7, links.py:16).  Here they are one dataclass with
two families of presets:

* ``reference_compat`` — the reference's own values (d^ = 0.095, kappa = 1e5, k = 4e8, gravity
  -10 along y, dt = 8 ms, 16 Newton iterations, 1e-5 tolerance, floor 1e-2).
* ``north_star`` — the benchmark configurations (d^ = 1e-3, kappa = 1e7, gravity 9.8 m/s^2
  downwards, ground plane, CCD-filtered line search, cubic barrier).
"""
from dataclasses import dataclass, field, asdict, replace

from ._abi import AbdParams


@dataclass
class Params:
    dt: float = 8e-3
    gravity: tuple = (0.0, -9.8, 0.0)
    dhat: float = 1e-3
    kappa: float = 1e7
    k_contact: float = 1e18
    newton_tol: float = 1e-5
    max_newton: int = 16
    cg_max_iter: int = 2000
    cg_tol: float = 1e-5
    ground_enabled: bool = False
    ccd_enabled: bool = True
    ground_normal: tuple = (0.0, 1.0, 0.0)
    ground_offset: float = 0.0
    ccd_gap_ratio: float = 0.1
    ccd_max_iter: int = 256
    precision: int = 0  # 0 = FP64, 1 = mixed (FP32 assembly/PCG)
    ls_floor: float = 1e-2
    pinned_diag: float = 100.0
    max_candidates: int = 1 << 20
    solver: int = 0  # 0 = PCG, 1 = direct (dense Cholesky per scene), 2 = PCG + direct fallback
    direct_max_dofs: int = 6144  # largest scene (12 x bodies) the direct solve takes
    predictor: int = 0  # first Newton iterate: 0 = q_now (the reference), 1 = q_now + dt qdot,
    #                     2 = the inertial target q~, 3 = learned (NewtonAbd.set_predictor)
    hess_proj: int = 0  # contact Hessian projection: 0 = clamp eigenvalues at 0, 1 = absolute values
    k_max: float = 0.0  # adaptive contact stiffness cap (<= k_contact: fixed stiffness)
    k_raise_gap: float = 0.5  # raise the stiffness when a contact gets closer than this * dhat
    cand_margin: float = None  # candidate-cache margin (length); None: 8 dhat; <= 0: no cache
    pcg_x0: int = 0  # PCG initial guess: 0 = zero (the reference), 1 = the untaken part of the previous direction
    stall_alpha: float = 0.0  # > 0: a scene stops iterating after an accepted step alpha below it (0: the reference)
    debug: int = 0  # bit 0: synchronise and check after every launch; bit 1: also log them;
    #                 bit 2: print broad-phase list sizes; bit 3: one PCG launch per iteration;
    #                 bit 4: raise the persistent PCG's failure flag before its launch (fallback test);
    #                 bit 5: the multi-tile persistent PCG even where one tile per wave fits;
    #                 bit 6: no HIP graph capture of the Newton iteration (eager launches)

    def to_c(self) -> AbdParams:
        p = AbdParams()
        p.dt = self.dt
        for k in range(3):
            p.gravity[k] = self.gravity[k]
            p.ground_normal[k] = self.ground_normal[k]
        p.dhat = self.dhat
        p.kappa = self.kappa
        p.k_contact = self.k_contact
        p.newton_tol = self.newton_tol
        p.max_newton = self.max_newton
        p.cg_max_iter = self.cg_max_iter
        p.cg_tol = self.cg_tol
        p.ground_enabled = int(self.ground_enabled)
        p.ccd_enabled = int(self.ccd_enabled)
        p.ground_offset = self.ground_offset
        p.ccd_gap_ratio = self.ccd_gap_ratio
        p.ccd_max_iter = self.ccd_max_iter
        p.precision = self.precision
        p.ls_floor = self.ls_floor
        p.pinned_diag = self.pinned_diag
        p.max_candidates = self.max_candidates
        p.reserved0 = self.debug
        p.solver = self.solver
        p.direct_max_dofs = self.direct_max_dofs
        p.predictor = self.predictor
        p.hess_proj = self.hess_proj
        p.k_max = self.k_max
        p.k_raise_gap = self.k_raise_gap
        p.cand_margin = 8.0 * self.dhat if self.cand_margin is None else self.cand_margin
        p.pcg_x0 = self.pcg_x0
        p.stall_alpha = self.stall_alpha
        return p

    def replace(self, **kw):
        return replace(self, **kw)

    def as_dict(self):
        return asdict(self)


def reference_compat(**kw) -> Params:
    """abd.py:27-33, barrier.py:4-5, ortho.py:5, links.py:16, gauss_newton.py:110."""
    base = Params(dt=8e-3, gravity=(0.0, -10.0, 0.0), dhat=0.0475 * 2, kappa=1e5, k_contact=4e8,
                  newton_tol=1e-5, max_newton=16, ground_enabled=False, ccd_enabled=False,
                  ls_floor=1e-2, pinned_diag=100.0)
    return replace(base, **kw)


def north_star(**kw) -> Params:
    """Benchmark parameters: d^ 1e-3, kappa 1e7, g = 9.8 m/s^2 along -y, CCD line search."""
    base = Params()
    return replace(base, **kw)
