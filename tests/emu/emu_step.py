"""Runs the NewtonAbd step on the CPU emulation of the HIP sources (tests/emu/_build/libabd_emu.so,
AddressSanitizer on every allocation) and compares it with the oracle.  TEST INFRASTRUCTURE ONLY:
started by tests/test_emu.py in a child process whose LD_PRELOAD holds the sanitizer runtime.
Prints `result <max relative state error> <active contacts emulated> <active oracle>` and, for two
steps with the learned step predictor (a small untrained model), `learned <feature error> <state
error> <predictor alpha emulated> <predictor alpha oracle>`."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

from colb_amd import _lib  # noqa: E402

_lib.LIB_PATH = os.path.join(HERE, "_build", "libabd_emu.so")

from colb_amd.abd import NewtonAbd  # noqa: E402
from colb_amd.params import north_star  # noqa: E402
from oracle import oracle as O  # noqa: E402
from test_oracle_scene import two_cube_scene  # noqa: E402


def main():
    sc = two_cube_scene(gap=0.01)
    p = north_star(dhat=0.05, k_contact=1e12, kappa=1e5, cg_tol=1e-12, newton_tol=1e-9, cg_max_iter=500,
                   predictor=1, k_max=1e14, k_raise_gap=0.45)
    g = NewtonAbd(p.dt, sc, params=p)
    o = O.Oracle(sc, p)
    for _ in range(2):
        sg, so = g.step()[0], o.step()[0]
    qg, qo = g.get_state()[0], o.get_state()[0]
    print("result", np.abs(qg - qo).max() / np.abs(qo).max(), sg["n_active"], so["n_active"], flush=True)
    g.close()
    # the multi-tile persistent PCG (debug bit 32; one body per tile: 8 tiles over the 4 waves of
    # the emulated grid) against the oracle, on the cubes on the ground
    from colb_amd import scene as S
    sg8 = S.cubes_on_ground(lift=0.02)
    pg = north_star(ground_enabled=True, dhat=0.05, k_contact=1e14, newton_tol=1e-9, cg_tol=1e-10, cg_max_iter=500,
                    max_newton=3)
    gm, om = NewtonAbd(pg.dt, sg8, params=pg.replace(debug=32)), O.Oracle(sg8, pg)
    gm.set_pcg_layout(1, False)
    gm.step(), om.step()
    qm, qo = gm.get_state()[0], om.get_state()[0]
    print("multi", np.abs(qm - qo).max() / np.abs(qo).max(), flush=True)
    gm.close()
    # the assembly's PSD projection (matrix-core sign iteration) against eigendecompositions
    from test_oracle_math import eig_projection, hip_psd_project, psd_cases
    errs = []
    for prec in (0, 1):
        e = 0.0
        for m in (3, 6, 9, 12):
            A = psd_cases(m, seed=m, n=3)
            for absolute in (False, True):
                out = hip_psd_project(_lib.load(), A, prec, absolute)
                e = max(e, max(np.abs(out[i] - eig_projection(A[i], absolute)).max() / np.linalg.norm(A[i])
                               for i in range(len(A))))
        errs.append(e)
    print("psd", *errs, flush=True)
    learned(sc, p)


def learned(sc, p):
    from colb_amd.learned import StepPredictor
    from oracle import mlp_oracle
    sp = StepPredictor(dims=(49, 32, 12), seed=1)
    rng = np.random.default_rng(2)
    sp.feat_mean = rng.normal(size=49).astype(np.float32)
    sp.feat_std = np.exp(rng.normal(size=49)).astype(np.float32)
    sp.out_std = np.full(12, 1e-4, dtype=np.float32)
    p3 = p.replace(predictor=3)
    g = NewtonAbd(p.dt, sc, params=p3)
    o = O.Oracle(sc, p3)
    g.set_predictor(sp)
    g.log_steps(True)
    ferr = serr = 0.0
    ncont = 0
    for _ in range(2):
        fo = o.features()
        ncont += int(fo[:, 24].sum())
        qn, qd = o.get_state()[1:]
        y = mlp_oracle.forward(sp.mlp.get_params(), sp.dims, sp._inputs(fo))[1][-1]
        o.set_prediction(p.dt * qd.reshape(-1, 12) + (y * sp.out_std + sp.out_mean).astype(np.float64))
        sg, so = g.step()[0], o.step()[0]
        fg = g.step_samples()[0][-fo.shape[0]:]
        scale = np.maximum(np.abs(fo).max(0), 1e-30)
        ferr = max(ferr, float((np.abs(fg - fo) / scale).max()))
        qg, qo = g.get_state()[0], o.get_state()[0]
        serr = max(serr, float(np.abs(qg - qo).max() / np.abs(qo).max()))
    assert o.take_error() == 0 and ncont > 0
    print("learned", ferr, serr, sg["predictor_alpha"], so["predictor_alpha"], flush=True)
    g.close()


if __name__ == "__main__":
    main()
