"""Runs the NewtonAbd step on the CPU emulation of the HIP sources (tests/emu/_build/libabd_emu.so,
AddressSanitizer on every allocation) and compares it with the oracle.  TEST INFRASTRUCTURE ONLY:
started by tests/test_emu.py in a child process whose LD_PRELOAD holds the sanitizer runtime.
Prints one line `result <max relative state error> <active contacts emulated> <active oracle>`."""
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


if __name__ == "__main__":
    main()
