"""Writes floor_history_64x32x16.npz: the reference routines (flang build of
the reference sources, oracle/_ref) run 24,000 implicit stages of the
inclined-free-stream wall block 64x32x16 (configs.adjoint_wall with its
default far field, utnrd.f90 incidence angles alpha = 5, beta = -10; CFL 10,
no bump, no state perturbation; unlimited MUSCL, implimf_3d with
lmax = 10), every routine of the stage the reference's own, as in
make_converged_wall.py.  Stored: the density-residual drop
||(rho^{n+1} - rho^n)/dt||_2 / (its value at stage 1) every 500 stages.  The
drop levels off near 3.8e-6 (minimum near stage 22,000): the residual floor
of this synthetic case is the reference's own.

Run from the repository root:  python tests/golden/make_floor_history.py
(about 13 minutes on one core)."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np

import cpu_stage
import oracle
from codens_amd import configs

STAGES = 24000
EVERY = 500


def main():
    assert oracle.ref_lib() is not None, "the reference build (oracle/_ref) is needed"
    c = configs.adjoint_wall(64, 32, 16, eta=10.0, amp=0.0, bump=0.0)
    st = cpu_stage.CpuStage(c, use_ref=True)
    inner = c.layout.interior()

    def res():
        return float(np.sqrt(np.sum((st.d[0, inner] / st.dt[inner]) ** 2)))
    st.stage()
    r1 = res()
    hist = []
    for it in range(2, STAGES + 1):
        st.stage()
        if it % EVERY == 0:
            hist.append((it, res() / r1))
            print(it, "%.4e" % hist[-1][1], flush=True)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "floor_history_64x32x16.npz"),
                        drop=np.array(hist))


if __name__ == "__main__":
    main()
