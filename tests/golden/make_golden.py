"""Write golden vectors for the oracle from the reference's own Fortran.

Runs the reference's Clawpack-4.6 patch step (clawpack46_step2_wrap with the
shallow-water, Euler and advection Riemann solvers) and the default physical
BC routine, compiled in place by `make -C oracle ref`, on small seeded
inputs, and stores inputs and outputs as .npz fixtures beside this script.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from helpers import random_aux, random_state  # noqa: E402
from oracle_lib import (RP_ADVECTION, RP_EULER, RP_SHALLOW, make_claw, ref,  # noqa: E402
                        ref_available, step2_ref)

CASES = [
    # (name, solver, meqn, mwaves, maux, mx, my, limiter)
    ("swe_mc_8x8", RP_SHALLOW, 3, 3, 0, 8, 8, 4),
    ("swe_superbee_8x6", RP_SHALLOW, 3, 3, 0, 8, 6, 2),
    ("swe_genminmod_6x8", RP_SHALLOW, 3, 3, 0, 6, 8, 6),
    ("euler_mc_8x8", RP_EULER, 4, 4, 0, 8, 8, 4),
    ("euler_vanleer_6x6", RP_EULER, 4, 4, 0, 6, 6, 3),
    ("adv_mc_8x8", RP_ADVECTION, 1, 1, 2, 8, 8, 4),
    ("adv_minmod_8x8", RP_ADVECTION, 1, 1, 2, 8, 8, 1),
]


def main():
    if not ref_available():
        raise SystemExit("oracle/_ref/libref46.so missing: run `make -C oracle ref`")
    for k, (name, solver, meqn, mwaves, maux, mx, my, lim) in enumerate(CASES):
        rng = np.random.default_rng(1000 + k)
        q = random_state(solver, 1, mx, my, rng)[0]
        aux = random_aux(1, mx, my, rng)[0] if maux else None
        dx, dy = 1.0 / mx, 1.0 / my
        dt = 0.3 * min(dx, dy)
        claw = make_claw(solver, meqn, maux, mwaves, mthlim=[lim] * mwaves)
        qout, cfl, (fm, fp, gm, gp) = step2_ref(claw, q, aux, dx, dy, dt, want_fluxes=True)
        np.savez(os.path.join(HERE, "step2_%s.npz" % name), solver=solver, meqn=meqn,
                 mwaves=mwaves, maux=maux, mx=mx, my=my, limiter=lim, dx=dx, dy=dy, dt=dt,
                 q_in=q, aux=np.zeros(0) if aux is None else aux, q_out=qout, cfl=cfl,
                 fm=fm, fp=fp, gm=gm, gp=gp)
    # default physical BC: extrapolation (1) and wall (3) on each face
    rng = np.random.default_rng(77)
    q = random_state(RP_SHALLOW, 1, 6, 5, rng)[0]
    for mth in ([1, 1, 1, 1], [3, 3, 3, 3], [1, 3, -1, 3], [3, -1, 1, -1]):
        out = np.ascontiguousarray(q.copy())
        ref().ref46_bc2(6, 5, 2, 3, out.reshape(-1), np.array(mth, dtype=np.int32))
        np.savez(os.path.join(HERE, "bc2_%s.npz" % "_".join(str(m) for m in mth)),
                 mthbc=np.array(mth), q_in=q, q_out=out)
    print("wrote golden vectors to", HERE)


if __name__ == "__main__":
    main()
