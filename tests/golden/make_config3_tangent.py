"""Writes tests/golden/config3_tangent.json: the oracle's FP64
tangent-linear run of the 64^3 hyperbolic-sinh stretched box (m = 3),
20 RK4 steps from the seeded initial state, and the derivative of the grain
x-drag J at the final step along three directions: dM = 1, dQ = 1 (n_d =
Q G) and a seeded density direction.  The GPU adjoint's gradients are
checked against these numbers (tests/test_gpu_configs.py).

The inputs are rebuilt from their seeds in the test, so the fixture holds
only the run's parameters and the resulting scalars.

    python tests/golden/make_config3_tangent.py      (a few minutes per direction)
"""
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

CASE = {"kind": "hyst", "n": 64, "m": 3.0, "seed": 0, "amp": 1e-3, "Q": -0.1, "M": 1.0, "tau": 0.1,
        "theta": 100.0, "sv_on": False, "pen_on": True, "dt": 1e-3, "steps": 20, "poisson_tol": 1e-13,
        "v_seed": 31, "v_scale": 1e-2}


def direction(og, case):
    """The seeded density direction: dealiased spectrum of Philox normals."""
    from oracle import oracle as O
    z = O.normals(case["v_seed"], int(np.prod(og.rshape))).reshape(og.rshape)
    return og.dealias23(og.r2c(z)) * case["v_scale"]


def run(which):
    from oracle import oracle as O
    from helpers import cell_volumes, initial_state
    c = CASE
    O.set_threads(max(1, (os.cpu_count() or 3) // 3))
    state, nd, mask, og = initial_state(c["kind"], c["n"], seed=c["seed"], amp=c["amp"], Q=c["Q"], m=c["m"])
    r = O.Rhs(og, M=c["M"], tau=c["tau"], theta=c["theta"], sv_on=c["sv_on"], pen_on=c["pen_on"],
              poisson_mode=O.POISSON_FAS, poisson_tol=c["poisson_tol"], poisson_max=600)
    r.set_fields(np.zeros(og.rshape), nd, mask)
    y = [np.array(f, dtype=np.complex128) for f in state]
    ty = [np.zeros(og.cshape, complex) for _ in range(4)]
    dM, dnd = 0.0, None
    if which == "M":
        dM = 1.0
    elif which == "Q":
        dnd = nd / c["Q"]
    else:
        ty[3] = direction(og, c)
    t0 = time.time()
    for s in range(c["steps"]):
        st = r.rk4_tl(c["dt"], y, ty, dM, dnd)
        assert st == 0, st
        print(which, s, round(time.time() - t0, 1), flush=True)
    J, dJ = r.objective_tl(y[3], ty[3], cell_volumes(og), dnd)
    return which, J, dJ


def main():
    with ProcessPoolExecutor(3) as ex:
        res = list(ex.map(run, ["M", "Q", "v"]))
    out = dict(CASE)
    for which, J, dJ in res:
        out["J"] = J
        out["dJ_d" + which] = dJ
    with open(os.path.join(HERE, "config3_tangent.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
