"""Times the reference's own CPU physics (Physics.update of src/integrator.js
under Node 12, one thread) beside the oracle on one core, on the same inputs:
the benchmark's configuration-2 force set in an open box without relativity
(the reference's relativity turns on signal delay, whose warm-up leaves no
leaf forces), at the reference's constants (theta 0.5, softening 8).  It
shows what the oracle's CPU baseline in bench.py stands for.  Build container
only: the reference is read from its checkout and never travels.

    python tests/golden/time_reference.py [/root/reference] [N] [updates]
"""
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from physsim_amd import _abi as A, workloads as W  # noqa: E402
from oracle import bind  # noqa: E402
from make_reference_goldens import _toggles, DESPAWN, PASS, TORUS  # noqa: E402


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    n = int(sys.argv[2]) if len(sys.argv) > 2 else 16384
    updates = int(sys.argv[3]) if len(sys.argv) > 3 else 2
    st, prm, dt = W.config2(n, open_box=True)
    prm = dict(prm, toggles0=A.GRAVITY | A.COULOMB | A.YUKAWA | A.MAGNETIC | A.BARNES_HUT, theta=0.5,
               softening=8.0, signal_delay=A.DELAY_OFF)
    case = dict(name="time", timeOnly=True, updates=updates, dt=dt, collisionMode=PASS, boundaryMode=DESPAWN,
                yukawaMu=prm["yukawa_mu"],
                topology=TORUS, W=prm["domain_w"], H=prm["domain_h"],
                toggles=_toggles("gravityEnabled", "coulombEnabled", "yukawaEnabled", "magneticEnabled",
                                 "barnesHutEnabled"),
                particles={k: np.asarray(st[k], np.float64).tolist() for k in
                           ("x", "y", "wx", "wy", "mass", "charge", "angw")})
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.json")
        with open(inp, "w") as f:
            json.dump({"cases": [case]}, f)
        env = dict(os.environ, GEON_REF_SRC=os.path.join(ref, "src"))
        r = subprocess.run(["node", "--experimental-loader", "./ref_loader.mjs", "reference_goldens.mjs", inp],
                           cwd=HERE, env=env, capture_output=True, text=True)
        if r.returncode != 0:
            raise SystemExit(r.stderr)
        js = json.loads(r.stdout)["time"]
    bind.set_threads(1)
    o = bind.Oracle(bind.default_params(**prm), n)
    o.upload(st["x"], st["y"], st["wx"], st["wy"], st["mass"], st["charge"], st["angw"])
    o.compute_forces()
    steps, t0 = 0, time.perf_counter()
    for _ in range(updates):
        n0 = o.count()
        steps += o.update(dt) * n0
    sec = time.perf_counter() - t0
    print(json.dumps({"n": n, "updates": updates, "workload": "configuration-2 forces without relativity, open "
                      "%gx%g box, theta 0.5, softening 8, dt %g" % (prm["domain_w"], prm["domain_h"], dt),
                      "reference_node_1thread": {"particle_steps_per_s": js["particle_steps"] / js["seconds"],
                                                 "seconds": js["seconds"]},
                      "oracle_1core": {"particle_steps_per_s": steps / sec, "seconds": sec}}))


if __name__ == "__main__":
    main()
