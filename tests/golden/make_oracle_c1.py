"""March the oracle (the numpy/scipy FP64 restatement, not the reference)
through the whole 64x96 eta_T run and record its end state.

    python tests/golden/make_oracle_c1.py [out.npz]

Workload: eta_T viscosity, deta_T = 1e3, Ra = 1e4, Nx = 64, Nz = 96, Lx = 2
(the unit box on its mirror-doubled carrier), BFBT Stokes route,
theta0 = conductive + 3e-2 cos(pi x) sin(pi z), march_adaptive to
tstop = 5e-3 with the production CFL policy.  The file holds the final
coefficients, the diagnostics, the per-step dt, Stokes iterations and route,
and the states at steps 100, 300 and 525.  It takes a few CPU hours (the
stagnating Stokes solves of the middle of the run go through the dense
rescue).  Device runs of the same workload are compared against it.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

WORKLOAD = dict(viscosity_model="eta_T", deta_T=1e3, Ra=1e4, Nx=64, Nz=96, Lx=2.0,
                precond="bfbt")
TSTOP = 5e-3
KEEP = (100, 300, 525)


def main(out):
    m = oracle.build_oracle(WORKLOAD)
    th0 = m.seed(amp=3e-2, m=1, n=1)
    log, saved = [], {}
    t0 = time.perf_counter()

    def cb(n, t, s):
        log.append((n, t, m.last_bfbt_iters, m.last_stokes_route != "bfbt",
                    time.perf_counter() - t0))
        if n in KEEP:
            saved["s%d" % n] = np.array(s)
        if n % 10 == 0:
            print("step %d t=%.6e bfbt=%d route=%s wall=%.0fs" % (
                n, t, m.last_bfbt_iters, m.last_stokes_route, time.perf_counter() - t0),
                flush=True)

    s, info = oracle.march_adaptive(m, th0, TSTOP, on_step=cb)
    d = m.diagnostics(s)
    np.savez_compressed(out, final=s, log=np.array(log), info=json.dumps(info),
                        diag=json.dumps(d), **saved)
    print(json.dumps(dict(info=info, diag=d, wall=time.perf_counter() - t0)), flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else
         os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_c1_full_64x96.npz"))
