"""Times the oracle's C restatement on the workload BASELINE.md measured the
reference on, in this container: the periodic 64^3 gauge wave (A = 0.01,
harmonic gauge, bt_rk4, CFL 0.25) as 8 blocks of 32^3 cells, one block per
host thread, 8 threads.  The reference built from its sources ran that
workload at 2.47e5 point-steps/s here (zone-cycles/omp_wsecond, 8 OpenMP
threads, main.hpp:724-742).  The ratio of the two rates calibrates the
oracle's CPU figure that bench.py reports on the GPU box, where the
reference cannot run.

TEST INFRASTRUCTURE: times the checker, never the product.

Run: python oracle/calibrate_cpu.py [steps, default 8, as the reference's
8 cycles]   (prints one JSON line)
"""
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "gr-athena_amd"))

import oracle as orc  # noqa: E402
from z4c_amd.awa import COMMON, GAUGES  # noqa: E402
from z4c_amd.initial_data import gauge_wave_adm  # noqa: E402
from z4c_amd.z4c import MeshBlockGeom  # noqa: E402

REFERENCE_RATE = 2.47e5  # point-steps/s, 8 threads, measured here


def main(steps=2, threads=8):
    n, sub = 64, 32
    h = 1.0 / n
    o = orc.Params(**GAUGES["harmonic"], **COMMON)
    blocks = []
    for c in range(2):
        for b in range(2):
            for a in range(2):
                g = MeshBlockGeom.cube(sub, h, (-0.5 + a * sub * h, -0.5 + b * sub * h,
                                                -0.5 + c * sub * h))
                og = orc.Geom.block(g.nx, g.dx, g.xmin)
                u = orc.adm_to_z4c(o, og, np.ascontiguousarray(gauge_wave_adm(g, A=0.01, d=1.0)))
                blocks.append((og, u))
    dt = 0.25 * h

    def one(q):
        og, u = blocks[q]
        for _ in range(steps):
            u, _ = orc.rk4_step(o, og, u, dt)
        return u

    def warm(q):
        og, u = blocks[q]
        orc.rk4_step(o, og, u.copy(), dt)

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(warm, range(len(blocks))))  # first touch of every array
        t0 = time.perf_counter()
        list(ex.map(one, range(len(blocks))))
        T = time.perf_counter() - t0
    rate = n ** 3 * steps / T
    print(json.dumps({"workload": "64^3 gauge wave as 8 blocks of 32^3, bt_rk4, oracle C code, "
                                  f"{threads} threads, {steps} steps after one untimed step",
                      "oracle_rate": rate, "reference_rate": REFERENCE_RATE,
                      "oracle_over_reference": rate / REFERENCE_RATE, "seconds": T,
                      "cpus": os.cpu_count()}))


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
