"""CPU timing of the oracle's Joseph forward + back projection -- TEST INFRASTRUCTURE ONLY.

The CPU baseline of the benchmark: the restatement of elsa's CPU Joseph projector
(elsa/projectors/JosephsMethod.cpp:140-198, OpenMP over rays with atomic adds in the back
projection, as elsa's own) timed on a subset of the poses of a circular cone- or fan-beam
trajectory, with a given number of OpenMP threads: one warm-up run, then the median of
`reps` timed runs. Runs in a process of its own, so that OMP_NUM_THREADS / OMP_PROC_BIND
are read by this process's OpenMP runtime at start-up.

    python -m oracle.cpu_time --dim 3 --vol 256 256 256 --views 360 --det 384 384 \
        --s2c 25600 --c2d 1280 --poses 54 --threads 16 --reps 5

Prints one JSON line: samples per run, the run times, their median, Gsamples/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def cpu_model() -> str:
    """The host CPU's model name as /proc/cpuinfo gives it (what lscpu prints)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--dim", type=int, default=3)
    p.add_argument("--vol", type=int, nargs="+", default=[256, 256, 256])
    p.add_argument("--views", type=int, default=360)
    p.add_argument("--det", type=int, nargs="+", default=[384, 384])
    p.add_argument("--s2c", type=float, default=25600.0)
    p.add_argument("--c2d", type=float, default=1280.0)
    p.add_argument("--poses", type=int, required=True, help="poses of the trajectory timed, spread over the orbit")
    p.add_argument("--threads", type=int, required=True)
    p.add_argument("--reps", type=int, default=5)
    a = p.parse_args(argv)

    import oracle
    from elsa_amd.phantoms import modified_shepp_logan

    g = oracle.circle_geometry(a.dim, tuple(a.vol), a.views, 360, a.s2c, a.c2d, det=tuple(a.det))
    idx = np.linspace(0, a.views - 1, a.poses).round().astype(int)
    sub = oracle.Geom(g.dim, g.vol, g.det, g.P[idx], g.Pinv[idx], g.C[idx], g.ext[idx], g.sdd[idx])
    S = oracle.joseph_sample_count(sub)
    x = modified_shepp_logan(tuple(a.vol[::-1]))

    def run():
        t0 = time.perf_counter()
        y = oracle.joseph(sub, x, nthreads=a.threads)
        oracle.joseph(sub, y, adjoint=True, nthreads=a.threads)
        return time.perf_counter() - t0

    run()  # warm-up
    times = [run() for _ in range(a.reps)]
    med = float(np.median(times))
    print(json.dumps({"threads": a.threads, "poses": a.poses, "views": a.views, "samples": 2 * S,
                      "times_s": [round(t, 4) for t in times], "median_s": round(med, 4),
                      "gsamples_per_s": round(2 * S / med / 1e9, 5), "cpu_model": cpu_model(),
                      "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "")}), flush=True)


if __name__ == "__main__":
    main()
