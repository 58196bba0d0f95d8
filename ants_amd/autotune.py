"""Autotuner of the SyN kernels' launch shapes.

Measures, on the GPU it runs on, each candidate value of every launch knob
of the tuning table (include/ants_syn.h, ``ants_tuning_*``) -- inversion
blocks per problem, CC and smoothing z-chunk depth, CC and smoothing tile --
for every pyramid-level grid of a set of representative registrations (the
192x224x192 headline pair, a 128^3 serving request and the 96..192 sizes of
the batch workload), and keeps per grid-size bucket (floor(log2(voxels)))
the value with the shortest mean launch time.  Knobs are tuned one after the
other, each with the values already chosen for the previous ones.  Launch
times come from the library's own HIP-event spans per kernel class and level
(``ants_syn_get_kernel_stats_level``), over a forced schedule.

    python -m ants_amd.autotune [--out ants_amd/tuning.txt]

writes the table that the library loads from beside itself
(``tuning.txt``).  Every candidate computes the same registration: tiles and
smoothing z-chunks give the same bits, CC z-chunks and inversion block
counts change only the order of double-precision sums.
"""
from __future__ import annotations

import argparse
import collections
import ctypes as C
import datetime
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

from ants_amd import _lib as L  # noqa: E402
from ants_amd import phantoms as ph  # noqa: E402
from ants_amd.registration import SyNRegistration, SyNStage  # noqa: E402

KNOBS = [
    # name, knob id, kernel class timed, candidate values (0 = the built-in model)
    ("inv_blocks", 0, "inv", [0, 128, 256, 384, 512, 768, 1024]),
    ("cc_tile", 3, "cc", [1, 2]),
    ("cc_zchunk", 1, "cc", [0, 4, 6, 8, 12, 16, 24, 32, 48, 64]),
    ("smooth_tile", 4, "smooth", [1, 2]),
    ("smooth_zchunk", 2, "smooth", [0, 4, 8, 12, 16, 24, 32, 48, 64]),
]
N_BUCKETS = 40

# representative registrations: (size, CC radius) -- their level grids cover
# the grid-size buckets of the headline, serving and batch workloads
SHAPES = [((192, 224, 192), 4), ((128, 128, 128), 2), ((96, 96, 96), 4), ((144, 144, 144), 4),
          ((176, 120, 160), 4), ((112, 168, 104), 2)]
ITERS = [12, 8, 4]


def bucket(voxels: int) -> int:
    return int(np.floor(np.log2(max(voxels, 1))))


def set_all(lib, knob: int, value: int):
    for b in range(N_BUCKETS):
        L.check(lib.ants_tuning_set(knob, b, value))


def measure(cases, cls: str):
    """Mean launch time (ms) of kernel class `cls` per grid-size bucket."""
    acc = collections.defaultdict(list)
    for reg, pair in cases:
        reg.register(pair.fixed, pair.moving)
        for lv in range(len(ITERS)):
            st = reg.kernel_stats(lv)[cls]
            if st["launches"] > 0:
                n = reg.stats()["level_voxels"][lv]
                acc[bucket(n)].append(st["ms"] / st["launches"])
    return {b: float(np.mean(v)) for b, v in acc.items()}


def tune(out_path: str, verbose: bool = True) -> str:
    lib = L.lib()
    L.check(lib.ants_tuning_clear())
    cases = []
    t0 = time.perf_counter()
    for k, (size, r) in enumerate(SHAPES):
        pair = ph.phantom_pair_3d(900 + k, size=size)
        stage = SyNStage.from_args("SyN[0.1,3,0]", f"CC[f,m,1,{r}]", "[%s,1e-6,10]" % "x".join(map(str, ITERS)),
                                   "4x2x1", "2x1x0vox")
        reg = SyNRegistration(stage, 3, force_iterations=True)
        reg.set_profiling(0x1F, every=1)
        reg.register(pair.fixed, pair.moving)  # warm-up: buffers and code objects
        cases.append((reg, pair))
    if verbose:
        print(f"autotune: {len(cases)} representative registrations ready in {time.perf_counter() - t0:.1f} s",
              flush=True)
    chosen = {}
    report = []
    for name, knob, cls, cands in KNOBS:
        times = {}
        for c in cands:
            set_all(lib, knob, c)
            times[c] = measure(cases, cls)
        buckets = sorted(set().union(*[set(t) for t in times.values()]))
        best = {}
        for b in buckets:
            vals = {c: times[c][b] for c in cands if b in times[c]}
            cbest = min(vals, key=vals.get)
            best[b] = cbest
            report.append(f"# {name:13s} bucket {b:2d}: " + ", ".join(
                f"{c}:{vals[c] * 1e3:.1f}us" for c in cands if c in vals) + f" -> {cbest}")
        set_all(lib, knob, 0)
        for b, c in best.items():
            L.check(lib.ants_tuning_set(knob, b, c))
        chosen[name] = best
        if verbose:
            print(f"autotune: {name}: " + ", ".join(f"2^{b}:{c}" for b, c in sorted(best.items())), flush=True)
    for reg, _ in cases:
        reg.close()
    lines = ["# Launch-shape table of libants_syn.so, written by the project's autotuner:",
             "#   python -m ants_amd.autotune --out ants_amd/tuning.txt",
             f"# on {datetime.date.today().isoformat()}, one MI355X; representative registrations "
             + ", ".join(f"{'x'.join(map(str, s))} r={r}" for s, r in SHAPES)
             + f", forced {'x'.join(map(str, ITERS))} iterations, 4x2x1 levels.",
             "# Lines: knob, grid-size bucket floor(log2(level voxels)), value (0 = the built-in launch model).",
             "# Candidate mean launch times per bucket (HIP events):"] + report
    for name, knob, cls, cands in KNOBS:
        for b, c in sorted(chosen[name].items()):
            if c:
                lines.append(f"{name} {b} {c}")
    text = "\n".join(lines) + "\n"
    if out_path:
        with open(out_path, "w") as fh:
            fh.write(text)
    return text


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.txt"))
    a = ap.parse_args(argv)
    print(tune(a.out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
