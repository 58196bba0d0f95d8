"""Benchmark: 3-D SyN with the CC metric (r=4) on one 192x224x192 pair.

One step = one complete SyN stage (SyN[0.1,3,0], CC r=4, -c [100x70x20,1e-6,10],
-f 4x2x1, -s 2x1x0vox) on a synthetic ellipsoid-phantom pair whose images
are already resident in HBM.  The metric is voxel-iterations per second
(sum over levels of level voxels x iterations run, per wall second).

Prints one JSON line.  Also reports the HBM roofline of the dominant kernel
class (HIP events on the engine's stream over the timed steps) and the CPU
restatement timed on a bounded sample of the same workload.  The timed steps
launch their iterations kernel by kernel so that the events can bracket the
dominant class (on every 4th iteration of each level: an event between two
launches stalls the second); `untimed_replay` adds the same registration
without events, launched kernel by kernel (the library's default) and
replayed from iteration graphs, measured after the timed region;
`host_buffer_drop_in` the same registration through the host-buffer entry
(PCIe transfers included); `bspline_syn` the same pair and schedule with the
BSplineSyN field model (BSplineSyN[0.1,26,0,3]); `linear_stages` the
Rigid and Affine stages antsRegistrationSyN.sh runs before SyN; `field_stages`
GaussianDisplacementField, Exponential and their B-spline forms with the CC
metric; `apply_transforms` the consumers of the output field (resampling
with each interpolator, CreateJacobianDeterminantImage).  The other half of the metric comes from two
bounded legs on the same GPU: `batch` (64 mixed-size pairs, batched versus
one at a time) and `serve` (a Poisson request stream at a fixed QPS: p50/p99
latency, pairs/s and capacity), summarised in `batch_summary` and
`serve_summary`.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# kernel that carries each profiler class's work (the name rocprofv3 reports)
CLASS_KERNEL = {"warp": "k_warp_pair", "cc": "k_cc3d", "smooth": "k_smooth3d", "compose": "k_compose",
                "inv": "k_inv_step"}
METRIC = "SyN CC voxel-iters/s @192x224x192 vs CPU; pairs/s + p50/p99 latency @QPS"
SCHEDULE = dict(transform="SyN[0.1,3,0]", metric="CC[fixed,moving,1,4]", convergence="[100x70x20,1e-6,10]",
                shrink="4x2x1", smoothing="2x1x0vox")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# iterations of each level whose dominant-class launches are timed with events
SPAN_EVERY = 4

def cpu_baseline(pair, gpu_iters, scale=0.5):
    """The oracle's FP32 build of the same algorithm (images, fields, CC
    window sums and the per-voxel CC formula in float as ITK's float metric
    computes them; sampling coordinates in double: a restatement, not
    ANTs), all host threads, on
    the same inputs with the GPU run's per-level iteration counts scaled by
    `scale` and forced, so that the level mix matches the timed GPU work
    (about 1.2e8 voxel-iterations, some 10-15 s on 16 threads).  Rates are
    reported per level from the oracle's per-iteration wall times."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as o
    iters = [max(1, int(round(scale * n))) if n > 0 else 0 for n in gpu_iters]
    P = o.SynParams(shrink=[4, 2, 1], sigma=[2, 1, 0], iterations=iters, learning_rate=0.1, update_variance=3.0,
                    total_variance=0.0, radius=4, force_iterations=True)
    g = o.Grid(list(pair.size), list(pair.spacing), list(pair.origin))
    threads = o.max_threads()
    t0 = time.perf_counter()
    r = o.syn_register(P, g, pair.fixed, g, pair.moving, precision="f32")
    dt = time.perf_counter() - t0
    vox = [o.shrink_grid(g, s).count for s in P.shrink]
    work = float(sum(v * it for v, it in zip(vox, iters)))
    per_level = {}
    for lv in range(3):
        sec = float(r.seconds[r.level == lv].sum())
        if sec > 0:
            per_level[str(lv)] = {"voxels": vox[lv], "iterations": iters[lv], "seconds": sec,
                                  "voxel_iters_per_s": vox[lv] * iters[lv] / sec}
    return {"value": work / dt, "unit": "voxel-iters/s", "cores": threads, "kind": "port",
            "sample": f"the 192x224x192 pair with the GPU run's per-level iteration counts {list(gpu_iters)} "
                      f"scaled by {scale} and forced ({iters}); FP32 CPU restatement of the SyN stage "
                      f"(oracle/syn_oracle.cpp: float images, fields and CC sums, double coordinates; not ANTs), {threads} OpenMP "
                      f"threads, {work:.3g} voxel-iterations in {dt:.1f} s",
            "per_level": per_level, "seconds": dt}


def measured_traffic(kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed counter summary
    (two rocprofv3 passes, FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel
    of known size), or None when no summary exists for it."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"{kernel}_*pmc_traffic*.txt"))):
        vals = {}
        for ln in open(path):
            if ":" in ln:
                k, v = ln.split(":", 1)
                vals[k.strip()] = v.strip()
        if vals.get("kernel") == kernel and "traffic_bytes_per_launch" in vals:
            return float(vals["traffic_bytes_per_launch"]), os.path.relpath(path, ROOT)
    return None, None


def _gen_pair(args):
    """Process-pool worker: one phantom pair (fields dropped to keep pickling
    small), optionally with a Voronoi label map of the moving phantom."""
    seed, size = args[:2]
    n_labels = args[2] if len(args) > 2 else 0
    from ants_amd import phantoms as ph
    p = ph.phantom_pair_3d(seed, size=size)
    lab = ph.voronoi_labels(seed + 1, p.labels_moving > 0, n_labels) if n_labels else p.labels_moving
    return p.fixed, p.moving, lab


def _pool_pairs(specs):
    import concurrent.futures as cf
    workers = min(16, os.cpu_count() or 4, len(specs))
    with cf.ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(_gen_pair, specs))


QUICK_ARGS = ("SyN[0.1,3,0]", "CC[f,m,1,2]", "[40x20x0,1e-6,10]", "4x2x1", "2x1x0vox")


def batch_leg(args) -> dict:
    """C4: 64 independent 3-D pairs, dims U{96..192} per axis; 16 quick pairs
    (CC r=2, 40x20x0) and 48 with the 192x224x192 schedule.  The batched
    scheduler versus the same pairs one at a time on the same GPU."""
    from ants_amd.registration import SyNStage
    from ants_amd.serving import Batch, Request, run_all
    rng = np.random.default_rng(args.seed)
    n = args.pairs
    specs = [(1000 + k, tuple(int(v) for v in rng.integers(96, 193, 3))) for k in range(n)]
    t0 = time.perf_counter()
    data = _pool_pairs(specs)
    log(f"batch: generated {n} pairs in {time.perf_counter() - t0:.1f} s")
    quick = SyNStage.from_args(*QUICK_ARGS)
    full = SyNStage.from_args(**SCHEDULE)
    reqs = [Request(quick if k < n // 4 else full, mov, fixed=fix, precision=args.precision)
            for k, (fix, mov, _) in enumerate(data)]
    plain = [Request(r.stage, r.moving, fixed=r.fixed, precision="fp32") for r in reqs]
    out = {}
    for name, slots, rq in (("batched", args.slots, reqs), ("one_at_a_time", 1, plain)):
        b = Batch(slots=slots)
        run_all(b, rq[: min(4, n)])  # warm-up: graphs and buffers
        t = time.perf_counter()
        res = run_all(b, rq)
        wall = time.perf_counter() - t
        b.close()
        bad = [r.error for r in res if r.status != 0]
        if bad:
            raise RuntimeError(bad[0])
        vi = sum(r.voxel_iterations for r in res)
        by = sum(r.algorithmic_bytes for r in res)
        out[name] = {"slots": slots, "precision": rq[0].precision, "wall_s": wall, "pairs_per_s": n / wall,
                     "voxel_iters_per_s": vi / wall, "algorithmic_gbs": by / wall / 1e9,
                     "hbm_frac": by / wall / 1e9 / HBM_PEAK_GBS}
        log(f"batch {name}: {n} pairs in {wall:.2f} s ({n / wall:.2f} pairs/s)")
    return {"workload": f"{n} phantom pairs, dims U{{96..192}}^3, {n // 4} quick (r=2, 40x20x0), "
                        f"{n - n // 4} at r=4 100x70x20, images uploaded from host per request",
            "batched": out["batched"], "one_at_a_time": out["one_at_a_time"],
            "speedup_vs_one_at_a_time": out["batched"]["pairs_per_s"] / out["one_at_a_time"]["pairs_per_s"]}


def serve_leg(args) -> dict:
    """C5: Poisson request stream at a fixed QPS: 70% quick SyN at 128^3 (r=2,
    40x20x0), 30% SyN at 192x224x192 (r=4, 100x70x20), against resident
    templates, each warping a 32-label Voronoi map nearest-neighbour.  The
    continuous-batching scheduler versus one request at a time: latency at the
    offered QPS, and capacity (every request queued at once)."""
    from ants_amd import phantoms as ph
    from ants_amd.registration import SyNStage
    from ants_amd.serving import Batch, Request, latency_stats, run_all, serve
    rng = np.random.default_rng(args.seed + 7)
    n = args.requests
    kinds = rng.random(n) < 0.7
    t0 = time.perf_counter()
    big_t = ph.phantom_pair_3d(5000, size=(192, 224, 192))
    small_t = ph.phantom_pair_3d(5001, size=(128, 128, 128))
    specs = [(6000 + k, (128, 128, 128) if kinds[k] else (192, 224, 192), 32) for k in range(n)]
    data = _pool_pairs(specs)
    log(f"serve: generated templates and {n} requests in {time.perf_counter() - t0:.1f} s")
    quick = SyNStage.from_args(*QUICK_ARGS)
    full = SyNStage.from_args(**SCHEDULE)

    def requests(precision):
        return [Request(quick if kinds[k] else full, data[k][1], fixed=None, template_index=1 if kinds[k] else 0,
                        moving_labels=data[k][2], precision=precision) for k in range(n)]

    arrivals = np.cumsum(rng.exponential(1.0 / args.qps, n))
    arrivals -= arrivals[0]
    out = {}
    for name, slots, prec in (("continuous_batching", args.slots, args.precision), ("one_at_a_time", 1, "fp32")):
        b = Batch(slots=slots)
        b.set_template(0, big_t.fixed)
        b.set_template(1, small_t.fixed)
        run_all(b, requests(prec)[:2])  # warm-up
        res = serve(b, requests(prec), arrivals)
        bad = [r.error for r in res if r.status != 0]
        if bad:
            raise RuntimeError(bad[0])
        st = latency_stats(res)
        cap = latency_stats(run_all(b, requests(prec)))
        b.close()
        st.update({"slots": slots, "precision": prec, "capacity_pairs_per_s": cap["pairs_per_s"]})
        out[name] = st
        log(f"serve {name}: p50 {st['p50_s'] * 1e3:.1f} ms p99 {st['p99_s'] * 1e3:.1f} ms {st['pairs_per_s']:.2f} "
            f"pairs/s at {args.qps} QPS; capacity {cap['pairs_per_s']:.2f} pairs/s")
    return {"workload": f"{n} Poisson requests at {args.qps} QPS: 70% 128^3 r=2 40x20x0, 30% 192x224x192 r=4 "
                        "100x70x20, resident templates, 32-label NN warp",
            "qps": args.qps, "continuous_batching": out["continuous_batching"], "one_at_a_time": out["one_at_a_time"],
            "capacity_speedup_vs_one_at_a_time": out["continuous_batching"]["capacity_pairs_per_s"]
            / out["one_at_a_time"]["capacity_pairs_per_s"]}


def bench_batch(args):
    leg = batch_leg(args)
    line = {"metric": "batched SyN pairs/s (64 mixed-size 3-D pairs)", "value": leg["batched"]["pairs_per_s"],
            "unit": "pairs/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": leg["workload"], "slots": args.slots}, **leg}
    print(json.dumps(line), flush=True)


def bench_serve(args):
    leg = serve_leg(args)
    line = {"metric": "served SyN: pairs/s and p50/p99 latency at fixed QPS",
            "value": leg["continuous_batching"]["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1,
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": leg["workload"], "slots": args.slots}, **leg}
    print(json.dumps(line), flush=True)


def apply_leg(pair, seed):
    """antsApplyTransforms subset: a 192x224x192 image (and a 32-label map)
    resampled onto the same grid through [displacement field, affine] with
    Linear, NearestNeighbor, GenericLabel, BSpline[3] (including its
    coefficient decomposition) and LanczosWindowedSinc interpolation, and
    CreateJacobianDeterminantImage of the field.  Algorithmic bytes per
    output voxel: field 3*4 (read in place) + input 4 (gathered once) +
    output 4.  Returns (per-operation timings, the set-up for the CPU leg)."""
    import ctypes as C
    from scipy import ndimage
    from ants_amd import _lib as L
    from ants_amd import apply_transforms as AT
    from ants_amd import phantoms as ph
    g = AT.Geometry(pair.size, pair.spacing, pair.origin)
    rng = np.random.default_rng(seed)
    fld = ndimage.gaussian_filter(rng.standard_normal(g.shape + (3,)), (6, 6, 6, 0)).astype(np.float32)
    fld *= 4.0 / np.abs(fld).max()
    A = np.eye(3) * 1.02 + 0.01 * rng.standard_normal((3, 3))
    ctr = [0.5 * s for s in pair.size]
    labels = ph.voronoi_labels(seed + 1, pair.labels_moving > 0, 32).astype(np.float32)
    d_img, d_lab = L.DeviceArray.from_host(pair.moving), L.DeviceArray.from_host(labels)
    d_fld = L.DeviceArray.from_host(fld)
    d_out = L.DeviceArray(g.shape)
    xf = [AT.DisplacementFieldTransform(d_fld, g), AT.AffineTransform(A, [1.0, -0.5, 0.3], ctr)]
    N = float(np.prod(g.shape))
    bytes_per_voxel = 3 * 4 + 4 + 4
    res = {}
    e0, e1 = L.Event(), L.Event()
    for interp, src in (("Linear", d_img), ("NearestNeighbor", d_lab), ("GenericLabel", d_lab),
                        ("BSpline[3]", d_img), ("LanczosWindowedSinc", d_img)):
        for _ in range(2):
            AT.apply_transforms(g, src, g, xf, interp, out=d_out)
        L.check(L.lib().ants_device_synchronize())
        e0.record()
        reps = 10
        for _ in range(reps):
            AT.apply_transforms(g, src, g, xf, interp, out=d_out)
        e1.record()
        L.check(L.lib().ants_device_synchronize())
        ms = e0.elapsed_ms(e1) / reps
        res[interp] = {"ms": ms, "voxels_per_s": N / (ms * 1e-3),
                       "achieved_gbs": N * bytes_per_voxel / (ms * 1e-3) / 1e9}
        log(f"{interp}: {ms:.3f} ms per 192x224x192 resample")
    # CreateJacobianDeterminantImage of the same field: 12 bytes read per
    # voxel (its neighbours come from the caches) + 4 written
    d_jac = L.DeviceArray(g.shape)
    gdesc = g.desc()
    for _ in range(2):
        L.check(L.lib().ants_jacobian_determinant(C.byref(gdesc), d_fld.ptr, 0, 0, d_jac.ptr, None, None))
    L.check(L.lib().ants_device_synchronize())
    e0.record()
    for _ in range(10):
        L.check(L.lib().ants_jacobian_determinant(C.byref(gdesc), d_fld.ptr, 0, 0, d_jac.ptr, None, None))
    e1.record()
    L.check(L.lib().ants_device_synchronize())
    jms = e0.elapsed_ms(e1) / 10
    res["CreateJacobianDeterminantImage"] = {"ms": jms, "voxels_per_s": N / (jms * 1e-3),
                                             "achieved_gbs": N * 16 / (jms * 1e-3) / 1e9}
    log(f"CreateJacobianDeterminantImage: {jms:.3f} ms per 192x224x192 field")
    return res, (g, fld, A, ctr, N, bytes_per_voxel)


def bench_apply(args):
    """The apply leg alone, with its CPU baseline and roofline line."""
    from ants_amd import phantoms as ph
    pair = ph.phantom_pair_3d(args.seed)
    res, (g, fld, A, ctr, N, bytes_per_voxel) = apply_leg(pair, args.seed)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as o
        ogg = o.og_geom(g.size, g.spacing, g.origin)
        t0 = time.perf_counter()
        o.apply_transforms(ogg, ogg, pair.moving, [("field", ogg, fld), ("affine", A, [1.0, -0.5, 0.3], ctr)], "linear")
        dt = time.perf_counter() - t0
        cpu = {"value": N / dt, "unit": "voxels/s", "cores": o.max_threads(), "kind": "port",
               "sample": f"the full 192x224x192 Linear resample, FP64 CPU restatement, {o.max_threads()} OpenMP threads, "
                         f"{dt:.2f} s"}
    lin = res["Linear"]
    line = {"metric": "antsApplyTransforms voxels/s (192x224x192, [field, affine])", "value": lin["voxels_per_s"],
            "unit": "voxels/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64 coordinates, f32 samples",
            "data": "synthetic", "config": {"workload": "192x224x192 phantom through a displacement field and an affine, "
                                                         "images resident in HBM"},
            "roofline": {"bound": "hbm", "kernel": "k_apply_transforms", "achieved": lin["achieved_gbs"],
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": lin["achieved_gbs"] / HBM_PEAK_GBS,
                         "traffic": None, "algorithmic_bytes_per_voxel": bytes_per_voxel},
            "per_interpolator": res, "cpu_baseline": cpu,
            "speedup_vs_cpu": lin["voxels_per_s"] / cpu["value"] if cpu else None}
    print(json.dumps(line), flush=True)


def bspline_leg(pair, gf, d_fixed, d_moving, d_fwd, d_inv, steps) -> dict:
    """The same pair and schedule with the BSplineSyN field model as
    antsRegistrationSyN.sh -t b sets it (BSplineSyN[0.1,26,0,3]: knot spacing
    26 mm, no total-field fit): timed registrations, then one profiled run
    for the B-spline fit passes (the smoothing class of this model)."""
    from ants_amd import _lib as L
    from ants_amd.registration import SyNRegistration, SyNStage
    sched = dict(SCHEDULE, transform="BSplineSyN[0.1,26,0,3]")
    stage = SyNStage.from_args(**sched, fixed_size=pair.size, fixed_spacing=pair.spacing)
    reg = SyNRegistration(stage, 3)
    reg.register_device(gf, d_fixed, gf, d_moving, d_fwd, d_inv)
    ts = []
    for _ in range(steps):
        L.check(L.lib().ants_device_synchronize())
        t0 = time.perf_counter()
        reg.register_device(gf, d_fixed, gf, d_moving, d_fwd, d_inv)
        L.check(L.lib().ants_device_synchronize())
        ts.append(time.perf_counter() - t0)
    st = reg.stats()
    ms = sum(ts) / len(ts) * 1e3
    reg.set_profiling(0x1F)
    reg.register_device(gf, d_fixed, gf, d_moving, d_fwd, d_inv)
    ks = reg.kernel_stats()
    fit = ks["smooth"]
    reg.close()
    out = {"transform": sched["transform"], "update_mesh_first_level": stage.transform.update_mesh,
           "ms_per_step": ms, "value": st["voxel_iterations"] / (ms * 1e-3), "unit": "voxel-iters/s",
           "iterations_run": st["iterations"],
           "bspline_fit": {"ms": fit["ms"], "launches": fit["launches"],
                           "achieved_gbs": fit["bytes"] / max(fit["ms"], 1e-9) / 1e6,
                           "note": "algorithmic bytes: each fitted field read and written once"},
           "kernel_breakdown_ms": {k: round(v["ms"], 3) for k, v in ks.items()}}
    log(f"BSplineSyN leg: {ms:.1f} ms per registration, iterations {st['iterations']}, "
        f"B-spline fits {fit['ms']:.2f} ms in {fit['launches']} launches")
    return out


def linear_leg(pair, d_fixed, d_moving) -> dict:
    """The linear stages antsRegistrationSyN.sh runs before SyN, on the same
    pair: -r [fixed,moving,1], then Rigid[0.1] and Affine[0.1], each with
    MI[fixed,moving,1,32,Regular,0.25] (Mattes), -c [1000x500x250x100,1e-6,10],
    -f 8x4x2x1, -s 3x2x1x0vox; both stages timed with the host clock."""
    from ants_amd import _lib as L
    from ants_amd import linear as LN
    from ants_amd.apply_transforms import Geometry
    from ants_amd.registration import initial_transform_from_images
    geo = Geometry(pair.size, pair.spacing, pair.origin, np.eye(3))
    stages = [LN.LinearStage.from_args(f"{t}[0.1]", "MI[fixed,moving,1,32,Regular,0.25]",
                                       "[1000x500x250x100,1e-6,10]", "8x4x2x1", "3x2x1x0vox", seed=1)
              for t in ("Rigid", "Affine")]

    def run():
        # cold start: the Rigid stage makes and uploads its sample points, the
        # Affine stage reuses them (same seed and levels), as in one
        # antsRegistrationSyN.sh run
        LN.release_points()
        chain = [initial_transform_from_images(pair.fixed, geo, pair.moving, geo, 1)]
        out = []
        for st in stages:
            L.check(L.lib().ants_device_synchronize())
            t0 = time.perf_counter()
            r = LN.register_linear(st, d_fixed, geo, d_moving, geo, chain)
            L.check(L.lib().ants_device_synchronize())
            out.append((time.perf_counter() - t0, r))
            chain = [r.transform] + chain
        return out
    run()  # warm-up: library set-up and first-touch of the level buffers
    res = run()
    legs = {}
    for st, (sec, r) in zip(stages, res):
        legs[st.transform] = {"ms": sec * 1e3, "iterations_run": r.iterations, "metric_evaluations": r.evaluations,
                              "ms_per_evaluation": sec * 1e3 / max(r.evaluations, 1)}
    log("linear stages: " + ", ".join(f"{k} {v['ms']:.1f} ms ({v['iterations_run']}, {v['metric_evaluations']} "
                                      f"metric passes)" for k, v in legs.items()))
    return {"stages": "-r [f,m,1]; Rigid[0.1] then Affine[0.1], MI[f,m,1,32,Regular,0.25], "
                      "-c [1000x500x250x100,1e-6,10] -f 8x4x2x1 -s 3x2x1x0vox (antsRegistrationSyN.sh)",
            **legs, "total_ms": sum(v["ms"] for v in legs.values())}


def field_leg(pair, d_fixed, d_moving) -> dict:
    """The dense field-transform stages on the same pair and schedule with the
    CC metric r=4: GaussianDisplacementField[0.25,3,0] (conjugate-gradient
    line search), Exponential[40,3,0.5,0] (gradient descent on the velocity,
    exp by scaling and squaring; its step multiplies the raw CC derivative,
    hence the large value) and their B-spline forms with a 26 mm knot
    spacing; one timed run each after a warm-up."""
    from ants_amd import _lib as L
    from ants_amd import fieldmodels as FM
    from ants_amd.apply_transforms import Geometry
    geo = Geometry(pair.size, pair.spacing, pair.origin, np.eye(3))
    out = {}
    for tr in ("GaussianDisplacementField[0.25,3,0]", "Exponential[40,3,0.5,0]",
               "BSplineDisplacementField[0.25,26,0,3]", "BSplineExponential[40,26,0,0,3]"):
        st = FM.FieldStage.from_args(tr, "CC[fixed,moving,1,4]", SCHEDULE["convergence"], SCHEDULE["shrink"],
                                     SCHEDULE["smoothing"], pair.size, pair.spacing)
        FM.register_field(st, d_fixed, geo, d_moving, geo)
        L.check(L.lib().ants_device_synchronize())
        t0 = time.perf_counter()
        res = FM.register_field(st, d_fixed, geo, d_moving, geo)
        sec = time.perf_counter() - t0
        work = sum(int(np.prod(FM.level_shape(st, geo, lv))) * it for lv, it in enumerate(res.iterations))
        out[tr] = {"ms": sec * 1e3, "iterations_run": res.iterations, "voxel_iters_per_s": work / sec,
                   "energy_first_last": [res.records[0][2], res.records[-1][2]] if res.records else None}
        log(f"field stage {tr}: {sec * 1e3:.1f} ms, iterations {res.iterations}")
    return out


def bench_stages(args):
    """The linear and field-transform stage legs on the headline pair, one
    JSON line (for profiling them apart from the SyN headline)."""
    from ants_amd import _lib as L
    from ants_amd import phantoms as ph
    pair = ph.phantom_pair_3d(args.seed)
    d_fixed = L.DeviceArray.from_host(pair.fixed)
    d_moving = L.DeviceArray.from_host(pair.moving)
    line = {"mode": "stages", "linear_stages": linear_leg(pair, d_fixed, d_moving),
            "field_stages": field_leg(pair, d_fixed, d_moving)}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--mode", choices=["pair", "batch", "serve", "apply", "stages"], default="pair",
                    help="pair: the 192x224x192 headline; batch: 64 mixed pairs; serve: Poisson request stream; "
                         "apply: antsApplyTransforms resampling; stages: the linear and field-transform stage "
                         "legs alone")
    ap.add_argument("--pairs", type=int, default=64)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--qps", type=float, default=30.0)
    ap.add_argument("--precision", choices=["fp32", "bf16"], default="fp32",
                    help="batch/serve modes: image pyramid precision of the batched form")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="pair mode: first copy MIB MiB three times with the calibration copy kernel "
                         "(the known-size reference for HBM byte counters)")
    ap.add_argument("--no-batch-serve", action="store_true",
                    help="pair mode: skip the batched (64 pairs) and served (Poisson stream) legs")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="HIP hardware queues (0: the runtime's default of 4, which measured fastest for 8 slots)")
    args = ap.parse_args()
    if args.hw_queues > 0 and "GPU_MAX_HW_QUEUES" not in os.environ:
        # read by the HIP runtime at initialisation, before the first device call
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(args.hw_queues, 32))
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for {args.gpus}); "
              "pairs would shard across GPUs with no communication", file=sys.stderr)
        sys.exit(2)
    if args.mode == "batch":
        return bench_batch(args)
    if args.mode == "serve":
        return bench_serve(args)
    if args.mode == "apply":
        return bench_apply(args)
    if args.mode == "stages":
        return bench_stages(args)

    from ants_amd import _lib as L
    from ants_amd import phantoms as ph
    from ants_amd.registration import SyNRegistration, SyNStage

    t0 = time.perf_counter()
    pair = ph.phantom_pair_3d(args.seed)
    log(f"generated 192x224x192 phantom pair in {time.perf_counter() - t0:.1f} s")
    stage = SyNStage.from_args(**SCHEDULE)
    reg = SyNRegistration(stage, 3)
    gf = L.desc(pair.size, pair.spacing, pair.origin)
    d_fixed = L.DeviceArray.from_host(pair.fixed)
    d_moving = L.DeviceArray.from_host(pair.moving)
    out_shape = pair.fixed.shape + (3,)
    d_fwd = L.DeviceArray(out_shape)
    d_inv = L.DeviceArray(out_shape)
    if args.copy_calibration > 0:
        nb = args.copy_calibration << 20
        src, dst = L.DeviceArray((nb // 4,)).zero(), L.DeviceArray((nb // 4,))
        for _ in range(3):
            L.check(L.lib().ants_copy_calibration(dst.ptr, src.ptr, nb, None))
        L.check(L.lib().ants_device_synchronize())
        src.free()
        dst.free()
        log(f"calibration copy: 3 x {nb} bytes")

    # profiled run (all kernel classes) to name the dominant class
    reg.set_profiling(0x1F)
    reg.register_device(gf, d_fixed, gf, d_moving, d_fwd, d_inv)
    ks = reg.kernel_stats()
    dominant = max(ks, key=lambda k: ks[k]["ms"])
    # per kernel class and pyramid level: mean launch time and algorithmic
    # bytes per launch -> achieved GB/s against the 8 TB/s HBM peak
    st0 = reg.stats()
    per_level = {}
    for lv in range(st0["levels_run"]):
        kl = reg.kernel_stats(lv)
        lg = reg.level_geometry(pair.size, pair.spacing, pair.origin, lv)
        grid = "x".join(str(int(lg.size[d])) for d in range(3))
        row = {}
        for k, v in kl.items():
            if v["launches"] == 0 or v["ms"] <= 0:
                continue
            gbs = v["bytes"] / (v["ms"] * 1e-3) / 1e9
            row[CLASS_KERNEL[k]] = {"launches": v["launches"], "mean_us": v["ms"] * 1e3 / v["launches"],
                                    "algorithmic_mb_per_launch": v["bytes"] / v["launches"] / 1e6,
                                    "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS}
        per_level[grid] = row
    log("per level (kernel: mean us, GB/s):", json.dumps({g: {k: (round(v["mean_us"], 1), round(v["achieved_gbs"]))
                                                             for k, v in r.items()} for g, r in per_level.items()}))
    log("kernel classes (ms, launches, GB/s):",
        {k: (round(v["ms"], 2), v["launches"], round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)) for k, v in ks.items()})
    # the timed steps time the dominant class on every 4th iteration of each
    # level: every timing span stalls the launch after it until the kernels
    # before it finish, so spans on all iterations would slow the steps
    reg.set_profiling(1 << L.KC_NAMES.index(dominant), every=SPAN_EVERY)

    for _ in range(max(args.warmup - 1, 0)):
        reg.register_device(gf, d_fixed, gf, d_moving, d_fwd, d_inv)
    L.check(L.lib().ants_device_synchronize())

    times, work, traces, kacc = [], 0.0, [], {"ms": 0.0, "launches": 0, "bytes": 0.0}
    for k in range(args.steps):
        L.check(L.lib().ants_device_synchronize())
        t = time.perf_counter()
        tr = reg.register_device(gf, d_fixed, gf, d_moving, d_fwd, d_inv)
        L.check(L.lib().ants_device_synchronize())
        dt = time.perf_counter() - t
        st = reg.stats()
        kd = reg.kernel_stats()[dominant]
        for key in kacc:
            kacc[key] += kd[key]
        times.append(dt)
        work += st["voxel_iterations"]
        traces.append(tr)
        log(f"step {k}: {dt * 1e3:.1f} ms, iterations {st['iterations']}, inner inversion steps {st['inv_steps']}, "
            f"graph/eager iterations {st['graph_launches']}/{st['eager_iterations']}")
    total = sum(times)
    value = work / total
    st = reg.stats()

    # After the timed region: the same registration with no timing spans, as
    # the library runs it by default (kernel-by-kernel launches), and with
    # its iterations replayed from captured graphs (ANTS_SYN_GRAPHS=1; one
    # untimed run first, which captures them).
    reg.set_profiling(0, every=1)
    saved_env = os.environ.get("ANTS_SYN_GRAPHS")

    def untimed_replay(graphs: bool):
        os.environ["ANTS_SYN_GRAPHS"] = "1" if graphs else "0"
        reg.register_device(gf, d_fixed, gf, d_moving, d_fwd, d_inv)
        ts = []
        for _ in range(args.steps):
            L.check(L.lib().ants_device_synchronize())
            t0 = time.perf_counter()
            reg.register_device(gf, d_fixed, gf, d_moving, d_fwd, d_inv)
            L.check(L.lib().ants_device_synchronize())
            ts.append(time.perf_counter() - t0)
        rs = reg.stats()
        ms = sum(ts) / len(ts) * 1e3
        log(f"{'graph replay' if graphs else 'default launches'} (no timing spans): {ms:.1f} ms per registration, "
            f"graph/eager iterations {rs['graph_launches']}/{rs['eager_iterations']}")
        return {"ms_per_step": ms, "value": rs["voxel_iterations"] / (ms * 1e-3)}

    # inverse consistency of the returned pair, FP64 sums on the device, away
    # from the faces (16 voxels), in voxel units
    mx, mn, cnt = C.c_double(), C.c_double(), C.c_int64()
    L.check(L.lib().ants_op_inverse_consistency(C.byref(gf), d_fwd.ptr, d_inv.ptr, 16, C.byref(mx), C.byref(mn),
                                                C.byref(cnt), None))
    ic = {"margin_voxels": 16, "max_voxels": mx.value, "mean_voxels": mn.value, "voxels": cnt.value}
    log(f"inverse consistency |fwd + inv(x + fwd)|: max {mx.value:.3f}, mean {mn.value:.4f} voxel")

    plain_replay = untimed_replay(False)
    graph_replay = untimed_replay(True)

    # the drop-in with host buffers (ants_syn_register): uploads of both
    # images and downloads of both fields over PCIe included; never `value`
    reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    th = []
    for _ in range(min(args.steps, 5)):
        t0 = time.perf_counter()
        reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
        th.append(time.perf_counter() - t0)
    host_ms = sum(th) / len(th) * 1e3
    host_drop_in = {"ms_per_step": host_ms, "value": reg.stats()["voxel_iterations"] / (host_ms * 1e-3),
                    "note": "ants_syn_register with host buffers: 2 image uploads and 2 field downloads over PCIe "
                            "included"}
    log(f"host-buffer drop-in: {host_ms:.1f} ms per registration")
    if saved_env is None:
        os.environ.pop("ANTS_SYN_GRAPHS", None)
    else:
        os.environ["ANTS_SYN_GRAPHS"] = saved_env
    bsp = bspline_leg(pair, gf, d_fixed, d_moving, d_fwd, d_inv, args.steps)
    lin = linear_leg(pair, d_fixed, d_moving)
    fld = field_leg(pair, d_fixed, d_moving)
    apl, _ = apply_leg(pair, args.seed)
    avg_ms = kacc["ms"] / max(kacc["launches"], 1)
    bytes_per_launch = kacc["bytes"] / max(kacc["launches"], 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    whole_gbs = st["algorithmic_bytes"] / (total / args.steps) / 1e9
    traffic, traffic_src = measured_traffic(CLASS_KERNEL[dominant])

    cpu = None
    if not args.no_cpu:
        log("timing the CPU restatement on a bounded sample ...")
        cpu = cpu_baseline(pair, st["iterations"][:3])
        log(f"cpu baseline: {cpu['value']:.4g} voxel-iters/s in {cpu['seconds']:.1f} s")
        cpu.pop("seconds")

    legs = {}
    if not args.no_batch_serve:
        reg.close()
        legs["batch"] = batch_leg(args)
        legs["serve"] = serve_leg(args)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3-D SyN CC r=4 on one 192x224x192 ellipsoid-phantom pair (1 mm), SyN[0.1,3,0], "
                               "-c [100x70x20,1e-6,10] -f 4x2x1 -s 2x1x0vox, images resident in HBM",
                   "global_batch": 1, "iterations_run": st["iterations"], "inv_steps": st["inv_steps"],
                   "parallelism": "single GPU"},
        "pairs_per_s": args.steps / total,
        "untimed_replay": {"default_launches": plain_replay, "graph_replay": graph_replay,
                           "note": "after the timed region, no event spans; default = kernel-by-kernel launches"},
        "whole_job_algorithmic_gbs": whole_gbs,
        "roofline": {"bound": "hbm", "kernel": CLASS_KERNEL[dominant], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src, "avg_launch_ms": avg_ms,
                     "timed_launches_per_step": kacc["launches"] / args.steps,
                     "timed_iterations": f"every {SPAN_EVERY}th iteration of each level",
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "kernel_breakdown_ms": {k: round(v["ms"], 3) for k, v in ks.items()},
        "per_level_kernels": per_level,
        "host_buffer_drop_in": host_drop_in,
        "inverse_consistency": ic,
        "bspline_syn": bsp,
        "linear_stages": lin,
        "field_stages": fld,
        "apply_transforms": {k: {"ms": round(v["ms"], 4), "voxels_per_s": v["voxels_per_s"]} for k, v in apl.items()},
    }
    if legs:
        sv = legs["serve"]["continuous_batching"]
        line["serve_summary"] = {"qps": args.qps, "pairs_per_s": sv["pairs_per_s"], "p50_ms": sv["p50_s"] * 1e3,
                                 "p99_ms": sv["p99_s"] * 1e3, "capacity_pairs_per_s": sv["capacity_pairs_per_s"]}
        line["batch_summary"] = {"pairs_per_s": legs["batch"]["batched"]["pairs_per_s"],
                                 "speedup_vs_one_at_a_time": legs["batch"]["speedup_vs_one_at_a_time"]}
        line.update(legs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
