"""The batched (C4) and served (C5) workloads at their own sizes.

C4: pairs from the benchmark's 64-pair list (the same seeds and dims, drawn
U{96..192} per axis; quick pairs CC r=2 40x20x0, the others CC r=4
100x70x20) run in a batch; sampled results equal the same request alone,
bit for bit.  C5: a Poisson stream of 128^3 quick and 192x224x192 full
requests against the resident templates, each with a 32-label Voronoi map;
sampled results equal the request alone, and the returned labels equal the
oracle's nearest-neighbour warp of the moving labels by the returned field."""
import numpy as np
import pytest

import oracle as o
from ants_amd import phantoms as ph
from ants_amd.registration import SyNStage
from ants_amd.serving import Batch, Request, run_all, serve

pytestmark = pytest.mark.gpu

QUICK = SyNStage.from_args("SyN[0.1,3,0]", "CC[f,m,1,2]", "[40x20x0,1e-6,10]", "4x2x1", "2x1x0vox")
FULL = SyNStage.from_args("SyN[0.1,3,0]", "CC[f,m,1,4]", "[100x70x20,1e-6,10]", "4x2x1", "2x1x0vox")


def c4_specs(n=64, seed=0):
    rng = np.random.default_rng(seed)   # as bench.py's batch leg draws them
    return [(1000 + k, tuple(int(v) for v in rng.integers(96, 193, 3))) for k in range(n)]


def alone(req):
    b = Batch(slots=1)
    r = run_all(b, [req])[0]
    b.close()
    return r


@pytest.mark.timeout(600)
def test_c4_batch_sample_equals_alone():
    specs = c4_specs()
    picks = [0, 1, 16, 17, 18, 19]       # two quick pairs, four full-schedule pairs
    reqs = []
    for k in picks:
        seed, size = specs[k]
        p = ph.phantom_pair_3d(seed, size=size)
        reqs.append(Request(QUICK if k < 16 else FULL, p.moving, fixed=p.fixed, want_fields=True, tag=str(size)))
    b = Batch(slots=4)
    res = run_all(b, reqs)
    b.close()
    assert all(r.status == 0 for r in res), [r.error for r in res]
    spans = sorted((r.t_start, r.t_end) for r in res)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))  # they overlapped
    for r in (res[1], res[3]):
        ref = alone(r.request)
        print(f"C4 pair {r.request.tag}: {r.voxel_iterations:.3g} voxel-iterations, same bits alone", flush=True)
        np.testing.assert_array_equal(r.fwd, ref.fwd)
        np.testing.assert_array_equal(r.inv, ref.inv)


@pytest.mark.timeout(600)
def test_c5_served_sample_equals_alone_and_labels_match_oracle():
    big_t = ph.phantom_pair_3d(5000, size=(192, 224, 192))
    small_t = ph.phantom_pair_3d(5001, size=(128, 128, 128))
    kinds = [True, True, False, True]     # 128^3 quick, quick, 192x224x192 full, quick
    reqs = []
    for k, quick in enumerate(kinds):
        size = (128, 128, 128) if quick else (192, 224, 192)
        p = ph.phantom_pair_3d(6000 + k, size=size)
        lab = ph.voronoi_labels(6001 + k, p.labels_moving > 0, 32)
        reqs.append(Request(QUICK if quick else FULL, p.moving, fixed=None, template_index=1 if quick else 0,
                            moving_labels=lab, want_fields=True))
    b = Batch(slots=4)
    b.set_template(0, big_t.fixed)
    b.set_template(1, small_t.fixed)
    res = serve(b, reqs, [0.0, 0.02, 0.03, 0.05])
    assert all(r.status == 0 for r in res), [r.error for r in res]
    solo = Batch(slots=1)
    solo.set_template(0, big_t.fixed)
    solo.set_template(1, small_t.fixed)
    for r in (res[0], res[2]):
        ref = run_all(solo, [r.request])[0]
        np.testing.assert_array_equal(r.fwd, ref.fwd)
        np.testing.assert_array_equal(r.labels, ref.labels)
        size = r.request.moving.shape[::-1]
        g = o.Grid(list(size))
        np.testing.assert_array_equal(r.labels, o.warp_labels_nn(g, r.fwd, g, r.request.moving_labels))
        print(f"C5 request {size}: latency {r.latency * 1e3:.1f} ms, same bits alone, labels = oracle NN warp",
              flush=True)
    solo.close()
    b.close()
