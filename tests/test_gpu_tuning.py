"""Launch shapes from the tuning table compute the same registration: the
smoothing tile and z-chunk give the same bits; the CC tile and z-chunk and
the inversion block count change only the order of double-precision sums
(energy within 1e-7 relative, fields within 1e-3 voxel: rounding differences
that the iterations carry forward, far inside the 1e-4 / 0.05 voxel parity bars)."""
import ctypes as C

import numpy as np
import pytest

from ants_amd import _lib as L
from ants_amd import phantoms as ph
from ants_amd.registration import SyNRegistration, SyNStage

pytestmark = pytest.mark.gpu


def run(pair, stage):
    reg = SyNRegistration(stage, 3, force_iterations=True)
    r = reg.register(pair.fixed, pair.moving)
    reg.close()
    return r


def with_table(entries, fn):
    lib = L.lib()
    v = C.c_int()
    saved = {}
    for k in range(5):
        for b in range(40):
            L.check(lib.ants_tuning_get(k, b, C.byref(v)))
            saved[(k, b)] = v.value
    try:
        L.check(lib.ants_tuning_clear())
        for k, val in entries:
            for b in range(40):
                L.check(lib.ants_tuning_set(k, b, val))
        return fn()
    finally:
        L.check(lib.ants_tuning_clear())
        for (k, b), val in saved.items():
            if val:
                L.check(lib.ants_tuning_set(k, b, val))


@pytest.mark.parametrize("r", [4, 2])
def test_launch_shapes_compute_the_same_registration(r):
    pair = ph.phantom_pair_3d(17, size=(72, 80, 64), max_disp_mm=4.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args("SyN[0.1,3,0]", f"CC[f,m,1,{r}]", "[8x5x3,1e-6,10]", "4x2x1", "2x1x0vox")
    ref = with_table([], lambda: run(pair, stage))
    smooth = with_table([(4, 2), (2, 5)], lambda: run(pair, stage))
    np.testing.assert_array_equal(smooth.fwd, ref.fwd)
    np.testing.assert_array_equal(smooth.energy, ref.energy)
    for entries in ([(3, 2)], [(1, 5)], [(3, 2), (1, 9)], [(0, 100)], [(0, 1024)]):
        got = with_table(entries, lambda: run(pair, stage))
        rel = np.abs(got.energy - ref.energy) / np.abs(ref.energy)
        dv = max(np.abs(got.fwd - ref.fwd).max(), np.abs(got.inv - ref.inv).max())
        print(f"table {entries}: energy rel {rel.max():.1e}, fields {dv:.1e} mm", flush=True)
        assert rel.max() < 1e-7 and dv < 1e-3
