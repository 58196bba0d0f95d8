"""FP32 and BF16-pyramid GPU forms against the FP64 oracle, per level
(CC energy, max/mean field error in voxels, Jacobian-determinant error).
The BF16 form stores the smoothed images in BF16 (relative step 2^-8); its
bar is looser than the FP32 form's and is written here."""
import numpy as np
import pytest

import oracle as o
from ants_amd import phantoms as ph
from ants_amd.analysis import field_errors
from ants_amd.registration import SyNRegistration, SyNStage

pytestmark = pytest.mark.gpu

# measured on this 48x56x48 case: energy 5.5e-5 relative, fields 0.011 voxel
BF16_ENERGY_RTOL = 2e-4
BF16_FIELD_MAX_VOX = 0.04


def per_level(precision, pair, iters):
    out = []
    for lv in range(len(iters)):
        it = [n if k <= lv else 0 for k, n in enumerate(iters)]
        st = SyNStage.from_args("SyN[0.1,3,0]", "CC[f,m,1,4]", "[%s,1e-6,10]" % "x".join(map(str, it)), "4x2x1",
                                "2x1x0vox")
        if precision == "oracle":
            P = o.SynParams(shrink=[4, 2, 1], sigma=[2, 1, 0], iterations=it, learning_rate=0.1, radius=4,
                            force_iterations=True)
            g = o.Grid(list(pair.size))
            r = o.syn_register(P, g, pair.fixed, g, pair.moving)
            out.append((r.energy[r.level == lv], r.fwd))
        else:
            reg = SyNRegistration(st, 3, precision=precision, force_iterations=True)
            r = reg.register(pair.fixed, pair.moving)
            reg.close()
            out.append((r.energy[r.levels == lv], r.fwd))
    return out


def test_precision_forms_per_level():
    pair = ph.phantom_pair_3d(7, size=(48, 56, 48), max_disp_mm=4.0, svf_sigma_mm=6.0)
    iters = [10, 6, 3]
    ref = per_level("oracle", pair, iters)
    rows = []
    for prec in ("fp32", "bf16"):
        res = per_level(prec, pair, iters)
        for lv, ((e, f), (er, fr)) in enumerate(zip(res, ref)):
            erel = float(np.max(np.abs(e - er) / np.abs(er)))
            fe = field_errors(f, fr, pair.spacing)
            rows.append((prec, lv, erel, fe["max_vox"], fe["mean_vox"], fe["jac_max_abs"]))
            if prec == "fp32":
                assert erel < 1e-4 and fe["max_vox"] < 0.05
            else:
                assert erel < BF16_ENERGY_RTOL and fe["max_vox"] < BF16_FIELD_MAX_VOX
    print("\nprecision  level  energy_rel  field_max_vox  field_mean_vox  jacdet_max_abs")
    for r in rows:
        print("%-9s  %5d  %10.2e  %13.4f  %14.5f  %14.5f" % r)
