"""Command lines over the library, reading and writing files as ANTs does.

``python -m ants_amd.cli registration -d 3 -m CC[fixed.nii.gz,moving.nii.gz,1,4]
  -t SyN[0.1,3,0] -c [100x70x20,1e-6,10] -f 4x2x1 -s 2x1x0vox
  -o [out_,out_Warped.nii.gz,out_InverseWarped.nii.gz]``
runs one SyN stage with the CC metric (the antsRegistration subset this
library replaces) and writes ``out_0Warp.nii.gz`` and
``out_0InverseWarp.nii.gz`` (naming of Examples/
antsRegistrationTemplateHeader.cxx:74-89) plus the optional warped images.
``-r affine.mat`` (repeatable, ``[file,1]`` inverts; ``[fixed,moving,f]``
computes one from the images: f = 0 image centres, 1 centres of mass, 2
origins) gives the stage its
moving initial transform, as the preceding linear stages do; the outputs are
then ``out_0GenericAffine.mat`` (the initial transforms composed),
``out_1Warp.nii.gz`` and ``out_1InverseWarp.nii.gz``.

Linear stages come first, each with its own option group, as
antsRegistrationSyN.sh writes them (Scripts/antsRegistrationSyN.sh:487-497):
``-t Rigid[0.1] -m Mattes[fixed,moving,1,32,Regular,0.25] -c [1000x500x250x100,1e-6,10]
  -f 8x4x2x1 -s 3x2x1x0vox -t Affine[0.1] -m ... -t SyN[0.1,3,0] -m CC[...] ...``
(Translation, Rigid or Affine with MeanSquares, GC or Mattes;
``--random-seed`` seeds their metric sampling).  Their transforms, composed
with the -r initial transforms, become ``out_0GenericAffine.mat``; with no
SyN stage that file is the whole output.  The last stage may instead be a
dense field-transform stage, ``-t GaussianDisplacementField[0.25,3,0]``,
``-t Exponential[0.25,3,0.5,0]``, ``-t BSplineDisplacementField[0.25,26,0,3]``
or ``-t BSplineExponential[0.25,26,0,0,3]`` with ``-m CC[...]``: it writes
``out_{k}Warp.nii.gz`` (the exponential forms also ``out_{k}InverseWarp.nii.gz``).

``python -m ants_amd.cli apply -d 3 -i moving.nii.gz -r fixed.nii.gz
  -o out.nii.gz -t out_0Warp.nii.gz -t [affine.mat,1] -n Linear
  --default-value 0``
is the antsApplyTransforms subset (Examples/antsApplyTransforms.cxx).
Transforms ending in .mat/.txt/.tfm are affine files, .h5 files ITK
composite transforms (their components in the order they apply); other paths
are displacement-field images.  ``-o [field.nii.gz,1]`` (or
``DisplacementField[field.nii.gz]``; ``.h5`` writes it as a transform) writes
the composite as one displacement field on the reference grid,
``-o Linear[affine.mat(,1)]`` the affine list collapsed into one (inverted
with 1) and ``-o CompositeTransform[out.h5]`` the list as a composite file;
those need no input image.  ``-n BSpline[3]`` selects B-spline
interpolation of order 0..5.

``apply -e 3`` resamples every volume of a 4-D time series (the output keeps
the input's time spacing and offset).

``python -m ants_amd.cli apply-points -d 3 -i points.csv -o out.csv
  -t [affine.mat,1] -t out_0InverseWarp.nii.gz`` is antsApplyTransformsToPoints
(Examples/antsApplyTransformsToPoints.cxx): the x,y[,z] columns of the CSV
(physical LPS coordinates, a header row) through the transforms in -t order,
the other columns copied.

``python -m ants_amd.cli jacobian 3 out_0Warp.nii.gz jac.nii.gz [doLog]
  [useGeometric] [deformationGradient]`` is CreateJacobianDeterminantImage;
``similarity -d 3 -m CC[fixed.nii.gz,moving.nii.gz,1,4]`` prints
MeasureImageSimilarity's CC value; ``smooth-field 3 in.nii.gz out.nii.gz 2.0``
(a variance) or ``... 8x8x8`` (a B-spline mesh) is SmoothDisplacementField.

``registration ... -a 1`` writes ``out_Composite.h5`` and
``out_InverseComposite.h5`` (--write-composite-transform) instead of the
separate files; ``--save-state state.h5`` writes the registration state
(the linear initial transforms, then FixedToMiddle and MovingToMiddle with
their inverses) and ``--restore-state state.h5`` starts the stage from one
(Examples/antsRegistrationTemplateHeader.h:293-327, 1308-1348).
"""
from __future__ import annotations

import argparse
import re
import sys
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import io as aio
from .apply_transforms import AffineTransform, DisplacementFieldTransform, Geometry, apply_transforms, compose_affines
from .fieldmodels import FieldStage, register_field
from .fieldmodels import level_geometry as field_level_geometry


def _bracket_list(s: str) -> List[str]:
    m = re.fullmatch(r"\s*\[(.*)\]\s*", s)
    return [p.strip() for p in (m.group(1) if m else s).split(",")]


def load_transforms(spec: str) -> list:
    """-t path | -t [path,invert]: the transforms of one option in the order
    they apply (an .h5 composite lists its components newest-first, the
    order ITK's CompositeTransform applies them)."""
    parts = _bracket_list(spec)
    path = parts[0]
    if len(parts) == 3 and parts[2] in ("0", "1", "2"):
        # [fixed,moving,feature]: an initial transform computed from the images
        from .registration import initial_transform_from_images
        fi, gfi = aio.read_nifti(parts[0])
        mi, gmi = aio.read_nifti(parts[1])
        return [initial_transform_from_images(np.asarray(fi, np.float32), gfi, np.asarray(mi, np.float32), gmi,
                                              int(parts[2]))]
    if path.endswith(".h5"):
        from . import h5transform
        if len(parts) > 1 and parts[1] not in ("0", "false", "False", ""):
            raise ValueError(f"{path}: inverting a composite file is not supported")
        return list(reversed(h5transform.read_composite(path)))
    t = load_transform(spec)
    return [] if t is None else [t]


def load_transform(spec: str):
    """-t path | -t [path,invert]"""
    parts = _bracket_list(spec)
    path = parts[0]
    invert = len(parts) > 1 and parts[1] not in ("0", "false", "False", "")
    if path.lower() == "identity":
        return None
    if path.endswith(".h5"):
        ts = load_transforms(spec)
        if len(ts) != 1:
            raise ValueError(f"{path}: a composite of {len(ts)} transforms where one is expected")
        return ts[0]
    if path.endswith((".mat", ".txt", ".tfm")):
        t = aio.read_transform(path)
        t.invert = invert
        return t
    if invert:
        raise ValueError(f"{path}: a displacement field cannot be inverted here (use its InverseWarp file)")
    fld, g = aio.read_nifti(path)
    return DisplacementFieldTransform(np.asarray(fld, np.float32), g)


def parse_apply_output(spec: str) -> Tuple[str, str, bool]:
    """-o forms of antsApplyTransforms (Examples/antsApplyTransforms.cxx:643-770):
    ("image", path, _) for `out` / `[out,0]`; ("field", path, _) for
    `[out,1]` / `DisplacementField[out]`; ("linear", path, invert) for
    `Linear[out(,1)]`; ("composite", path, _) for `CompositeTransform[out]`."""
    m = re.fullmatch(r"\s*([A-Za-z]*)\s*\[(.*)\]\s*", spec)
    if not m:
        return "image", spec.strip(), False
    name, parts = m.group(1).lower(), [p.strip() for p in m.group(2).split(",")]
    flag = len(parts) > 1 and parts[1] not in ("", "0")
    if name == "linear":
        return "linear", parts[0], flag
    if name == "compositetransform":
        return "composite", parts[0], False
    if name == "displacementfield" or (name == "" and flag):
        return "field", parts[0], False
    if name == "":
        return "image", parts[0], False
    raise ValueError(f"unknown output form {spec!r}")


def cmd_apply(argv: Sequence[str]) -> int:
    from .apply_transforms import collapse_linear, compose_transforms_to_field
    ap = argparse.ArgumentParser(prog="apply")
    ap.add_argument("-d", "--dimensionality", type=int, default=None)
    ap.add_argument("-i", "--input", default=None)
    ap.add_argument("-r", "--reference-image", default=None)
    ap.add_argument("-o", "--output", required=True)
    ap.add_argument("-t", "--transform", action="append", default=[])
    ap.add_argument("-n", "--interpolation", default="Linear")
    ap.add_argument("--default-value", type=float, default=0.0)
    ap.add_argument("-e", "--input-image-type", type=int, default=0,
                    help="0 scalar, 3 time series (a 4-D image, each volume resampled)")
    ap.add_argument("-v", "--verbose", type=int, default=0)
    a = ap.parse_args(argv)
    if a.input_image_type not in (0, 3):
        raise SystemExit("-e: scalar (0) and time-series (3) inputs are supported")
    kind, path, invert = parse_apply_output(a.output)
    xf = [x for t in a.transform for x in load_transforms(t)]
    if kind == "linear":
        D = a.dimensionality or (np.asarray(xf[0].matrix).shape[0] if xf and xf[0] is not None else 3)
        t = collapse_linear(xf, D)
        A, tr, c = t.parameters()
        if invert:
            t = AffineTransform(A, tr, c, invert=True)
            A, tr, c = t.parameters()
        aio.write_transform(path, AffineTransform(A, tr, c))
        return 0
    if kind == "composite":
        from .h5transform import write_composite
        # queue order: the transform applied first is added last
        write_composite(path, [t for t in reversed(xf) if t is not None])
        return 0
    if not a.reference_image:
        raise SystemExit("-r is required")
    ref, gr = aio.read_nifti(a.reference_image)
    if kind == "field":
        fld = compose_transforms_to_field(gr, xf)
        if path.endswith((".h5", ".hdf5")):
            from .h5transform import write_composite
            write_composite(path, [DisplacementFieldTransform(fld, gr)])
        else:
            aio.write_nifti(path, fld, gr, vector=True)
        return 0
    if not a.input:
        raise SystemExit("Output is not a transform, and there is no input image")
    if a.input_image_type == 3:
        from .apply_transforms import apply_transforms_series
        series, gs, t_sp, t_org = aio.read_nifti_series(a.input)
        out = apply_transforms_series(gr, series, gs, xf, a.interpolation, a.default_value)
        aio.write_nifti_series(path, out, gr, t_sp, t_org)
        return 0
    img, gi = aio.read_nifti(a.input)
    if a.dimensionality and a.dimensionality != gi.dim:
        raise SystemExit(f"-d {a.dimensionality} but the input is {gi.dim}-D")
    out = apply_transforms(gr, np.asarray(img, np.float32), gi, xf, a.interpolation, a.default_value)
    if a.interpolation.lower().startswith(("nearest", "genericlabel")) and np.issubdtype(np.asarray(img).dtype,
                                                                                         np.integer):
        out = out.astype(np.asarray(img).dtype)
    aio.write_nifti(path, out, gr)
    return 0


def cmd_registration(argv: Sequence[str]) -> int:
    from .linear import LinearStage, register_linear
    from .registration import SyNRegistration, SyNStage
    ap = argparse.ArgumentParser(prog="registration")
    ap.add_argument("-d", "--dimensionality", type=int, required=True)
    # one -t/-m/-c/-f/-s group per stage, in the order the stages run
    ap.add_argument("-m", "--metric", action="append", required=True)
    ap.add_argument("-t", "--transform", action="append", required=True)
    ap.add_argument("-c", "--convergence", action="append", required=True)
    ap.add_argument("-f", "--shrink-factors", action="append", required=True)
    ap.add_argument("-s", "--smoothing-sigmas", action="append", required=True)
    ap.add_argument("-o", "--output", required=True)
    ap.add_argument("-r", "--initial-moving-transform", action="append", default=[],
                    help="affine transform file or [file,useInverse]; repeated options compose in -t order")
    ap.add_argument("-q", "--initial-fixed-transform", action="append", default=[],
                    help="affine transform file or [file,useInverse] applied to the fixed image's points")
    ap.add_argument("--float", default="1")
    ap.add_argument("--precision", choices=["fp32", "bf16"], default="fp32")
    ap.add_argument("--random-seed", type=int, default=0, help="seed of the linear stages' metric sampling")
    ap.add_argument("-a", "--write-composite-transform", type=int, default=0,
                    help="1: write <prefix>Composite.h5 and <prefix>InverseComposite.h5")
    ap.add_argument("--save-state", default=None, help="write the registration state (.h5)")
    ap.add_argument("--restore-state", default=None, help="start from a saved registration state (.h5)")
    ap.add_argument("--print-similarity-measure-interval", type=int, default=0,
                    help="full-scale CC r=4 of the original images every N iterations (0: off)")
    ap.add_argument("-v", "--verbose", type=int, default=0)
    a = ap.parse_args(argv)
    n = len(a.transform)
    if not (len(a.metric) == len(a.convergence) == len(a.shrink_factors) == len(a.smoothing_sigmas) == n):
        raise SystemExit("each stage needs one -t, -m, -c, -f and -s (in stage order)")
    linear = [LinearStage.is_linear(t) for t in a.transform]
    if any(not lin for lin in linear[:-1]):
        raise SystemExit("a SyN / BSplineSyN / GaussianDisplacementField / Exponential stage must be the last stage")
    images = {}

    def image(name):
        if name not in images:
            images[name] = aio.read_nifti(name)
        return images[name]

    mparts = _bracket_list(re.sub(r"^\s*[A-Za-z]+", "", a.metric[-1]))
    fixed, gf = image(mparts[0])
    moving, gm = image(mparts[1])
    for g, name in ((gf, mparts[0]), (gm, mparts[1])):
        if g.dim != a.dimensionality:
            raise SystemExit(f"{name} is {g.dim}-D, -d {a.dimensionality}")
    outs = _bracket_list(a.output)
    prefix = outs[0]

    def initial(specs):
        return [t for spec in specs for t in load_transforms(spec)]

    init = initial(a.initial_moving_transform)
    finit = initial(a.initial_fixed_transform)
    restored = None
    if a.restore_state:
        if init or finit:
            # antsRegistrationTemplateHeader.h:295-305
            raise SystemExit("restore-state option is mutually exclusive with initial-moving-transform & "
                             "initial-fixed-transform options")
        if any(linear):
            raise SystemExit("restore-state continues a SyN stage; linear stages cannot follow it here")
        from . import h5transform
        lin, state, gst = h5transform.read_syn_state(a.restore_state)
        init = list(reversed(lin))
        restored = (state, gst)
    # linear stages: each stage's transform is applied before the transforms
    # gathered so far (the composite applies the newest transform first)
    for k in range(n):
        if not linear[k]:
            continue
        st = LinearStage.from_args(a.transform[k], a.metric[k], a.convergence[k], a.shrink_factors[k],
                                   a.smoothing_sigmas[k], a.random_seed)
        fk, gfk = image(st.fixed)
        mk, gmk = image(st.moving)
        if any(isinstance(t, DisplacementFieldTransform) for t in init + finit):
            raise SystemExit("linear stages take affine initial transforms only")
        res = register_linear(st, np.asarray(fk, np.float32), gfk, np.asarray(mk, np.float32), gmk, init, finit)
        if a.verbose:
            print(f"*** linear stage {k}: {a.transform[k]} {a.metric[k]} ***")
            for ln in res.diagnostic_lines():
                print(ln)
        init = [res.transform] + init
    if all(linear):
        if finit and (len(outs) > 1 and any(outs[1:])):
            raise SystemExit("-q with warped-image outputs is not supported (write the transform only)")
        if a.save_state or a.restore_state or a.print_similarity_measure_interval:
            raise SystemExit("--save-state / --restore-state / the full-scale CC need a SyN stage")
        aff = compose_affines(init) if init else AffineTransform(np.eye(gf.dim), np.zeros(gf.dim), np.zeros(gf.dim))
        if a.write_composite_transform:
            from . import h5transform
            h5transform.write_composite(prefix + "Composite.h5", [aff])
            h5transform.write_composite(prefix + "InverseComposite.h5",
                                        [AffineTransform(aff.matrix, aff.translation, aff.center, invert=True)])
        else:
            aio.write_transform(prefix + "0GenericAffine.mat", aff)
        if len(outs) > 1 and outs[1]:
            aio.write_nifti(outs[1], apply_transforms(gf, np.asarray(moving, np.float32), gm, [aff], "Linear"), gf)
        if len(outs) > 2 and outs[2]:
            inv = AffineTransform(aff.matrix, aff.translation, aff.center, invert=True)
            aio.write_nifti(outs[2], apply_transforms(gm, np.asarray(fixed, np.float32), gf, [inv], "Linear"), gm)
        return 0
    if FieldStage.is_field(a.transform[-1]):
        return _field_stage_outputs(a, fixed, gf, moving, gm, init, finit, outs, prefix)
    # a BSplineSyN knot spacing becomes a mesh with the fixed image's geometry
    # (antsRegistrationTemplateHeader.h:709-731)
    stage = SyNStage.from_args(a.transform[-1], a.metric[-1], a.convergence[-1], a.shrink_factors[-1],
                               a.smoothing_sigmas[-1], gf.size, gf.spacing)
    if finit and (len(outs) > 1 and any(outs[1:])):
        raise SystemExit("-q with warped-image outputs is not supported (write the warp files only)")
    if any(isinstance(t, DisplacementFieldTransform) for t in init) and not a.write_composite_transform:
        raise SystemExit("a displacement field among the initial transforms needs -a 1 (composite outputs)")
    reg = SyNRegistration(stage, a.dimensionality, precision=a.precision)
    reg.set_initial_transforms(init, finit)
    if restored is not None:
        state, gst = restored
        reg.set_initial_state(state, gst.size, gst.spacing, gst.origin, gst.direction)
    reg.set_full_scale_cc_interval(a.print_similarity_measure_interval)
    res = reg.register(np.asarray(fixed, np.float32), np.asarray(moving, np.float32), gf.spacing, gf.origin,
                       gm.spacing, gm.origin, gf.direction, gm.direction)
    lg = reg.level_geometry(gf.size, gf.spacing, gf.origin, direction=gf.direction)  # the fields' grid
    lshape = tuple(int(lg.size[k]) for k in reversed(range(gf.dim))) + (gf.dim,)
    state = reg.state(lshape) if a.save_state else None
    reg.close()
    if a.verbose:
        for ln in res.diagnostic_lines():
            print(ln)
    D = gf.dim
    g_out = Geometry(tuple(int(lg.size[k]) for k in range(D)), tuple(lg.spacing[k] for k in range(D)),
                     tuple(lg.origin[k] for k in range(D)), np.array([lg.direction[k] for k in range(D * D)]).reshape(D, D))
    if state is not None:
        from . import h5transform
        # the state composite lists the linear transforms first (applied last)
        h5transform.write_syn_state(a.save_state, state, g_out, [t for t in reversed(init)
                                                                 if not isinstance(t, DisplacementFieldTransform)])
    # with initial transforms the output composite is [initial affine, SyN field]:
    # <prefix>0GenericAffine.mat, <prefix>1Warp.nii.gz, <prefix>1InverseWarp.nii.gz
    k = 0
    fwd_chain = [DisplacementFieldTransform(res.fwd, g_out)]
    inv_chain = [DisplacementFieldTransform(res.inv, g_out)]
    if a.write_composite_transform:
        # --write-composite-transform: the stage's composite and its inverse,
        # components newest-first
        from . import h5transform
        fwd_chain = fwd_chain + list(init)
        inv_init = []
        for t in reversed(init):
            if isinstance(t, DisplacementFieldTransform):
                raise SystemExit("the inverse composite of a displacement-field initial transform is not available")
            inv_init.append(AffineTransform(t.matrix, t.translation, t.center, invert=not t.invert))
        inv_chain = inv_init + inv_chain
        h5transform.write_composite(prefix + "Composite.h5", list(reversed(fwd_chain)))
        h5transform.write_composite(prefix + "InverseComposite.h5", list(reversed(inv_chain)))
    elif init:
        aff = compose_affines(init)
        aio.write_transform(prefix + "0GenericAffine.mat", aff)
        k = 1
        fwd_chain = fwd_chain + [aff]
        inv_chain = [AffineTransform(aff.matrix, aff.translation, aff.center, invert=True)] + inv_chain
    if not a.write_composite_transform:
        aio.write_nifti(prefix + f"{k}Warp.nii.gz", res.fwd, g_out, vector=True)
        aio.write_nifti(prefix + f"{k}InverseWarp.nii.gz", res.inv, g_out, vector=True)
    if len(outs) > 1 and outs[1]:
        warped = apply_transforms(gf, np.asarray(moving, np.float32), gm, fwd_chain, "Linear")
        aio.write_nifti(outs[1], warped, gf)
    if len(outs) > 2 and outs[2]:
        inv_warped = apply_transforms(gm, np.asarray(fixed, np.float32), gf, inv_chain, "Linear")
        aio.write_nifti(outs[2], inv_warped, gm)
    return 0


def _field_stage_outputs(a, fixed, gf, moving, gm, init, finit, outs, prefix) -> int:
    """A dense field-transform last stage: its field (and for the exponential
    forms the inverse) written as <prefix>{k}Warp.nii.gz /
    <prefix>{k}InverseWarp.nii.gz after 0GenericAffine.mat when initial
    transforms precede it, as antsRegistration names a displacement-field
    transform; the displacement-field forms keep no inverse field."""
    st = FieldStage.from_args(a.transform[-1], a.metric[-1], a.convergence[-1], a.shrink_factors[-1],
                              a.smoothing_sigmas[-1], gf.size, gf.spacing)
    if any(isinstance(t, DisplacementFieldTransform) for t in init + finit):
        raise SystemExit("field-transform stages take affine initial transforms only")
    if a.save_state or a.restore_state or a.print_similarity_measure_interval or a.write_composite_transform:
        raise SystemExit("--save-state / --restore-state / the full-scale CC / -a 1 need a SyN stage here")
    if finit and (len(outs) > 1 and any(outs[1:])):
        raise SystemExit("-q with warped-image outputs is not supported (write the warp files only)")
    res = register_field(st, np.asarray(fixed, np.float32), gf, np.asarray(moving, np.float32), gm, init, finit)
    if a.verbose:
        for (lev, it, v, cv, lr) in res.records:
            print(f" 2DIAGNOSTIC, {it:5d}, {v:.12e}, {cv:.12e}, ")
    g_out = field_level_geometry(st, gf, len(st.shrink) - 1)
    k = 0
    fwd_chain = [DisplacementFieldTransform(res.field, g_out)]
    inv_chain = [DisplacementFieldTransform(res.inverse, g_out)] if res.inverse is not None else None
    if init:
        aff = compose_affines(init)
        aio.write_transform(prefix + "0GenericAffine.mat", aff)
        k = 1
        fwd_chain = fwd_chain + [aff]
        if inv_chain is not None:
            inv_chain = [AffineTransform(aff.matrix, aff.translation, aff.center, invert=True)] + inv_chain
    aio.write_nifti(prefix + f"{k}Warp.nii.gz", res.field, g_out, vector=True)
    if res.inverse is not None:
        aio.write_nifti(prefix + f"{k}InverseWarp.nii.gz", res.inverse, g_out, vector=True)
    if len(outs) > 1 and outs[1]:
        aio.write_nifti(outs[1], apply_transforms(gf, np.asarray(moving, np.float32), gm, fwd_chain, "Linear"), gf)
    if len(outs) > 2 and outs[2]:
        if inv_chain is None:
            raise SystemExit("a displacement-field stage keeps no inverse: no inverse-warped output")
        aio.write_nifti(outs[2], apply_transforms(gm, np.asarray(fixed, np.float32), gf, inv_chain, "Linear"), gm)
    return 0


def cmd_apply_points(argv: Sequence[str]) -> int:
    """antsApplyTransformsToPoints -d D -i in.csv -o out.csv -t T1 -t T2 ..."""
    from .apply_transforms import apply_transforms_to_points
    ap = argparse.ArgumentParser(prog="apply-points")
    ap.add_argument("-d", "--dimensionality", type=int, required=True)
    ap.add_argument("-i", "--input", required=True)
    ap.add_argument("-o", "--output", required=True)
    ap.add_argument("-t", "--transform", action="append", default=[])
    a = ap.parse_args(argv)
    names, pts = aio.read_points_csv(a.input)
    if pts.shape[1] < a.dimensionality:
        raise SystemExit(f"the point file has fewer than {a.dimensionality} columns")
    xf = [x for t in a.transform for x in load_transforms(t)]
    aio.write_points_csv(_bracket_list(a.output)[0], names, apply_transforms_to_points(pts, xf, a.dimensionality))
    return 0


def cmd_jacobian(argv: Sequence[str]) -> int:
    """CreateJacobianDeterminantImage imageDimension deformationField
    outputImage [doLogJacobian=0] [useGeometric=0] [deformationGradient=0]
    (Examples/CreateJacobianDeterminantImage.cxx:215-240)."""
    from .analysis import create_jacobian_determinant_image
    if len(argv) < 3:
        raise SystemExit("usage: jacobian imageDimension deformationField outputImage [doLogJacobian=0] "
                         "[useGeometric=0] [deformationGradient=0]")
    dim = int(argv[0])
    flags = [int(v) for v in argv[3:6]] + [0] * (3 - len(argv[3:6]))
    if flags[1] and flags[2]:
        raise SystemExit("useGeometric has no deformation-gradient output")
    fld, g = aio.read_nifti(argv[1])
    if g.dim != dim:
        raise SystemExit(f"imageDimension {dim} but the field is {g.dim}-D")
    out = create_jacobian_determinant_image(np.asarray(fld, np.float32), g, do_log=bool(flags[0]),
                                            deformation_gradient=bool(flags[2]), geometric=bool(flags[1]))
    if flags[2]:
        # row-major J[0][0], J[0][1], ... as a vector image
        aio.write_nifti(argv[2], out.reshape(out.shape[:-2] + (dim * dim,)), g, vector=True)
    else:
        aio.write_nifti(argv[2], out, g)
    return 0


def cmd_smooth_field(argv: Sequence[str]) -> int:
    """SmoothDisplacementField imageDimension inputField outputField
    variance_or_mesh_size_base_level [numberOfLevels=1] [splineOrder=3]
    [estimateInverse=0] (Examples/SmoothDisplacementField.cxx:185-200):
    one value is a Gaussian variance, dim values (AxBxC) a B-spline mesh."""
    from .analysis import smooth_displacement_field
    if len(argv) < 4:
        raise SystemExit("usage: smooth-field imageDimension inputField outputField variance_or_mesh "
                         "[numberOfLevels=1] [splineOrder=3] [estimateInverse=0]")
    dim = int(argv[0])
    fld, g = aio.read_nifti(argv[1])
    if g.dim != dim:
        raise SystemExit(f"imageDimension {dim} but the field is {g.dim}-D")
    vals = [float(v) for v in argv[3].split("x")]
    levels = int(argv[4]) if len(argv) > 4 else 1
    order = int(argv[5]) if len(argv) > 5 else 3
    if len(argv) > 6 and int(argv[6]) != 0:
        raise SystemExit("estimateInverse is not supported")
    if len(vals) == 1:
        out = smooth_displacement_field(fld, g, variance=vals[0])
    elif len(vals) == dim:
        if levels != 1:
            raise SystemExit("only one B-spline fitting level is supported")
        out = smooth_displacement_field(fld, g, mesh=[int(v) for v in vals], order=order)
    else:
        raise SystemExit(f"give one variance or {dim} mesh sizes")
    aio.write_nifti(argv[2], out, g, vector=True)
    return 0


def cmd_similarity(argv: Sequence[str]) -> int:
    """MeasureImageSimilarity -d D -m CC[fixed,moving,weight,radius]: prints
    the metric value (the CC form; other metrics are not built here)."""
    from .analysis import measure_cc_similarity
    ap = argparse.ArgumentParser(prog="similarity")
    ap.add_argument("-d", "--dimensionality", type=int, required=True)
    ap.add_argument("-m", "--metric", required=True)
    a = ap.parse_args(argv)
    m = re.fullmatch(r"\s*([A-Za-z]+)\s*\[(.*)\]\s*", a.metric)
    if not m or m.group(1).lower() != "cc":
        raise SystemExit("similarity: only CC[fixed,moving,weight,radius] is supported")
    parts = [p.strip() for p in m.group(2).split(",")]
    if len(parts) < 4:
        raise SystemExit("CC needs [fixed,moving,weight,radius]")
    fixed, gf = aio.read_nifti(parts[0])
    moving, gm = aio.read_nifti(parts[1])
    if gf.dim != a.dimensionality:
        raise SystemExit(f"-d {a.dimensionality} but the fixed image is {gf.dim}-D")
    print(repr(measure_cc_similarity(fixed, gf, moving, gm, int(parts[3]))))
    return 0


def main(argv: Optional[Sequence[str]] = None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    cmds = {"registration": cmd_registration, "apply": cmd_apply, "apply-points": cmd_apply_points,
            "jacobian": cmd_jacobian, "similarity": cmd_similarity, "smooth-field": cmd_smooth_field}
    if not argv or argv[0] not in cmds:
        print(__doc__)
        return 2
    return cmds[argv[0]](argv[1:])


if __name__ == "__main__":
    sys.exit(main())
