"""Writes gaussian_kernels.json: half kernels of ITK's GaussianOperator for
the variances and maximum errors the SyN stage uses (image pyramid: error
0.01; fields: error 0.001), from the published definition
c_n = e^-t I_n(t) (scipy.special.ive) normalised to sum 1 over -R..R, with
the radius R that GaussianOperator's growth rule gives for each case.

Run from the repository root:  python tests/golden/make_gaussian_kernels.py
"""
import json
import os

import numpy as np
from scipy import special

# (variance t, maximum error): the image pyramid's and the field smoothing's
# error bounds at the variances the stage's schedules use
CASES = [(3.0, 0.001), (4.0, 0.01), (1.0, 0.01), (0.5, 0.001), (1.0, 0.001), (0.25, 0.01)]


def radius_for(t, max_error):
    """GaussianOperator's growth rule: add taps while the kernel's mass
    1 + 2 sum(c_1..c_R) (in units of c_0 normalised by the e^-t I_n(t)
    series, whose total is 1) is below 1 - max_error."""
    mass = special.ive(0, t)
    r = 0
    while mass < 1.0 - max_error:
        r += 1
        mass += 2.0 * special.ive(r, t)
    return r


def half_kernel(t, radius):
    c = special.ive(np.arange(radius + 1), t)
    return c / (c[0] + 2.0 * c[1:].sum())


def main():
    out = []
    for t, e in CASES:
        r = radius_for(t, e)
        out.append({"variance": t, "max_error": e, "radius": r, "half_kernel": [float(v) for v in half_kernel(t, r)]})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gaussian_kernels.json")
    with open(path, "w") as fh:
        json.dump({"source": "scipy.special.ive, normalised; radius from GaussianOperator's rule", "kernels": out},
                  fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
