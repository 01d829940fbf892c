"""Writes tests/golden/reference_fixtures.json.

Inputs are the data tables of the reference's own NUFFT tests (read as text from
lib/tests/tests1.cpp: test_deterministic_1d 16-56, test_deterministic_3d 58-216,
compare_normal_methods 351-550), built into float32 coordinates exactly as those tests
do (-pi + 2 pi u, stored in a float tensor). The reference tests print their outputs
and assert nothing, so the expected outputs here are the FP64 direct sums of the
definitions (oracle.nudft_*), with each case's own signs and the reference's
1/sqrt(prod nmodes) scaling.

Run from the repo root: python tests/golden/make_golden.py [/path/to/reference]
"""
import json
import math
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def tables(src: str):
    """The `fixed_coords = { {a, b, c}, ... };` tables, in file order."""
    out = []
    for m in re.finditer(r"fixed_coords\s*=\s*\{(.*?)\};", src, flags=re.S):
        rows = re.findall(r"\{\s*([-0-9.eE+]+)\s*,\s*([-0-9.eE+]+)\s*,\s*([-0-9.eE+]+)\s*\}", m.group(1))
        out.append([[float(a), float(b), float(c)] for a, b, c in rows])
    return out


def f32_coords(rows, n):
    u = np.asarray(rows[:n], dtype=np.float32).astype(np.float64)
    return (-math.pi + 2.0 * math.pi * u).T.astype(np.float32)  # [3, n]


def cplx(a):
    a = np.asarray(a)
    return {"re": a.real.ravel().tolist(), "im": a.imag.ravel().tolist(), "shape": list(a.shape)}


def main(ref_root):
    src = open(os.path.join(ref_root, "lib", "tests", "tests1.cpp")).read()
    t3d, tcmp = tables(src)[:2]
    fx = {"source": "lib/tests/tests1.cpp (input tables); expected outputs = FP64 direct sums"}

    # test_deterministic_1d: 8 Cartesian points -pi + 2 pi x / 8, type 2, positive = true, ones.
    nx = 8
    c1 = np.asarray([-math.pi + 2 * x * math.pi / nx for x in range(nx)], dtype=np.float32)[None, :]
    f1 = np.ones((1, nx), dtype=np.complex128)
    out1 = O.nudft_type2(c1.astype(np.float64), f1[0], isign=+1) / math.sqrt(nx)
    fx["deterministic_1d"] = {"coords": c1.tolist(), "nmodes": [1, nx], "positive": True, "input": cplx(f1),
                              "type2": cplx(out1[None])}

    # test_deterministic_3d: n1=n2=n3=2 -> nf = 8 points, nmodes {1, 4, 4, 4};
    # c[0,0,0,1] = 2, c[0,2,1,3] = 3; type 2 (negative) then type 1 (positive), tol 1e-6.
    co = f32_coords(t3d, 8)
    nm = [1, 4, 4, 4]
    f = np.zeros(nm, dtype=np.complex128)
    f[0, 0, 0, 1] = 2.0
    f[0, 2, 1, 3] = 3.0
    k2 = O.nudft_type2(co.astype(np.float64), f[0], isign=-1) / 8.0
    back = O.nudft_type1(co.astype(np.float64), k2, nm[1:], isign=+1) / 8.0
    fx["deterministic_3d"] = {"coords": co.tolist(), "nmodes": nm, "input": cplx(f), "type2": cplx(k2[None]),
                              "type1_of_type2": cplx(back[None])}

    # compare_normal_methods: nx=5, ny=2, nz=3 -> 30 points, nmodes {1, 3, 2, 5}, c = ones.
    co = f32_coords(tcmp, 30)
    nm = [1, 3, 2, 5]
    nvox = 30.0
    ones = np.ones(nm, dtype=np.complex128)
    cd = co.astype(np.float64)
    # out1: Nufft type 2 with positive=true, then Nufft type 1 with positive=false (as the test builds them)
    k = O.nudft_type2(cd, ones[0], isign=+1) / math.sqrt(nvox)
    out1 = O.nudft_type1(cd, k, nm[1:], isign=-1) / math.sqrt(nvox)
    # out2: CUDANufftNormal with the default signs
    k = O.nudft_type2(cd, ones[0], isign=-1) / math.sqrt(nvox)
    out2 = O.nudft_type1(cd, k, nm[1:], isign=+1) / math.sqrt(nvox)
    # out3: CUDANormalNufftToeplitz::apply = 2^-3 * out2 in exact arithmetic
    diag = O.toeplitz_diagonal(cd, nm, exact=True)
    out3 = O.toeplitz_apply(diag, ones)
    fx["compare_normal_methods"] = {"coords": co.tolist(), "nmodes": nm, "explicit_flipped_signs": cplx(out1[None]),
                                    "normal": cplx(out2[None]), "toeplitz": cplx(out3),
                                    "toeplitz_diagonal": cplx(diag)}

    with open(os.path.join(HERE, "reference_fixtures.json"), "w") as fh:
        json.dump(fx, fh)
    print("wrote", os.path.join(HERE, "reference_fixtures.json"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
