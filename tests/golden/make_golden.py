"""Write golden vectors from the compiled reference sweeper.

Runs ``oracle/_ref/nm{4,16}/ref_harness`` (the reference's serial KBA sources
compiled in place by ``oracle/Makefile ref``) on small configurations with
perturbed or random input states, and stores inputs and outputs as
``tests/golden/<name>.npz`` (plain arrays, no pickles).

Run in the development container only (the reference tree does not travel):
    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref")

# (name, NM build, argstring, seed, perturb)  perturb: 0 manufactured,
# 1 vi*(1+u/4), 2 vi=u with u uniform in [-1,1)
CASES = [
    ("c1_3x4x2_ne2_na8_nm4_p1", 4,
     "--ncell_x 3 --ncell_y 4 --ncell_z 2 --ne 2 --na 8", 1, 1),
    ("c2_3x4x2_ne2_na8_nm16_p2", 16,
     "--ncell_x 3 --ncell_y 4 --ncell_z 2 --ne 2 --na 8", 2, 2),
    ("c3_5x4x5_ne3_na7_nm4_p1", 4,
     "--ncell_x 5 --ncell_y 4 --ncell_z 5 --ne 3 --na 7", 3, 1),
    ("c4_2x3x4_ne2_na33_nm16_it2_p1", 16,
     "--ncell_x 2 --ncell_y 3 --ncell_z 4 --ne 2 --na 33 --niterations 2", 4, 1),
    ("c5_4x4x4_ne2_na64_nm16_bz2_p1", 16,
     "--ncell_x 4 --ncell_y 4 --ncell_z 4 --ne 2 --na 64 --nblock_z 2", 5, 1),
    ("c6_1x2x1_ne1_na1_nm4_p2", 4,
     "--ncell_x 1 --ncell_y 2 --ncell_z 1 --ne 1 --na 1", 6, 2),
    ("c7_6x5x7_ne1_na32_nm16_p1", 16,
     "--ncell_x 6 --ncell_y 5 --ncell_z 7 --ne 1 --na 32", 7, 1),
    ("c8_9x3x2_ne1_na16_nm4_it3_p2", 4,
     "--ncell_x 9 --ncell_y 3 --ncell_z 2 --ne 1 --na 16 --niterations 3", 8, 2),
    ("c9_8x8x8_ne4_na32_nm16_manufactured", 16,
     "--ncell_x 8 --ncell_y 8 --ncell_z 8 --ne 4 --na 32", 0, 0),
]


def run_case(nm_build, argstring, seed, perturb):
    exe = os.path.join(REF, f"nm{nm_build}", "ref_harness")
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.bin")
        subprocess.check_call([exe, argstring, str(seed), str(perturb), out],
                              stdout=subprocess.DEVNULL)
        raw = open(out, "rb").read()
    hdr = np.frombuffer(raw[:32], dtype=np.int32)
    ncx, ncy, ncz, ne, nm, na, nu, nit = (int(x) for x in hdr)
    norms = np.frombuffer(raw[32:48], dtype=np.float64)
    n = ncx * ncy * ncz * ne * nm * nu
    arr = np.frombuffer(raw[48:], dtype=np.float64)
    assert arr.size == 3 * n
    return dict(dims=np.array([ncx, ncy, ncz, ne, nm, na, nu, nit], np.int64),
                normsq=norms[0], normsqdiff=norms[1],
                vi_in=arr[:n].copy(), vi=arr[n:2 * n].copy(),
                vo=arr[2 * n:].copy(), argstring=np.array(argstring))


def main():
    if not os.path.isdir(REF):
        print("oracle/_ref missing: run `make -C oracle ref` first", file=sys.stderr)
        return 1
    for name, nm_build, argstring, seed, perturb in CASES:
        d = run_case(nm_build, argstring, seed, perturb)
        if perturb == 0:
            # manufactured input: arrays are reproducible from the formulas;
            # keep only the norms the reference printed.
            for k in ("vi_in", "vi", "vo"):
                d.pop(k)
        np.savez(os.path.join(HERE, name + ".npz"), **d)
        print(f"{name}: normsq={d['normsq']:.8e} diff={d['normsqdiff']:.3e}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
