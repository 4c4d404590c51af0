"""Writes tests/golden/codec_streams.json: inputs, streams and decoded grids of the
reference's analysis codec (native/src/aero_lbm_analysis_codec.cpp with the zfp
sources it vendors), compiled from its sources into oracle/_ref/libcodec_ref.so by
`make -C oracle ref` in the build container.

    python tests/golden/make_codec_golden.py
"""
import base64
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import codec_cases as CC  # noqa: E402

CASES = [
    ("smooth", (7, 9, 13), 1e-3, 1),
    ("noise", (8, 8, 8), CC.VELOCITY_TOL, 2),
    ("special", (9, 5, 6), 1e-3, 3),
    ("zero", (7, 6, 5), CC.PRESSURE_TOL, 4),
    ("smooth", (4, 4, 4), 5e-324, 5),
    ("mixed", (4, 8, 4), 1e300, 6),
    ("velocity", (16, 16, 16), CC.VELOCITY_TOL, 7),
    ("pressure", (16, 16, 16), CC.PRESSURE_TOL, 8),
    ("subnormal", (5, 3, 6), 1e-44, 9),
]


def b64(a: bytes) -> str:
    return base64.b64encode(a).decode("ascii")


def main():
    ref = CC.reference()
    if ref is None:
        sys.exit("oracle/_ref/libcodec_ref.so is missing: run `make -C oracle ref` first")
    out = []
    for kind, shape, tol, seed in CASES:
        v = CC.grid(kind, shape, seed)
        s = ref.compress(v, tol)
        d = ref.decompress(s, shape[2], shape[1], shape[0])
        out.append({"kind": kind, "shape": list(shape), "tolerance": tol, "seed": seed,
                    "input": b64(v.tobytes()), "stream": b64(s), "decoded": b64(d.tobytes())})
        print(kind, shape, tol, len(s), "bytes")
    with open(os.path.join(HERE, "codec_streams.json"), "w") as f:
        json.dump({"source": "native/src/aero_lbm_analysis_codec.cpp + native/third_party/zfp (1.0.1)",
                   "cases": out}, f, indent=1)


if __name__ == "__main__":
    main()
