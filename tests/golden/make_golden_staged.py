"""Writes tests/golden/landmark_staged.npz from the reference build
(oracle/_ref, `make -C oracle ref`): LandmarkIndex end to end as the reference
runs on an AVX-512 VNNI host, where prepare() keeps sign planes from 32768
vectors and search() takes the staged sign-plane branch
(src/search/search.cpp:1379-1510). Needs an AVX-512 VNNI/VBMI host; the
fixture then serves the tests anywhere.

    python tests/golden/make_golden_staged.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(HERE))

from cases import STAGED_CASES, landmark_inputs  # noqa: E402
from oracle import Oracle  # noqa: E402


def main():
    R = Oracle("reference")
    out = {}
    for tag, n, nq, dim, m, seed, k in STAGED_CASES:
        x, q = landmark_inputs(n, nq, dim)
        s, i, info = R.lm_search(x, q, k, scan_bits=4, n_landmarks=m, seed=seed, planes=True)
        out[f"{tag}_scores"] = s
        out[f"{tag}_ids"] = i
    np.savez_compressed(HERE / "landmark_staged.npz", **out)
    print("wrote landmark_staged.npz", sorted(out))


if __name__ == "__main__":
    main()
