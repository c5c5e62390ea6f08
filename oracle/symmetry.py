"""Symmetry projection of a design field (oracle, FP64; test infrastructure only).

Restates parse_symmetry / symmetrize_vol_mirror / symmetrize_vol_cubic
(main_IHD_v3.py:221-337) with numpy: "mirror" averages the field with its
reflection about the centre plane of each selected axis in turn (x -> n-1-x);
"cubic" averages the 48 images under the 6 axis permutations times the 8
reflections.  Pinned by tests/test_oracle_symmetry.py through group
properties (idempotence, invariance under every generator, self-adjointness,
mean preservation) and a brute-force orbit average on a tiny cell.
"""
from __future__ import annotations

import itertools
from typing import List, Tuple

import numpy as np


def parse_symmetry(symmetry: str) -> Tuple[str, List[int]]:
    sym = (symmetry or "none").lower().strip()
    if sym in ("none", ""):
        return "none", []
    if sym in ("octant", "1/8", "one_eighth"):
        sym = "xyz"
    if sym in ("cubic", "iso", "isotropic"):
        return "cubic", [0, 1, 2]
    axes = sorted({a for a, ch in enumerate("xyz") if ch in sym})
    if not axes:
        return "none", []
    return "mirror", axes


def symmetrize_vol(vol: np.ndarray, kind: str, axes: List[int]) -> np.ndarray:
    if kind == "none":
        return vol.copy()
    if kind == "mirror":
        out = vol
        for ax in axes:
            out = 0.5 * (out + np.flip(out, axis=ax))
        return out
    if kind == "cubic":
        acc = np.zeros_like(vol)
        for p in itertools.permutations((0, 1, 2)):
            v = np.transpose(vol, p)
            for fl in itertools.product((0, 1), repeat=3):
                ax = tuple(a for a in range(3) if fl[a])
                acc = acc + (np.flip(v, axis=ax) if ax else v)
        return acc / 48.0
    raise ValueError(f"unknown symmetry kind {kind}")


def symmetrize(x: np.ndarray, r: int, kind: str, axes: List[int]) -> np.ndarray:
    return symmetrize_vol(np.asarray(x, dtype=np.float64).reshape(r, r, r), kind, axes).ravel()
