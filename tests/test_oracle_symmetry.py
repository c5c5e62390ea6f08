"""Pins the oracle's symmetry projection (main_IHD_v3.py:221-337): it is the
orbit average of a group of index maps, hence an orthogonal projection."""
import itertools

import numpy as np
import pytest

from oracle import symmetry as osym


def test_parse_symmetry_aliases():
    assert osym.parse_symmetry("none") == ("none", [])
    assert osym.parse_symmetry("") == ("none", [])
    assert osym.parse_symmetry("octant") == ("mirror", [0, 1, 2])
    assert osym.parse_symmetry("XY") == ("mirror", [0, 1])
    assert osym.parse_symmetry("z") == ("mirror", [2])
    assert osym.parse_symmetry("iso") == ("cubic", [0, 1, 2])
    assert osym.parse_symmetry("abc") == ("none", [])


@pytest.mark.parametrize("kind,axes", [("mirror", [0]), ("mirror", [0, 2]), ("mirror", [0, 1, 2]), ("cubic", [0, 1, 2])])
def test_projection_properties(kind, axes):
    r = 6
    rng = np.random.default_rng(0)
    v = rng.standard_normal((r, r, r))
    w = rng.standard_normal((r, r, r))
    P = lambda a: osym.symmetrize_vol(a, kind, axes)  # noqa: E731
    pv = P(v)
    np.testing.assert_allclose(P(pv), pv, atol=1e-14)                    # idempotent
    assert np.isclose((pv * w).sum(), (v * P(w)).sum())                   # self-adjoint
    assert np.isclose(pv.mean(), v.mean())                                # mean preserving
    gens = [(lambda a, ax=ax: np.flip(a, axis=ax)) for ax in axes]
    if kind == "cubic":
        gens += [lambda a: np.transpose(a, (1, 0, 2)), lambda a: np.transpose(a, (1, 2, 0))]
    for g in gens:
        np.testing.assert_allclose(g(pv), pv, atol=1e-14)                 # invariant


def test_cubic_brute_force_orbit():
    r = 3
    v = np.random.default_rng(1).standard_normal((r, r, r))
    out = np.zeros_like(v)
    for x, y, z in itertools.product(range(r), repeat=3):
        imgs = []
        for p in itertools.permutations((x, y, z)):
            for fl in itertools.product((0, 1), repeat=3):
                q = tuple(r - 1 - c if f else c for c, f in zip(p, fl))
                imgs.append(v[q])
        out[x, y, z] = np.mean(imgs)
    np.testing.assert_allclose(osym.symmetrize_vol(v, "cubic", [0, 1, 2]), out, atol=1e-14)
