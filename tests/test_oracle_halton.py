"""Pins for the oracle's HaltonSampler (halton.rs:48-107) and radical_inverse
(montecarlo.rs:8-21): the reference's radical-inverse vectors, the inverse
worked exactly here by digit reversal in rational arithmetic, and the
sampler's sequence rebuilt from it: every kept sample at (delta u, delta v)
with u, v the base-3 and base-2 inverses rounded to f32, kept iff inside the
image, grouped by pixel in sequence order; the counts add up to the kept
samples, and the squares of the sequence fall into pixels as the closed form
says (for a square power-of-two image, each y row receives exactly
side * spp samples, since the base-2 inverse of k < side^2 spp is
stratified in 2^m strata)."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as ol

F32 = np.float32
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


def test_radical_inverse_reference_vectors():
    L = ol.lib()
    g = GOLD["radical_inverse"]
    for n, b, v in g["exact"]:
        assert L.orc_radical_inverse(n, b) == v, (n, b)
    for n, b, num, den in g["approx_1e-6"]:
        assert abs(L.orc_radical_inverse(n, b) - num / den) < 1e-6


def rinv_exact(n, b):
    v, base = Fraction(0), Fraction(1, b)
    while n:
        v += (n % b) * base
        n //= b
        base /= b
    return v


def test_radical_inverse_against_exact_digit_reversal():
    L = ol.lib()
    rng = np.random.default_rng(2)
    for n in list(range(200)) + [int(x) for x in rng.integers(0, 2 ** 40, 200)]:
        for b in (2, 3, 5, 7):
            got = L.orc_radical_inverse(n, b)
            want = rinv_exact(n, b)
            assert abs(Fraction(got) - want) <= Fraction(1, 2 ** 49), (n, b)
            if b == 2:
                assert Fraction(got) == want  # dyadic: exact in f64


@pytest.mark.parametrize("xres,yres,spp", [(16, 16, 4), (24, 10, 3), (7, 19, 1), (32, 32, 8)])
def test_halton_table_rebuilds_from_the_sequence(xres, yres, spp):
    K, cnt, pos = ol.halton_table(xres, yres, spp)
    side = max(xres, yres)
    want = [[] for _ in range(xres * yres)]
    for k in range(side * side * spp):
        u = F32(float(rinv_exact(k, 3)))
        v = F32(float(rinv_exact(k, 2)))
        ix = F32(F32(0) * (F32(1) - u)) + F32(F32(side) * u)  # lerp(0, delta, u) in f32
        iy = F32(F32(0) * (F32(1) - v)) + F32(F32(side) * v)
        if ix >= xres or iy >= yres:
            continue
        want[int(iy) * xres + int(ix)].append((ix, iy))
    assert K == max(len(w) for w in want)
    assert np.array_equal(cnt.reshape(-1), [len(w) for w in want])
    flat = pos.reshape(-1, K, 2)
    for p, w in enumerate(want):
        if w:
            assert np.array_equal(flat[p, :len(w)], np.array(w, F32)), p
        assert np.all(flat[p, len(w):] == 0)
    if xres == yres and (xres & (xres - 1)) == 0:
        assert np.all(cnt.sum(axis=1) == xres * spp)  # base-2 strata fill every row
    assert cnt.sum() == sum(len(w) for w in want)
