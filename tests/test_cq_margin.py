"""The ray-space record test of the traversal kernel (trace.hip: cq_axis_ray,
slab_ray) lowers near distances and raises far ones by a margin P meant to
cover every rounding difference from the reference's slab distance
fl(fl(b - o) * id) (bbox.rs:185-209) to the exact plane b of a box whose
quantized plane decodes (one rounded fma, as the device does) on the safe
side of b.  This checks that claim with exact rational arithmetic on
adversarial inputs (decoded plane equal to the exact one, ray origins on or
near the record origin, large origins with small steps, tiny and huge 1/d),
and that it fails when a term of P is dropped."""
from fractions import Fraction as Fr

import numpy as np
import pytest

F32 = np.float32


def r32(x):
    """x (a Fraction) rounded to the nearest float32, ties to even."""
    c = F32(float(x))
    best = None
    for cand in (np.nextafter(c, F32(-np.inf)), c, np.nextafter(c, F32(np.inf))):
        if not np.isfinite(cand):
            continue
        d = abs(Fr(float(cand)) - x)
        key = (d, int(np.frombuffer(cand.tobytes(), np.uint32)[0]) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return best[1]


def fma32(a, b, c):
    return r32(Fr(float(a)) * Fr(float(b)) + Fr(float(c)))


def margin_terms(o, s, ro, idv, drop=None):
    A = F32(s) * F32(idv)
    B = (F32(o) - F32(ro)) * F32(idv)
    t_o = F32(0) if drop == "o" else abs(F32(o))
    t_a = F32(0) if drop == "a" else abs(A)
    t_b = F32(0) if drop == "b" else abs(B)
    m = fma32(F32(255.0), t_a, fma32(t_o, abs(F32(idv)), t_b))
    P = fma32(m, F32(2.0 ** -20), F32(2.0 ** -100))
    return A, B - P, B + P


def check_plane(o, s, q, b, role, ro, idv, drop=None):
    """role 'lo': b is a box's low plane, the decoded plane fma(q, s, o) <= b;
    'hi': a high plane, decoded >= b.  Returns True when the ray-space
    distance is on the safe side of the reference's distance to b."""
    A, bn, bf = margin_terms(o, s, ro, idv, drop)
    tn = fma32(F32(q), A, bn)
    tf = fma32(F32(q), A, bf)
    ref = (F32(b) - F32(ro)) * F32(idv)
    near_side = (role == "lo") == (idv > 0)
    return bool(tn <= ref) if near_side else bool(tf >= ref)


def cases(rng, n):
    out = []
    for _ in range(n):
        mag = 2.0 ** rng.integers(-20, 40)
        o = F32(rng.uniform(-1, 1) * mag)
        e = int(rng.integers(-30, 12))
        e = max(e, int(np.floor(np.log2(abs(float(o)) + 1e-30))) - 30)
        s = F32(2.0 ** e)
        q = int(rng.integers(0, 256))
        dec = fma32(F32(q), s, o)
        # the exact plane on the safe side of the decoded one, often equal
        k = int(rng.integers(0, 3))
        lo_b = dec
        hi_b = dec
        for _ in range(k):
            lo_b = np.nextafter(lo_b, F32(np.inf))
            hi_b = np.nextafter(hi_b, F32(-np.inf))
        # ray origin: on the record origin, on the plane, or nearby
        sel = rng.integers(0, 4)
        if sel == 0:
            ro = o
        elif sel == 1:
            ro = dec
        elif sel == 2:
            ro = F32(float(o) + rng.normal() * float(s) * 300)
        else:
            ro = F32(rng.uniform(-1, 1) * 2.0 ** rng.integers(-10, 40))
        d = F32(rng.choice([-1, 1]) * 2.0 ** rng.uniform(-40, 20))
        idv = F32(1.0) / d
        if abs(idv) > 2.0 ** 40 or abs(ro) > 2.0 ** 40 or abs(o) > 2.0 ** 40:
            continue
        out.append((o, s, q, lo_b, hi_b, ro, idv))
    return out


@pytest.fixture(scope="module")
def sample():
    return cases(np.random.default_rng(7), 2500)


def test_margin_covers_reference_distance(sample):
    assert len(sample) > 1500
    for o, s, q, lo_b, hi_b, ro, idv in sample:
        assert check_plane(o, s, q, lo_b, "lo", ro, idv), (o, s, q, lo_b, ro, idv)
        assert check_plane(o, s, q, hi_b, "hi", ro, idv), (o, s, q, hi_b, ro, idv)


@pytest.mark.parametrize("drop", ["o", "a", "b"])
def test_dropping_a_margin_term_fails(sample, drop):
    bad = 0
    for o, s, q, lo_b, hi_b, ro, idv in sample:
        bad += not check_plane(o, s, q, lo_b, "lo", ro, idv, drop)
        bad += not check_plane(o, s, q, hi_b, "hi", ro, idv, drop)
    assert bad > 0


def test_margin_is_small(sample):
    # the widening stays a few parts in 2^20 of the terms it covers
    for o, s, q, lo_b, hi_b, ro, idv in sample[:300]:
        A, bn, bf = margin_terms(o, s, ro, idv)
        B = (F32(o) - F32(ro)) * F32(idv)
        scale = abs(float(B)) + abs(float(o) * float(idv)) + 255 * abs(float(A)) + 2.0 ** -100
        assert float(bf) - float(bn) <= 2.0 ** -18 * scale
