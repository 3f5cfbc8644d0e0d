"""Pins the CPU oracle (oracle/lj_oracle.c) before it is trusted as the checker.

Every check is independent of the oracle's own formulas: numpy's Philox, numpy's
IEEE float32 division and stable sort, exact rational arithmetic, closed forms of
the two-body potential, lattice sums, Newton's third law and finite differences.
"""
import ctypes
import json
import math
import pathlib
from fractions import Fraction

import numpy as np
import pytest

import ljamd
import pyoracle as po

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------------ Philox
def test_philox_known_answers():
    cases = json.loads((GOLDEN / "philox_kat.json").read_text())
    for c in cases:
        out = po.philox(np.array(c["ctr"], np.uint64), np.array(c["key"], np.uint64))
        assert [int(v) for v in out] == c["out"]


def test_philox_matches_numpy_live():
    for key in ([3, 5], [2**63 + 7, 11]):
        bg = np.random.Philox(key=np.array(key, np.uint64), counter=np.zeros(4, np.uint64))
        raw = bg.random_raw(12)
        for blk in range(3):
            out = po.philox(np.array([blk + 1, 0, 0, 0], np.uint64), np.array(key, np.uint64))
            assert np.array_equal(out, raw[4 * blk : 4 * blk + 4])


def test_generator_uniform_is_philox_top24():
    n = 5000
    x, y, z = ljamd.generate("uniform", 1234, n)
    bg = np.random.Philox(key=np.array([1234, 0], np.uint64), counter=np.array([2**64 - 1, 2**64 - 1, 2**64 - 1, 2**64 - 1], np.uint64))
    # counter (i, 0, 0, 0) for i = 0..n-1 is numpy's stream from counter -1 (all ones)
    raw = bg.random_raw(4 * n).reshape(n, 4)
    u = (raw[:, :3] >> np.uint64(40)).astype(np.float64) * 2.0**-24
    assert np.array_equal(x, u[:, 0].astype(np.float32))
    assert np.array_equal(y, u[:, 1].astype(np.float32))
    assert np.array_equal(z, u[:, 2].astype(np.float32))
    assert x.min() >= 0 and x.max() < 1


def test_generator_surface_and_clusters_formulas():
    n = 300
    x, y, z = ljamd.generate("surface", 99, n, [0.5, 0.5, 0.5, 1.0, 0.0])
    for i in range(n):
        w = po.philox(np.array([i, 1, 0, 0], np.uint64), np.array([99, 0], np.uint64))
        th = 2 * math.pi * (int(w[0]) >> 11) * 2.0**-53
        ph = math.acos(1 - 2 * (int(w[1]) >> 11) * 2.0**-53)
        assert x[i] == np.float32(math.sin(ph) * math.cos(th) + 0.5)
        assert y[i] == np.float32(math.sin(ph) * math.sin(th) + 0.5)
        assert z[i] == np.float32(math.cos(ph) + 0.5)
    r = np.sqrt((x - 0.5).astype(np.float64) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2)
    assert np.allclose(r, 1.0, atol=1e-6)
    x, y, z = ljamd.generate("clusters", 5, n, [4, 0.02, 0.1, 0.9])
    for i in range(0, n, 7):
        c = i % 4
        wc = po.philox(np.array([c, 2, 0, 0], np.uint64), np.array([5, 0], np.uint64))
        w = po.philox(np.array([i, 3, 0, 0], np.uint64), np.array([5, 0], np.uint64))
        u = [((int(v) >> 11) + 1) * 2.0**-53 for v in w]
        ctr = [0.1 + 0.8 * (int(v) >> 11) * 2.0**-53 for v in wc[:3]]
        r1 = math.sqrt(-2 * math.log(u[0]))
        g0 = r1 * math.cos(2 * math.pi * ((int(w[1]) >> 11) * 2.0**-53))
        assert x[i] == np.float32(ctr[0] + 0.02 * g0)


# ------------------------------------------------------------------ geometry / binning
def _np_keys(p, x, y, z):
    dims, _ = po.grid_dims(p)
    cw = np.float32(p.cell_edge if p.cell_edge > 0 else p.rc)
    cs = []
    for d, a in enumerate((x, y, z)):
        q = (a.astype(np.float32) - np.float32(p.box_lo[d])) / cw  # numpy float32: IEEE RN
        q = np.where(np.isnan(q), np.float32(0), q)
        q = np.clip(q, np.float32(0), np.float32(dims[d] - 1))
        cs.append(q.astype(np.int64))
    return ((cs[2] * dims[1] + cs[1]) * dims[0] + cs[0]).astype(np.uint32), dims


@pytest.mark.parametrize("rc,L", [(0.04, 1.0), (0.01, 1.0), (0.005, 1.0), (0.03, 2.0), (0.07, 0.9)])
def test_grid_dims_float32_rule(rc, L):
    p = po.params(rc, box_hi=(L, L, L))
    dims, nc = po.grid_dims(p)
    g = int(np.float32(L) / np.float32(rc))
    assert dims == (g, g, g) and nc == g**3
    # the engine's host helper agrees
    lp = ljamd.make_params(rc, box_hi=(L, L, L))
    assert ljamd.grid_dims(lp) == (dims, nc)


def test_largest_float_below_one_lands_in_cell_g_without_clamp():
    top = np.nextafter(np.float32(1), np.float32(0))
    for rc in (0.04, 0.01, 0.005):
        g = int(np.float32(1) / np.float32(rc))
        assert int(top / np.float32(rc)) == g  # the hazard the clamp removes
        p = po.params(rc)
        dims, _ = po.grid_dims(p)
        key = po.cell_key(p, dims, float(top), float(top), float(top))
        assert key == (g - 1) * (g * g + g + 1)


@pytest.mark.parametrize("case", ["uniform", "surface", "outside", "faces"])
def test_bin_matches_numpy(case):
    if case == "uniform":
        x, y, z = ljamd.generate("uniform", 1, 20000)
        p = po.params(0.04)
    elif case == "surface":
        x, y, z = ljamd.generate("surface", 2, 20000)
        p = po.params(0.05, box_lo=(-0.5,) * 3, box_hi=(1.5,) * 3)
    elif case == "outside":
        rng = np.random.default_rng(3)
        x, y, z = (rng.uniform(-0.3, 1.3, 5000).astype(np.float32) for _ in range(3))
        p = po.params(0.1)
    else:
        g = np.arange(0, 26, dtype=np.float32) * np.float32(0.04)
        x, y, z = (a.ravel() for a in np.meshgrid(g, g[:5], g[:3], indexing="ij"))
        p = po.params(0.04)
    off, perm, keys = po.bin(p, x, y, z)
    k_np, dims = _np_keys(p, x, y, z)
    assert np.array_equal(keys, k_np)
    cnt = np.bincount(k_np, minlength=int(np.prod(dims)))
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint32))
    assert np.array_equal(perm, np.argsort(k_np, kind="stable").astype(np.uint32))


def test_candidates_three_ways():
    x, y, z = ljamd.generate("uniform", 9, 3000)
    p = po.params(0.1)
    off, perm, keys = po.bin(p, x, y, z)
    dims, _ = po.grid_dims(p)
    c1 = po.candidates(p, off)
    c2 = po.candidates_direct(p, x, y, z, off)
    # brute force: ordered pairs (i != j) whose cells differ by at most 1 per axis
    cz, rem = np.divmod(keys.astype(np.int64), dims[0] * dims[1])
    cy, cx = np.divmod(rem, dims[0])
    near = (np.abs(cx[:, None] - cx[None]) <= 1) & (np.abs(cy[:, None] - cy[None]) <= 1) & (np.abs(cz[:, None] - cz[None]) <= 1)
    c3 = int(near.sum()) - x.size
    assert c1 == c2 == c3


# ------------------------------------------------------------------ predicate
def _round_f32(q: Fraction) -> Fraction:
    """Round a rational to the nearest float32 (ties to even), as an exact Fraction."""
    if q == 0:
        return Fraction(0)
    sign = -1 if q < 0 else 1
    a = abs(q)
    e = math.floor(math.log2(a.numerator) - math.log2(a.denominator))
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    e = max(e, -126)
    scale = Fraction(2) ** (e - 23)
    m = a / scale
    fl = math.floor(m)
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return sign * fl * scale


def test_canonical_d2_exact_rational():
    rng = np.random.default_rng(5)
    pts = rng.uniform(-1, 2, size=(300, 6)).astype(np.float32)
    pts[:100, 3:] = pts[:100, :3] + rng.uniform(-0.01, 0.01, (100, 3)).astype(np.float32)
    for r in pts:
        xi, yi, zi, xj, yj, zj = (Fraction(float(v)) for v in r)
        dx = _round_f32(xj - xi)
        dy = _round_f32(yj - yi)
        dz = _round_f32(zj - zi)
        d2 = _round_f32(dz * dz + _round_f32(dy * dy + _round_f32(dx * dx)))
        assert Fraction(po.d2(r[:3], r[3:])) == d2


def test_pairs_brute_equals_cellgrid_all_predicates():
    x, y, z = ljamd.generate("uniform", 17, 2500)
    for pred in ("aabb", "custom", "sphere", "triangle"):
        p = po.params(0.07, predicate=pred)
        rb, cb = po.pairs_brute(p, x, y, z)
        rg, cg = po.pairs_cellgrid(p, x, y, z)
        assert np.array_equal(rb, rg) and np.array_equal(cb, cg)
        # symmetric: j in row i iff i in row j
        ii = np.repeat(np.arange(x.size), np.diff(rb.astype(np.int64)))
        a = set(zip(ii.tolist(), cb.tolist()))
        assert a == {(j, i) for i, j in a}


def test_predicate_boundaries():
    rc = np.float32(0.25)
    rc2 = np.float32(rc * rc)
    for pred, expect in (("aabb", True), ("custom", False), ("sphere", False)):
        # a pair at exactly d2 == rc2: x-offset rc, others 0
        assert bool(po.lib().orc_accept(po.PRED[pred], float(rc2), float(rc), 0, 1)) is expect
    assert not po.lib().orc_accept(0, 1e-17, 0.25, 0, 1)  # below 1e-16
    assert po.lib().orc_accept(0, 1e-16, 0.25, 0, 1)
    assert not po.lib().orc_accept(4, float(np.float32(1e-4) ** 2), 0.25, 0, 1)  # triangle: d must exceed 1e-4
    assert po.lib().orc_accept(2, 0.0, 0.25, 0, 1) and not po.lib().orc_accept(2, 0.0, 0.25, 3, 3)


# ------------------------------------------------------------------ ray-hit models
def _mult(pred, c, pi, pj):
    q = po.params(c, predicate=pred)
    return po.lib().orc_pair_mult(ctypes.byref(q), *(float(v) for v in pi), *(float(v) for v in pj), 0, 1)


def _np_sphere_mult(c, pi, pj):
    """Interval form of the Sphere program's rule, written independently of the
    oracle's: the pair counts when d2 < c^2 and the ray along the axis closest to p_j
    reaches the sphere's surface within [0, 2h] (first crossing)."""
    c = np.float32(c)
    h = np.float32(c * np.float32(0.81649658092))
    rad = np.float32(h + np.float32(1e-5))
    d2 = np.float32(po.d2(pi, pj))
    if not d2 < np.float32(c * c):
        return 0
    sq = [np.float32(abs(np.float32(pi[k] - pj[k]))) ** 2 for k in range(3)]
    dist_axis = [np.float32(sq[1] + sq[2]), np.float32(sq[0] + sq[2]), np.float32(sq[0] + sq[1])]
    a = int(np.argmin(dist_axis))  # first minimum: x, then y, then z
    o_a = np.float32(pi[a] - h)
    along = float(pj[a]) - float(o_a)
    perp2 = sum((float(pj[b]) - float(pi[b])) ** 2 for b in range(3) if b != a)
    disc = float(rad) ** 2 - perp2
    if disc < 0:
        return 0
    t1, t2 = along - math.sqrt(disc), along + math.sqrt(disc)
    t = t1 if t1 >= 0 else t2
    return int(0 <= t <= 2.0 * float(h))


def _np_triangle_mult(c, pi, pj):
    """Interval form of the Triangle program's rule: a ray at (y_r, z_r) hits a panel
    when it lies inside the panel's y and z extents and the panel's x plane is within
    [tmin, tmax] of the ray origin; each hit counts when the centre q rebuilt from the
    panel satisfies 1e-4 < |q - p_i| < c and the ray-index rule."""
    f = np.float32
    c = f(c)
    hc = f(c / f(2))
    hi_off = f(hc + f(1e-5))
    lo_off = f(-hc - f(1e-5))
    y_lo, y_hi = f(pj[1] + lo_off), f(pj[1] + hi_off)
    z_lo, z_hi = f(pj[2] + lo_off), f(pj[2] + hi_off)
    panels = [(f(pj[0] + (-hc)), 0), (f(pj[0] + hc), 1)]
    ox = f(pi[0] - hc)
    tmin = float(f(1.19209290e-07))
    tmax = float(f(float(f(2) * hc) + float(f(2) * hc) * 0.00001))
    m = 0
    for r in range(4):
        yc, zc = (f(-1) if r & 1 else f(1)), (f(-1) if r & 2 else f(1))
        yr, zr = f(pi[1] + f(hc * yc)), f(pi[2] + f(hc * zc))
        if not (y_lo <= yr <= y_hi and z_lo <= zr <= z_hi):
            continue
        for xp, back in panels:
            t = float(xp) - float(ox)
            if not (tmin <= t <= tmax):
                continue
            q = (f(xp - hc) if back else f(xp + hc), f(f(y_lo + y_hi) / f(2)), f(f(z_lo + z_hi) / f(2)))
            d = f(np.sqrt(np.float32(po.d2(pi, q))))
            if not (d < c and float(d) > 1e-4):
                continue
            ok = ((pi[1] != q[1] and pi[2] != q[2]) or (pi[2] < q[2] and r == 0) or (pi[2] > q[2] and r == 2)
                  or (pi[1] < q[1] and r == 0) or (pi[1] > q[1] and r == 1) or r == 0)
            m += int(ok)
    return m


def _ray_pairs(c, n, seed):
    """Pairs around the rules' edges: generic offsets inside and beyond c, offsets
    within 2e-5 of zero on one or two axes (several rays or panels), near-coincident
    pairs, and x offsets within c 1e-5 (both panels)."""
    rng = np.random.default_rng(seed)
    pi = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
    off = rng.uniform(-1.3 * c, 1.3 * c, (n, 3))
    kind = rng.integers(0, 5, n)
    off[kind == 1, 1] = rng.uniform(-2e-5, 2e-5, (kind == 1).sum())
    k2 = kind == 2
    off[k2, 1] = rng.uniform(-2e-5, 2e-5, k2.sum())
    off[k2, 2] = rng.uniform(-2e-5, 2e-5, k2.sum())
    k3 = kind == 3
    off[k3] = rng.uniform(-3e-5, 3e-5, (k3.sum(), 3))
    off[kind == 4, 0] = rng.uniform(-2e-5 * c, 2e-5 * c, (kind == 4).sum())
    pj = (pi + off).astype(np.float32)
    return pi, pj


@pytest.mark.parametrize("pred", ["sphere", "triangle"])
def test_ray_models_match_independent_interval_form(pred):
    c = 0.05
    pi, pj = _ray_pairs(c, 3000, 8 if pred == "sphere" else 9)
    ref = _np_sphere_mult if pred == "sphere" else _np_triangle_mult
    got = np.array([_mult(pred, c, a, b) for a, b in zip(pi, pj)])
    want = np.array([ref(c, a, b) for a, b in zip(pi, pj)])
    assert np.array_equal(got, want), np.flatnonzero(got != want)[:10]
    if pred == "triangle":
        assert (got >= 2).sum() > 100 and (got == 4).sum() > 10  # the degenerate families occur
    else:
        assert (got == 0).sum() > 100 and got.max() == 1


def test_ray_models_hand_placed_pairs():
    c = 0.05
    p = np.array([0.4, 0.5, 0.6], np.float32)

    def at(dx, dy, dz):
        return (p + np.array([dx, dy, dz])).astype(np.float32)

    # Sphere: generic pairs count once inside c; a pair closer than the 1e-5 radius
    # margin along the ray's axis is missed (the whole ray segment lies inside the sphere)
    assert _mult("sphere", c, p, at(0.3 * c, 0.2 * c, 0.25 * c)) == 1
    assert _mult("sphere", c, p, at(0.0, 0.0, 0.999 * c)) == 1
    assert _mult("sphere", c, p, at(0.0, 0.0, 1.001 * c)) == 0
    assert _mult("sphere", c, p, at(0.0, 0.0, 1e-6)) == 0
    assert _mult("sphere", c, p, at(0.0, 0.0, 5e-5)) == 1
    assert _mult("sphere", c, p, p) == 0
    # Triangle: one ray and one panel for generic pairs; two rays when |dy| < 1e-5, four
    # when |dz| < 1e-5 too; both panels when 0 < dx < c 1e-5; none below d = 1e-4
    assert _mult("triangle", c, p, at(0.3 * c, 0.2 * c, 0.25 * c)) == 1
    assert _mult("triangle", c, p, at(-0.3 * c, -0.2 * c, -0.25 * c)) == 1
    assert _mult("triangle", c, p, at(0.3 * c, 5e-6, 0.25 * c)) == 2
    assert _mult("triangle", c, p, at(0.3 * c, 5e-6, -5e-6)) == 4
    assert _mult("triangle", c, p, at(2e-7, 0.2 * c, 0.25 * c)) == 2
    assert _mult("triangle", c, p, at(5e-5, 0.0, 0.0)) == 0
    assert _mult("triangle", c, p, at(0.3 * c, 0.2 * c, 1.1 * c)) == 0


def test_ray_models_reduce_to_distance_predicates_on_generic_pairs():
    """Away from the degenerate families the Sphere program finds exactly the pairs
    with d2 < c^2 and the Triangle program those with 1e-4 < d < c, each once."""
    c = 0.05
    rng = np.random.default_rng(12)
    pi = rng.uniform(0.2, 0.8, (4000, 3)).astype(np.float32)
    off = rng.uniform(-1.2 * c, 1.2 * c, (4000, 3))
    off = np.where(np.abs(off) < 1e-3, 1e-3, off)  # no axis within 1e-3
    pj = (pi + off).astype(np.float32)
    d2 = np.array([po.d2(a, b) for a, b in zip(pi, pj)], np.float32)
    keep = np.abs(np.sqrt(d2.astype(np.float64)) - c) > 1e-6 * c  # clear of the cutoff
    for a, b, q, k in zip(pi, pj, d2, keep):
        if not k:
            continue
        assert _mult("sphere", c, a, b) == int(q < np.float32(c * c))
        assert _mult("triangle", c, a, b) == int(np.sqrt(q) < c)


# ------------------------------------------------------------------ FP64 forces
def _two(r, sigma=0.3, eps=1.7):
    x = np.array([0.4, 0.4 + r], np.float32)
    y = np.array([0.5, 0.5], np.float32)
    z = np.array([0.5, 0.5], np.float32)
    p = po.params(0.9, sigma=sigma, epsilon=eps)
    return po.forces_f64(p, x, y, z), float(x[1]) - float(x[0])


def test_two_body_closed_forms():
    sigma, eps = np.float32(0.3), np.float32(1.7)
    # V(sigma) = 0 and |F(sigma)| = 24 eps / sigma
    r, d = _two(float(sigma), sigma, eps)
    s = float(sigma) / d
    assert r["e"][0] == pytest.approx(4 * float(eps) * (s**12 - s**6), abs=1e-9)
    assert abs(r["e"][0]) < 1e-5
    assert -r["fx"][0] == pytest.approx(24 * float(eps) * (2 * s**12 - s**6) / d, rel=1e-12)
    assert -r["fx"][0] == pytest.approx(24 * float(eps) / float(sigma), rel=1e-5)
    assert r["fx"][1] == pytest.approx(-r["fx"][0], rel=1e-15)
    # minimum at 2^(1/6) sigma: V = -eps, F = 0
    r, d = _two(2 ** (1 / 6) * float(sigma), sigma, eps)
    assert r["e"][0] == pytest.approx(-float(eps), rel=1e-6)
    assert abs(r["fx"][0]) < 1e-4 * 24 * float(eps) / float(sigma)
    assert r["e_total"] == pytest.approx(r["e"][0], rel=1e-15)  # 1/2 (e0 + e1), e0 == e1
    assert r["n_pairs"] == 2


def test_simple_cubic_lattice_sums():
    a = 0.0625  # exact in binary
    g = np.arange(1, 15) * a
    x, y, z = (v.ravel().astype(np.float32) for v in np.meshgrid(g, g, g, indexing="ij"))
    sigma, eps = float(np.float32(0.05)), 1.0
    p = po.params(1.2 * a, sigma=sigma, epsilon=eps)  # a < rc < sqrt(2) a: 6 neighbours
    r = po.forces_f64(p, x, y, z)
    interior = (x > 1.5 * a) & (x < 13.5 * a) & (y > 1.5 * a) & (y < 13.5 * a) & (z > 1.5 * a) & (z < 13.5 * a)
    s6 = (sigma / a) ** 6
    assert np.allclose(r["e"][interior], 6 * 4 * eps * (s6 * s6 - s6), rtol=1e-12)
    scale = 24 * eps * s6 * (2 * s6 - 1) / a
    assert np.abs(r["fx"][interior]).max() <= 1e-9 * abs(scale)
    # a surface particle feels the unbalanced bond: |F| = |24 eps s6 (2 s6 - 1) / a|
    face = (x == np.float32(a)) & interior.reshape(x.shape).__invert__() & (y > 1.5 * a) & (y < 13.5 * a) & (z > 1.5 * a) & (z < 13.5 * a)
    assert np.allclose(np.abs(r["fx"][face]), abs(scale), rtol=1e-10)
    # bonds: 3 * 14^2 * 13 unordered nearest-neighbour pairs
    assert r["n_pairs"] == 2 * 3 * 14 * 14 * 13


def test_newton_third_law_and_invariances():
    x, y, z = ljamd.generate("uniform", 4, 4000)
    # on a 2^-16 grid, adding 0.25 is exact, so translated differences are identical
    x, y, z = (np.floor(a * np.float32(65536)) / np.float32(65536) for a in (x, y, z))
    p = po.params(0.09)
    r = po.forces_f64(p, x, y, z)
    mag = np.abs(r["fx"]).sum()
    assert abs(r["fx"].sum()) < 1e-9 * mag and abs(r["fy"].sum()) < 1e-9 * mag
    # permutation: outputs follow the particles
    perm = np.random.default_rng(0).permutation(x.size)
    q = po.forces_f64(p, x[perm], y[perm], z[perm])
    assert np.allclose(q["fx"], r["fx"][perm], rtol=1e-12, atol=1e-12 * mag / x.size)
    assert q["e_total"] == pytest.approx(r["e_total"], rel=1e-12)
    # translation by an exact power of two keeps every difference exact
    t = np.float32(0.25)
    p2 = po.params(0.09, box_lo=(0.25,) * 3, box_hi=(1.25,) * 3)
    s = po.forces_f64(p2, x + t, y + t, z + t)
    assert s["n_pairs"] == r["n_pairs"]
    assert np.allclose(s["fx"], r["fx"], rtol=1e-12, atol=1e-12 * mag / x.size)


def test_force_is_minus_energy_gradient():
    rng = np.random.default_rng(8)
    x, y, z = (rng.uniform(0.3, 0.5, 40).astype(np.float32) for _ in range(3))
    p = po.params(0.9, sigma=0.02)  # cutoff beyond every distance: E is smooth
    r = po.forces_f64(p, x, y, z)
    for i in (0, 7, 21):
        h = np.float32(2.0**-20)
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        ep = po.forces_f64(p, xp, y, z)["e_total"]
        em = po.forces_f64(p, xm, y, z)["e_total"]
        hd = float(xp[i]) - float(x[i])
        fd = -(ep - em) / (2 * hd)
        assert fd == pytest.approx(r["fx"][i], rel=1e-5, abs=1e-8 * np.abs(r["fx"]).max())


def test_virial_is_half_sum_w_r2():
    x, y, z = ljamd.generate("uniform", 12, 1500)
    p = po.params(0.12)
    r = po.forces_f64(p, x, y, z)
    rows, cols = po.pairs_brute(p, x, y, z)
    ii = np.repeat(np.arange(x.size), np.diff(rows.astype(np.int64)))
    d = np.stack([a[cols].astype(np.float64) - a[ii] for a in (x, y, z)])
    r2 = (d**2).sum(0)
    s6 = (np.float64(p.sigma) ** 2 / r2) ** 3
    w = 24 * s6 * (2 * s6 - 1) / r2
    assert r["virial"] == pytest.approx(0.5 * (w * r2).sum(), rel=1e-10)
    assert r["e_total"] == pytest.approx(0.5 * (4 * (s6 * s6 - s6)).sum(), rel=1e-10)


def test_fp32_baseline_vs_fp64():
    x, y, z = ljamd.generate("uniform", 1234, 100_000)
    p = po.params(0.04)
    r64 = po.forces_f64(p, x, y, z)
    r32 = po.forces_f32(p, x, y, z, nthreads=4)
    assert r32["n_pairs"] == r64["n_pairs"] and r32["n_candidates"] == r64["n_candidates"]
    err = po.rel_errors((r32["fx"], r32["fy"], r32["fz"]), (r64["fx"], r64["fy"], r64["fz"]), r64["cond"])
    assert err["cond"].max() < 3e-6
    assert np.sqrt((err["rel"] ** 2).mean()) < 2e-6


def test_cudalj_energy_vs_numpy_restatement():
    """cudaLJ.cu:28-56 in FP64 numpy over the clipped 27-cell candidates."""
    x, y, z = ljamd.generate("uniform", 33, 600)
    p = po.params(0.25, sigma=1.0, epsilon=1.0, softening2=1e-5, predicate="cudalj")
    e = po.cudalj_energy(p, x, y, z)
    keys, dims = _np_keys(p, x, y, z)
    cz, rem = np.divmod(keys.astype(np.int64), dims[0] * dims[1])
    cy, cx = np.divmod(rem, dims[0])
    near = (np.abs(cx[:, None] - cx[None]) <= 1) & (np.abs(cy[:, None] - cy[None]) <= 1) & (np.abs(cz[:, None] - cz[None]) <= 1)
    np.fill_diagonal(near, False)
    d2 = sum((a[None, :].astype(np.float64) - a[:, None]) ** 2 for a in (x, y, z))
    r = np.sqrt(d2 + np.float64(np.float32(1e-5)))
    s6 = (1.0 / r) ** 6
    ref = np.where(near, 4 * (s6 * s6 - s6), 0).sum(1)
    assert np.allclose(e, ref, rtol=1e-12)
    # the FP32 cell-grid restatement in cudaLJ mode counts every candidate
    r32 = po.forces_f32(p, x, y, z)
    assert r32["n_pairs"] == r32["n_candidates"] == int(near.sum())


# ------------------------------------------------------------------ golden fixtures
@pytest.mark.parametrize("name", ["uniform", "surface", "clusters"])
def test_golden_small_systems(name):
    g = np.load(GOLDEN / "small_systems.npz")
    x, y, z = g[f"{name}_pos"]
    prm = g[f"{name}_params"]
    p = po.params(float(prm[0]), sigma=float(prm[1]), epsilon=float(prm[2]), box_lo=prm[3:6], box_hi=prm[6:9])
    kind, seed, n, gp = {"uniform": ("uniform", 1234, 3000, None), "surface": ("surface", 1235, 3000, None),
                         "clusters": ("clusters", 1238, 3000, [8, 0.02, 0.1, 0.9])}[name]
    gx, gy, gz = ljamd.generate(kind, seed, n, gp)
    assert np.array_equal(gx, x) and np.array_equal(gy, y) and np.array_equal(gz, z)
    off, perm, _ = po.bin(p, x, y, z)
    assert np.array_equal(off, g[f"{name}_offsets"]) and np.array_equal(perm, g[f"{name}_perm"])
    r = po.forces_f64(p, x, y, z)
    assert [r["n_candidates"], r["n_pairs"]] == list(g[f"{name}_counts"])
    f = g[f"{name}_f64"]
    assert np.array_equal(np.stack([r["fx"], r["fy"], r["fz"], r["e"]]), f)
    rb, _ = po.pairs_brute(p, x, y, z)
    assert int(rb[-1]) == r["n_pairs"]
