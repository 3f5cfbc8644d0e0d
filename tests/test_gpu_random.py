"""Seeded random configurations against the oracle, through the C ABI: sizes from 0 to
6,000, boxes of random origin and edge with some particles outside them, uniform,
clustered and sphere-surface positions (with exact duplicates), random cutoffs, every
cutoff predicate (the AABB and Custom distance rules, the Sphere and Triangle ray
models), random kernel shapes (force-grid refinement, workgroup size, LDS queue) and
both engines (cell grid, LBVH).

Bars as in the fixed-case tests: neighbour lists (CSR) bit-exact against the oracle's
cell grid (and its O(N^2) brute force for N <= 2,000), accepted-pair counts exact, the
rc-grid candidate count exact for the cell grid, forces within 1e-5 of the FP64 oracle
per particle (mixed precision), and the energy total within 1e-9 of the oracle's."""
import numpy as np
import pytest

import ljamd
import pyoracle as po

pytestmark = pytest.mark.gpu

PREDICATES = ["aabb", "custom", "sphere", "triangle"]
SHAPES = [(1, 1), (1, 2), (1, 4), (2, 2), (2, 4), (1, 8), (2, 8), (4, 4)]


def _config(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.choice([0, 1, 2, 33, 500, 2000, 6000]))
    lo = rng.uniform(-2.0, 2.0, 3)
    edge = float(rng.uniform(0.5, 3.0))
    hi = lo + edge
    kind = str(rng.choice(["uniform", "clusters", "surface"]))
    rc = edge * float(rng.uniform(0.02, 0.15))
    if kind == "uniform":
        P = rng.uniform(lo - 0.05 * edge, hi + 0.05 * edge, (n, 3))
    elif kind == "clusters":
        centres = rng.uniform(lo, hi, (4, 3))
        P = centres[rng.integers(0, 4, n)] + rng.normal(0.0, 0.05 * edge, (n, 3))
    else:
        v = rng.normal(size=(n, 3))
        v /= np.maximum(np.linalg.norm(v, axis=1, keepdims=True), 1e-30)
        P = (lo + hi) / 2 + 0.4 * edge * v
    P = P.astype(np.float32)
    if n >= 33:  # exact duplicates: d2 = 0 pairs the predicates must reject
        dup = rng.integers(0, n, n // 33)
        P[rng.integers(0, n, dup.size)] = P[dup]
    pred = str(rng.choice(PREDICATES))
    refine, refine_x = SHAPES[int(rng.integers(0, len(SHAPES)))]
    block = int(rng.choice([64, 128, 256]))
    depth = int(rng.choice([1, 1, 8]))
    lbvh = bool(rng.integers(0, 2))
    p = ljamd.make_params(rc, box_lo=tuple(float(v) for v in lo), box_hi=tuple(float(v) for v in hi),
                          predicate=pred, refine=refine, refine_x=refine_x, block=block, depth=depth)
    desc = (f"seed={seed} n={n} {kind} lo={np.round(lo, 3)} edge={edge:.3f} rc={rc:.4f} {pred} "
            f"shape=({refine},{refine_x},{block},{depth}) lbvh={lbvh}")
    return np.ascontiguousarray(P[:, 0]), np.ascontiguousarray(P[:, 1]), np.ascontiguousarray(P[:, 2]), p, lbvh, desc


@pytest.mark.parametrize("seed", range(24))
def test_random_configuration_matches_oracle(engine, seed):
    x, y, z, p, lbvh, desc = _config(seed)
    n = x.size
    print(f"[random] {desc}", flush=True)
    q = po.from_lj(p)
    dx, dy, dz = engine.to_device(x, y, z)
    rows, cols = engine.pairs(dx, dy, dz, p, lbvh=lbvh)
    out = engine.compute(dx, dy, dz, p, lbvh=lbvh)
    engine.torch.cuda.synchronize()
    rows_o, cols_o = po.pairs_cellgrid(q, x, y, z)
    assert np.array_equal(rows.cpu().numpy().astype(np.uint64), rows_o)
    assert np.array_equal(cols.cpu().numpy().view(np.uint32), cols_o)
    if n <= 2000:
        rows_b, cols_b = po.pairs_brute(q, x, y, z)
        assert np.array_equal(rows_b, rows_o) and np.array_equal(cols_b, cols_o)
    ref = po.forces_f64(q, x, y, z)
    res = ljamd.read_results(out["result"])[0]
    assert int(res["n_pairs"]) == ref["n_pairs"]
    if not lbvh:
        assert int(res["n_candidates"]) == ref["n_candidates"]
    if n == 0:
        return
    f = [out[k][:n].cpu().numpy() for k in ("fx", "fy", "fz")]
    fin = (np.abs(ref["fx"]) < 3e38) & (np.abs(ref["fy"]) < 3e38) & (np.abs(ref["fz"]) < 3e38)
    if fin.any():
        err = po.rel_errors([a[fin] for a in f], (ref["fx"][fin], ref["fy"][fin], ref["fz"][fin]))["rel"]
        assert err.max() <= 1e-5, desc
    if np.isfinite(ref["e_total"]) and ref["n_pairs"] > 0:
        assert float(res["e_total"]) == pytest.approx(ref["e_total"], rel=1e-9, abs=1e-300), desc


def test_random_batch_matches_oracle(engine):
    """Eight random systems (sizes 0..3,000, their own boxes and cutoffs; a batch shares
    one predicate and precision) in one batched call: each system's pair count and
    candidate count equal the oracle's, its forces are within 1e-5 of FP64."""
    torch = engine.torch
    systems, xs, ys, zs, refs = [], [], [], [], []
    first = 0
    for s in range(8):
        x, y, z, p, _, desc = _config(100 + s)
        x, y, z = x[:3000], y[:3000], z[:3000]
        p.predicate = ljamd.PRED["custom"]
        p.refine, p.refine_x, p.block, p.depth = 1, 2, 256, 1
        sy = ljamd.LJSystem()
        sy.first, sy.n, sy.params = first, x.size, p
        first += x.size
        systems.append(sy)
        xs.append(x), ys.append(y), zs.append(z)
        refs.append(po.forces_f64(po.from_lj(p), x, y, z))
    X, Y, Z = np.concatenate(xs), np.concatenate(ys), np.concatenate(zs)
    print(f"[random batch] n={X.size}", flush=True)
    dx, dy, dz = engine.to_device(X, Y, Z)
    out = engine.compute_batched(systems, dx, dy, dz)
    torch.cuda.synchronize()
    rb = ljamd.read_results(out["result"], len(systems))
    for s, (sy, ref) in enumerate(zip(systems, refs)):
        assert int(rb["n_pairs"][s]) == ref["n_pairs"], s
        assert int(rb["n_candidates"][s]) == ref["n_candidates"], s
        a, b = sy.first, sy.first + sy.n
        if sy.n == 0:
            continue
        f = [out[k][a:b].cpu().numpy() for k in ("fx", "fy", "fz")]
        fin = (np.abs(ref["fx"]) < 3e38) & (np.abs(ref["fy"]) < 3e38) & (np.abs(ref["fz"]) < 3e38)
        if fin.any():
            err = po.rel_errors([v[fin] for v in f], (ref["fx"][fin], ref["fy"][fin], ref["fz"][fin]))["rel"]
            assert err.max() <= 1e-5, s
