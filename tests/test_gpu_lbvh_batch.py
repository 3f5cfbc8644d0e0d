"""GPU parity of the LBVH path and of the batched path.

LBVH: neighbour lists identical (bit-exact CSR) to the cell grid and the O(N^2)
oracle; forces within 1e-5 of FP64.  Batched: every system identical to its own
single-system call (same sorted order and sums, so the same bits) and to the
oracle's counts and FP64 totals; a captured graph replays to the eager result.
"""
import numpy as np
import pytest

import ljamd
import pyoracle as po
from ljamd import workloads

pytestmark = pytest.mark.gpu


def _lbvh_case(name):
    if name == "uniform_100k":
        x, y, z = ljamd.generate("uniform", 1234, 100_000)
        return (x, y, z), ljamd.make_params(0.04)
    if name == "clusters_256k":
        x, y, z = ljamd.generate("clusters", 1238, 262_144, [16, 0.02, 0.1, 0.9])
        return (x, y, z), ljamd.make_params(0.005)
    if name == "surface_100k":
        x, y, z = ljamd.generate("surface", 1235, 100_000)
        return (x, y, z), ljamd.make_params(0.02, box_lo=(-0.5,) * 3, box_hi=(1.5,) * 3)
    if name == "duplicates":
        # many identical Morton codes and identical positions (d2 = 0 pairs rejected)
        g = np.array([0.2, 0.2, 0.2001, 0.5, 0.5, 0.5, 0.90], np.float32)
        xs, ys, zs = np.meshgrid(g, g, g, indexing="ij")
        return (xs.ravel(), ys.ravel(), zs.ravel()), ljamd.make_params(0.05)
    if name == "two":
        return (np.array([0.1, 0.12], np.float32), np.array([0.5, 0.5], np.float32), np.array([0.5, 0.5], np.float32)), ljamd.make_params(0.05)
    if name == "single":
        return (np.array([0.3], np.float32),) * 3, ljamd.make_params(0.05)
    raise KeyError(name)


LBVH_CASES = ["uniform_100k", "clusters_256k", "surface_100k", "duplicates", "two", "single"]


@pytest.mark.parametrize("name", LBVH_CASES)
def test_lbvh_pairs_equal_cellgrid_and_oracle(engine, name):
    (x, y, z), p = _lbvh_case(name)
    print(f"[lbvh pairs] {name} n={x.size}", flush=True)
    dx, dy, dz = engine.to_device(x, y, z)
    rows_l, cols_l = engine.pairs(dx, dy, dz, p, lbvh=True)
    rows_c, cols_c = engine.pairs(dx, dy, dz, p)
    engine.torch.cuda.synchronize()
    assert np.array_equal(rows_l.cpu().numpy(), rows_c.cpu().numpy())
    assert np.array_equal(cols_l.cpu().numpy(), cols_c.cpu().numpy())
    rows_o, cols_o = po.pairs_cellgrid(po.from_lj(p), x, y, z)
    assert np.array_equal(rows_l.cpu().numpy().astype(np.uint64), rows_o)
    assert np.array_equal(cols_l.cpu().numpy().view(np.uint32), cols_o)
    if x.size <= 4000:
        rows_b, cols_b = po.pairs_brute(po.from_lj(p), x, y, z)
        assert np.array_equal(rows_b, rows_o) and np.array_equal(cols_b, cols_o)


@pytest.mark.parametrize("name", ["uniform_100k", "clusters_256k", "surface_100k", "duplicates"])
def test_lbvh_forces_vs_fp64(engine, name):
    (x, y, z), p = _lbvh_case(name)
    print(f"[lbvh force] {name} n={x.size}", flush=True)
    dx, dy, dz = engine.to_device(x, y, z)
    out = engine.compute(dx, dy, dz, p, lbvh=True)
    engine.torch.cuda.synchronize()
    res = ljamd.read_results(out["result"])[0]
    ref = po.forces_f64(po.from_lj(p), x, y, z)
    assert int(res["n_pairs"]) == ref["n_pairs"]
    f = [out[k][: x.size].cpu().numpy() for k in ("fx", "fy", "fz")]
    fin = (np.abs(ref["fx"]) < 3e38) & (np.abs(ref["fy"]) < 3e38) & (np.abs(ref["fz"]) < 3e38)
    err = po.rel_errors([a[fin] for a in f], (ref["fx"][fin], ref["fy"][fin], ref["fz"][fin]))["rel"]
    assert err.max() <= 1e-5
    if np.all(fin):
        assert res["e_total"] == pytest.approx(ref["e_total"], rel=1e-9)


def test_lbvh_rejects_cudalj(engine):
    (x, y, z), p = _lbvh_case("two")
    p.predicate = ljamd.PRED["cudalj"]
    dx, dy, dz = engine.to_device(x, y, z)
    with pytest.raises(ljamd.LJError):
        engine.compute(dx, dy, dz, p, lbvh=True)


def test_lbvh_walk_guard_marks_results(engine):
    """With every walk capped at one step (LJ_FLAG_WALK_GUARD_TEST, set from the host)
    the step guard trips: the force call marks its results (n_pairs = UINT64_MAX, NaN
    totals) and the pair-list call returns LJ_ERR_OVERFLOW.  The next normal call on the
    same context is unaffected."""
    (x, y, z), p = _lbvh_case("clusters_256k")
    dx, dy, dz = engine.to_device(x, y, z)
    p.flags = ljamd.FLAG_WALK_GUARD_TEST
    print("[lbvh guard] forced", flush=True)
    out = engine.compute(dx, dy, dz, p, lbvh=True)
    engine.torch.cuda.synchronize()
    res = ljamd.read_results(out["result"])[0]
    assert int(res["n_pairs"]) == 2**64 - 1
    assert np.isnan(res["e_total"]) and np.isnan(res["virial"])
    with pytest.raises(ljamd.LJError) as ei:
        engine.pairs(dx, dy, dz, p, lbvh=True)
    assert ei.value.code == -5
    p.flags = 0
    out = engine.compute(dx, dy, dz, p, lbvh=True)
    engine.torch.cuda.synchronize()
    res = ljamd.read_results(out["result"])[0]
    ref = po.forces_f64(po.from_lj(p), x, y, z)
    assert int(res["n_pairs"]) == ref["n_pairs"] and np.isfinite(res["e_total"])


@pytest.mark.parametrize("name", ["clusters_256k", "surface_100k"])
def test_lbvh_deterministic_bits(engine, name):
    """Two LBVH calls on the same input give the same bits: forces, energies, the
    result record and the pair lists (the walk order is fixed by the tree, the totals
    are reduced in a fixed order, the counts are integers)."""
    (x, y, z), p = _lbvh_case(name)
    print(f"[lbvh determinism] {name} n={x.size}", flush=True)
    dx, dy, dz = engine.to_device(x, y, z)
    got = []
    for _ in range(2):
        out = engine.compute(dx, dy, dz, p, lbvh=True)
        rows, cols = engine.pairs(dx, dy, dz, p, lbvh=True)
        engine.torch.cuda.synchronize()
        got.append(([out[k][: x.size].cpu().numpy().view(np.uint32) for k in ("fx", "fy", "fz", "e")],
                    out["result"].cpu().numpy().tobytes(), rows.cpu().numpy(), cols.cpu().numpy()))
    (f0, r0, rows0, cols0), (f1, r1, rows1, cols1) = got
    for a, b in zip(f0, f1):
        assert np.array_equal(a, b)
    assert r0 == r1
    assert np.array_equal(rows0, rows1) and np.array_equal(cols0, cols1)


def _small_batch(nsys=48):
    return workloads.batch_c4(nsys=nsys, nmin=64, nmax=4096, seed=77)


@pytest.mark.parametrize("refine,refine_x", [(1, 1), (2, 2), (4, 4), (1, 8)])
def test_batched_equals_single_calls_and_oracle(engine, refine, refine_x):
    """One launch over all systems gives each system the bits of its own call with
    the same kernel shape, and the oracle's counts, totals and forces."""
    systems, X, Y, Z = _small_batch()
    for sy in systems:
        sy.params.refine = refine
        sy.params.refine_x = refine_x
    print(f"[batched] refine={refine} refine_x={refine_x}", flush=True)
    dx, dy, dz = engine.to_device(X, Y, Z)
    out = engine.compute_batched(systems, dx, dy, dz)
    engine.torch.cuda.synchronize()
    res = ljamd.read_results(out["result"], len(systems))
    fb = [out[k][: X.size].cpu().numpy() for k in ("fx", "fy", "fz", "e")]
    for s, sy in enumerate(systems):
        a, b = sy.first, sy.first + sy.n
        single = engine.compute(dx[a:b], dy[a:b], dz[a:b], sy.params)
        engine.torch.cuda.synchronize()
        r1 = ljamd.read_results(single["result"])[0]
        for k, arr in zip(("fx", "fy", "fz", "e"), fb):
            assert np.array_equal(single[k][: sy.n].cpu().numpy().view(np.uint32), arr[a:b].view(np.uint32)), (s, k)
        assert int(r1["n_pairs"]) == int(res[s]["n_pairs"])
        assert int(r1["n_candidates"]) == int(res[s]["n_candidates"])
        ref = po.forces_f64(po.from_lj(sy.params), X[a:b], Y[a:b], Z[a:b])
        assert int(res[s]["n_pairs"]) == ref["n_pairs"]
        assert int(res[s]["n_candidates"]) == ref["n_candidates"]
        assert res[s]["e_total"] == pytest.approx(ref["e_total"], rel=1e-9)
        assert res[s]["virial"] == pytest.approx(ref["virial"], rel=1e-9)
        err = po.rel_errors([arr[a:b] for arr in fb[:3]], (ref["fx"], ref["fy"], ref["fz"]))["rel"]
        assert err.max() <= 1e-5, s


def test_batched_small_and_empty_systems(engine):
    """Systems smaller than a wave of 64 sorted slots, several in one wave, empty ones,
    and boundaries on and off the wave boundaries: each system gets the bits and
    counts of its own call and the oracle's counts and totals."""
    sizes = [0, 1, 5, 63, 64, 65, 0, 2, 130, 7, 64, 1, 0, 200, 3, 61, 64, 128, 1, 1, 0]
    rng = np.random.default_rng(19)
    systems, parts, first = [], [], 0
    for s, n in enumerate(sizes):
        if n:
            x, y, z = ljamd.generate("uniform", 500 + s, n)
        else:
            x = y = z = np.zeros(0, np.float32)
        parts.append((x, y, z))
        # about 4-20 neighbours per particle
        rc = float(0.35 + 0.25 * rng.random()) if n < 20 else float((3.0 * 12 / (4 * np.pi * n)) ** (1 / 3))
        systems.append(ljamd.LJSystem(first, n, ljamd.make_params(rc)))
        first += n
    X, Y, Z = (np.concatenate([q[i] for q in parts]).astype(np.float32) for i in range(3))
    print(f"[batched small] systems={len(sizes)} n={X.size}", flush=True)
    dx, dy, dz = engine.to_device(X, Y, Z)
    out = engine.compute_batched(systems, dx, dy, dz)
    engine.torch.cuda.synchronize()
    res = ljamd.read_results(out["result"], len(systems))
    fb = [out[k][: X.size].cpu().numpy() for k in ("fx", "fy", "fz", "e")]
    for s, sy in enumerate(systems):
        a, b = sy.first, sy.first + sy.n
        ref = po.forces_f64(po.from_lj(sy.params), X[a:b], Y[a:b], Z[a:b])
        assert int(res[s]["n_pairs"]) == ref["n_pairs"], s
        assert int(res[s]["n_candidates"]) == ref["n_candidates"], s
        assert res[s]["e_total"] == pytest.approx(ref["e_total"], rel=1e-9, abs=1e-12), s
        assert res[s]["virial"] == pytest.approx(ref["virial"], rel=1e-9, abs=1e-12), s
        if sy.n == 0:
            continue
        single = engine.compute(dx[a:b], dy[a:b], dz[a:b], sy.params)
        engine.torch.cuda.synchronize()
        r1 = ljamd.read_results(single["result"])[0]
        assert int(r1["n_pairs"]) == int(res[s]["n_pairs"]), s
        for k, arr in zip(("fx", "fy", "fz", "e"), fb):
            assert np.array_equal(single[k][: sy.n].cpu().numpy().view(np.uint32), arr[a:b].view(np.uint32)), (s, k)


def test_batched_autotune(engine):
    """lj_autotune_batched writes one valid shape into every system; the tuned batch
    finds the same pairs as the default one."""
    systems, X, Y, Z = _small_batch(64)
    dx, dy, dz = engine.to_device(X, Y, Z)
    out = engine.compute_batched(systems, dx, dy, dz)
    engine.torch.cuda.synchronize()
    res0 = ljamd.read_results(out["result"], len(systems)).copy()
    print("[batched autotune]", flush=True)
    t = engine.autotune_batched(systems, dx, dy, dz, reps=2)
    assert t["refine"] in (1, 2, 4) and t["block"] in (64, 128, 256) and t["depth"] in (1, 8, 16, 32)
    assert all((sy.params.refine, sy.params.refine_x, sy.params.block, sy.params.depth) ==
               (t["refine"], t["refine_x"], t["block"], t["depth"]) for sy in systems)
    out = engine.compute_batched(systems, dx, dy, dz)
    engine.torch.cuda.synchronize()
    res1 = ljamd.read_results(out["result"], len(systems))
    assert np.array_equal(res0["n_pairs"], res1["n_pairs"])
    assert np.array_equal(res0["n_candidates"], res1["n_candidates"])
    assert np.allclose(res0["e_total"], res1["e_total"], rtol=1e-9, atol=0)


def test_batched_rejects_mixed_shapes(engine):
    systems, X, Y, Z = _small_batch(4)
    systems[1].params.depth = 8
    dx, dy, dz = engine.to_device(X, Y, Z)
    with pytest.raises(ljamd.LJError):
        engine.compute_batched(systems, dx, dy, dz)


def test_batched_graph_replays_match_eager(engine):
    torch = engine.torch
    systems, X, Y, Z = _small_batch(32)
    dx, dy, dz = engine.to_device(X, Y, Z)
    out = engine.compute_batched(systems, dx, dy, dz)
    torch.cuda.synchronize()
    eager = {k: out[k].clone() for k in ("fx", "fy", "fz", "e", "result")}
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        engine.compute_batched(systems, dx, dy, dz, out=out)
    torch.cuda.current_stream().wait_stream(s)
    with torch.cuda.graph(g):
        engine.compute_batched(systems, dx, dy, dz, out=out)
    for rep in range(3):
        for k in ("fx", "fy", "fz", "e", "result"):
            out[k].zero_()
        engine.compute(dx[:500], dy[:500], dz[:500], ljamd.make_params(0.1))  # unrelated eager work
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in ("fx", "fy", "fz", "e", "result"):
            assert torch.equal(out[k], eager[k]), (rep, k)


def test_workspace_generation_marks_stale_graphs():
    """A graph captured on a context keeps its workspace addresses.  The generation is
    unchanged by calls that fit the workspace (replays stay valid and match the eager
    result) and changes as soon as a larger call reallocates it, which is the caller's
    signal that the captured graph must not be replayed.  (A context of its own, so that
    earlier tests have not grown its workspace already.)"""
    engine = ljamd.Engine(0)
    torch = engine.torch
    systems, X, Y, Z = _small_batch(16)
    dx, dy, dz = engine.to_device(X, Y, Z)
    out = engine.compute_batched(systems, dx, dy, dz)
    engine.compute(dx[:300], dy[:300], dz[:300], ljamd.make_params(0.1))  # grows what this call needs
    torch.cuda.synchronize()
    eager = out["fx"].clone()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        engine.compute_batched(systems, dx, dy, dz, out=out)
    torch.cuda.current_stream().wait_stream(s)
    with torch.cuda.graph(g):
        engine.compute_batched(systems, dx, dy, dz, out=out)
    gen = engine.workspace_generation()
    engine.compute(dx[:300], dy[:300], dz[:300], ljamd.make_params(0.1))  # fits: no reallocation
    torch.cuda.synchronize()
    assert engine.workspace_generation() == gen
    out["fx"].zero_()
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(out["fx"], eager)
    big = 4 * X.size + 100_000
    bx, by, bz = engine.to_device(*ljamd.generate("uniform", 99, big))
    engine.compute(bx, by, bz, ljamd.make_params(0.02))  # larger: the workspace grows
    torch.cuda.synchronize()
    assert engine.workspace_generation() > gen
    del g
    engine.close()


def test_allocation_failure_reports_oom(monkeypatch):
    """With device allocations made to fail (LJ_TEST_FAIL_ALLOC_AFTER, read when a
    context is created) a call that must grow its workspace returns LJ_ERR_OOM and
    names the allocation; a context created without the variable works again."""
    x, y, z = ljamd.generate("uniform", 5, 20_000)
    monkeypatch.setenv("LJ_TEST_FAIL_ALLOC_AFTER", "2")
    eng = ljamd.Engine(0)
    try:
        dx, dy, dz = eng.to_device(x, y, z)
        with pytest.raises(ljamd.LJError) as ei:
            eng.compute(dx, dy, dz, ljamd.make_params(0.05))
        assert ei.value.code == -2 and "grow" in str(ei.value)
    finally:
        eng.close()
        monkeypatch.delenv("LJ_TEST_FAIL_ALLOC_AFTER")
        eng = ljamd.Engine(0)  # the injection ends with the next context creation
    out = eng.compute(dx, dy, dz, ljamd.make_params(0.05))
    eng.torch.cuda.synchronize()
    res = ljamd.read_results(out["result"])[0]
    assert int(res["n_pairs"]) == po.forces_f64(po.params(0.05), x, y, z)["n_pairs"]
    eng.close()


def test_batched_rejects_unpacked_systems(engine):
    systems, X, Y, Z = _small_batch(4)
    systems[2].first += 1
    dx, dy, dz = engine.to_device(X, Y, Z)
    with pytest.raises(ljamd.LJError):
        engine.compute_batched(systems, dx, dy, dz)
