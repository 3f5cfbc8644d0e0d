"""Parity at the benchmark sizes through properties that do not need an FP64 run over
every particle: bit-exact binning against the oracle at 16.7M particles, the closed-form
candidate count, Newton's third law and the energy total, LBVH = cell grid on the 1M
sphere and the 4.2M clusters, every one of the 4,096 batched systems equal to its own
call, and binning plus subsampled forces at 2^27 particles.  (Force errors against FP64
at the benchmark sizes: test_c3_precision_subsample.)
"""
import numpy as np
import pytest

import ljamd
import pyoracle as po
from ljamd import workloads

pytestmark = pytest.mark.gpu


def _forces(engine, dx, dy, dz, p, lbvh=False):
    out = engine.compute(dx, dy, dz, p, lbvh=lbvh)
    engine.torch.cuda.synchronize()
    n = dx.numel()
    res = ljamd.read_results(out["result"])[0]
    return [out[k][:n].cpu().numpy() for k in ("fx", "fy", "fz", "e")], res


def test_c3_bin_bit_exact_full(engine):
    """16,777,216 uniform particles, rc = 0.01 (10^6 cells): cell offsets, permutation
    and candidate count bit-exact against the oracle's stable counting sort."""
    n = 16_777_216
    x, y, z = ljamd.generate("uniform", 1236, n)
    p = ljamd.make_params(0.01)
    print(f"[full bin] c3 n={n}", flush=True)
    q = po.from_lj(p)
    off_ref, perm_ref, _ = po.bin(q, x, y, z)
    dx, dy, dz = engine.to_device(x, y, z)
    off, perm, cand = engine.bin(dx, dy, dz, p)
    engine.torch.cuda.synchronize()
    assert np.array_equal(off.cpu().numpy().view(np.uint32), off_ref)
    assert np.array_equal(perm[:n].cpu().numpy().view(np.uint32), perm_ref)
    assert int(cand.cpu()[0]) == po.candidates(q, off_ref)


def test_c3_newton_and_energy_total_full(engine):
    """Every pair adds equal and opposite terms, so sum_i F_i = 0 up to the FP32
    rounding of each F_i; e_total is half the sum of the per-particle energies;
    ordered pairs come in couples; two calls give the same bits."""
    n = 16_777_216
    x, y, z = ljamd.generate("uniform", 1236, n)
    p = ljamd.make_params(0.01)
    print(f"[full newton] c3 n={n}", flush=True)
    dx, dy, dz = engine.to_device(x, y, z)
    (fx, fy, fz, e), res = _forces(engine, dx, dy, dz, p)
    for f in (fx, fy, fz):
        f64 = f.astype(np.float64)
        assert np.all(np.isfinite(f64))
        assert abs(f64.sum()) <= 4 * 2.0**-24 * np.abs(f64).sum()
    assert res["e_total"] == pytest.approx(0.5 * e.astype(np.float64).sum(), rel=1e-6)
    assert int(res["n_pairs"]) % 2 == 0
    (gx, gy, gz, g_e), res2 = _forces(engine, dx, dy, dz, p)
    for a, b in ((fx, gx), (fy, gy), (fz, gz), (e, g_e)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert res.tobytes() == res2.tobytes()


@pytest.mark.parametrize("name", ["c2", "c5"])
def test_lbvh_equals_cellgrid_full(engine, name):
    """C2 (1,000,000 points on a sphere, rc = 0.01) and C5 (4,194,304 particles in 256
    clusters, rc = 0.005): the LBVH finds the cell grid's pairs and, summing the same
    terms in another order in FP64, the same forces to FP32 rounding."""
    w = workloads.single(name)
    n = w.x.size
    print(f"[full lbvh] {name} n={n}", flush=True)
    dx, dy, dz = engine.to_device(w.x, w.y, w.z)
    (cfx, cfy, cfz, ce), cres = _forces(engine, dx, dy, dz, w.params)
    (lfx, lfy, lfz, le), lres = _forces(engine, dx, dy, dz, w.params, lbvh=True)
    assert int(lres["n_pairs"]) == int(cres["n_pairs"])
    if name == "c2":
        # uniform density on the unit sphere: P / N = (N - 1) rc^2 / 4
        assert int(cres["n_pairs"]) == pytest.approx((n - 1) * 0.01**2 / 4 * n, rel=0.02)
    assert lres["e_total"] == pytest.approx(cres["e_total"], rel=1e-9)
    assert lres["virial"] == pytest.approx(cres["virial"], rel=1e-9)
    # the closest pairs' forces exceed the FP32 range: the same particles overflow to
    # the same infinities on both paths; the rest agree to FP32 rounding
    fin = np.isfinite(cfx) & np.isfinite(cfy) & np.isfinite(cfz)
    print(f"[full lbvh] {name} overflowing force rows: {(~fin).sum()}", flush=True)
    assert np.array_equal(np.isfinite(lfx) & np.isfinite(lfy) & np.isfinite(lfz), fin)
    for a, b in ((lfx, cfx), (lfy, cfy), (lfz, cfz)):
        assert np.array_equal(a[~fin], b[~fin])
    err = po.rel_errors((lfx[fin], lfy[fin], lfz[fin]), (cfx[fin], cfy[fin], cfz[fin]))["rel"]
    assert err.max() <= 1e-6, (err.max(), int(np.argmax(err)))
    efin = np.isfinite(ce)
    assert np.array_equal(np.isfinite(le), efin) and np.array_equal(le[~efin], ce[~efin])
    assert np.max(np.abs(le[efin] - ce[efin]) / np.maximum(np.abs(ce[efin]), 1e-30)) <= 1e-6


@pytest.mark.parametrize("name", ["c2", "c5"])
def test_full_pair_lists_cellgrid_equals_lbvh(engine, name):
    """C2 (1,000,000 points on a sphere) and C5 (4,194,304 clustered particles, about
    116M ordered pairs): the LBVH's CSR neighbour lists equal the cell grid's bit for
    bit over every row.  C2's lists equal the oracle's in full, and its pair and
    candidate counts equal the FP64 oracle's; a 4,096-row subsample of C5 equals the
    oracle's rows.  The LBVH force call's walks all finish inside their step guard."""
    w = workloads.single(name)
    n = w.x.size
    print(f"[full lists] {name} n={n}", flush=True)
    q = po.from_lj(w.params)
    dx, dy, dz = engine.to_device(w.x, w.y, w.z)
    rows_c, cols_c = engine.pairs(dx, dy, dz, w.params)
    rows_l, cols_l = engine.pairs(dx, dy, dz, w.params, lbvh=True)
    engine.torch.cuda.synchronize()
    rc_, cc_ = rows_c.cpu().numpy(), cols_c.cpu().numpy()
    assert np.array_equal(rc_, rows_l.cpu().numpy())
    assert np.array_equal(cc_, cols_l.cpu().numpy())
    del rows_l, cols_l
    print(f"[full lists] {name} pairs={cc_.size}", flush=True)
    (_, _, _, _), lres = _forces(engine, dx, dy, dz, w.params, lbvh=True)
    assert int(lres["n_pairs"]) == cc_.size and np.isfinite(lres["e_total"])
    if name == "c2":
        rows_o, cols_o = po.pairs_cellgrid(q, w.x, w.y, w.z)
        assert np.array_equal(rc_.astype(np.uint64), rows_o)
        assert np.array_equal(cc_.view(np.uint32), cols_o)
        ref = po.forces_f64(q, w.x, w.y, w.z, nthreads=16)
        (_, _, _, _), cres = _forces(engine, dx, dy, dz, w.params)
        assert int(cres["n_pairs"]) == ref["n_pairs"] == cc_.size
        assert int(cres["n_candidates"]) == ref["n_candidates"]
    else:
        sub = np.arange(0, n, n // 4096, dtype=np.int64)
        rows_o, cols_o = po.pairs_rows(q, w.x, w.y, w.z, sub, nthreads=16)
        got = [cc_[rc_[i]:rc_[i + 1]] for i in sub]
        assert np.array_equal(np.array([g.size for g in got], np.uint64), np.diff(rows_o))
        assert np.array_equal(np.concatenate(got).view(np.uint32), cols_o)


def test_c4_batched_equals_single_calls_full(engine):
    """All 4,096 systems of the batched workload (46.8M particles): one batched launch
    gives every system the bits and counts of its own call with the same shape.  With
    x-cells of cell_edge / 4 the batch spans 64M cells, a 26-bit key sorted in three
    9-bit passes (each single call sorts its own small key space in 8-bit passes)."""
    torch = engine.torch
    systems, X, Y, Z = workloads.batch_c4()
    for sy in systems:
        sy.params.refine, sy.params.refine_x, sy.params.block, sy.params.depth = 1, 4, 256, 1
    n = X.size
    print(f"[full batched] c4 systems={len(systems)} n={n}", flush=True)
    dx, dy, dz = engine.to_device(X, Y, Z)
    out = engine.compute_batched(systems, dx, dy, dz)
    single = {k: engine.empty_f32(n) for k in ("fx", "fy", "fz", "e")}
    one = engine.empty_result()
    results = torch.empty((len(systems), one.numel()), dtype=one.dtype, device=one.device)
    for s, sy in enumerate(systems):
        a, b = sy.first, sy.first + sy.n
        o = {k: single[k][a:b] for k in ("fx", "fy", "fz", "e")}
        o["result"] = results[s]
        engine.compute(dx[a:b], dy[a:b], dz[a:b], sy.params, out=o)
    torch.cuda.synchronize()
    rb = ljamd.read_results(out["result"], len(systems))
    rs = np.frombuffer(results.cpu().numpy().tobytes(), dtype=rb.dtype, count=len(systems))
    assert np.array_equal(rb["n_pairs"], rs["n_pairs"])
    assert np.array_equal(rb["n_candidates"], rs["n_candidates"])
    np.testing.assert_allclose(rb["e_total"], rs["e_total"], rtol=1e-12, atol=0)
    for k in ("fx", "fy", "fz", "e"):
        assert torch.equal(out[k][:n].view(torch.int32), single[k].view(torch.int32)), k


def test_large_n_bin_and_forces(engine):
    """2^27 = 134,217,728 uniform particles, rc = 0.005 (8,000,000 cells, ~9.4e9 pairs):
    eight times the largest benchmark input.  Binning bit-exact and the closed-form
    candidate count against the oracle, forces of a 4,096-particle subsample against
    FP64."""
    n = 1 << 27
    x, y, z = ljamd.generate("uniform", 4242, n)
    p = ljamd.make_params(0.005)
    q = po.from_lj(p)
    print(f"[large] n={n}", flush=True)
    off_ref, perm_ref, _ = po.bin(q, x, y, z)
    dx, dy, dz = engine.to_device(x, y, z)
    off, perm, cand = engine.bin(dx, dy, dz, p)
    engine.torch.cuda.synchronize()
    assert np.array_equal(off.cpu().numpy().view(np.uint32), off_ref)
    assert np.array_equal(perm[:n].cpu().numpy().view(np.uint32), perm_ref)
    assert int(cand.cpu()[0]) == po.candidates(q, off_ref)
    del off, perm
    (fx, fy, fz, e), res = _forces(engine, dx, dy, dz, p)
    sub = np.arange(0, n, n // 4096)
    ref = po.forces_f64(q, x, y, z, subset=sub, nthreads=16)
    fin = (np.abs(ref["fx"]) < 3e38) & (np.abs(ref["fy"]) < 3e38) & (np.abs(ref["fz"]) < 3e38)
    err = po.rel_errors([a[sub][fin] for a in (fx, fy, fz)], (ref["fx"][fin], ref["fy"][fin], ref["fz"][fin]))["rel"]
    print(f"[large] pairs={int(res['n_pairs'])} max rel {err.max():.3e}", flush=True)
    assert err.max() <= 1e-5
    assert int(res["n_pairs"]) % 2 == 0
    assert int(res["n_pairs"]) == pytest.approx(4 / 3 * np.pi * 0.005**3 * n * n, rel=0.02)
