"""lj_compute_sorted (the reference's data flow: particles moved into cell order each
step, MoveParticlesToNewCells cudaLJ.cu:441-471, outputs in that order) against
lj_compute on the same input, through the C ABI.

Bars: positions left in the oracle's stable cell order (x_out = x[perm], perm = the
oracle's permutation), outputs bit for bit those of lj_compute permuted by perm, the
same counts and the same bits of the energy total; a second call on the sorted
positions keeps their order (perm = identity) and gives the same outputs."""
import numpy as np
import pytest

import ljamd
import pyoracle as po

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "c1_uniform_100k":
        return ljamd.generate("uniform", 1234, 100_000), ljamd.make_params(0.04)
    if name == "c2_surface_200k":
        return ljamd.generate("surface", 1235, 200_000), ljamd.make_params(0.01, box_lo=(-0.5,) * 3, box_hi=(1.5,) * 3)
    if name == "c5_clusters_64k":
        return ljamd.generate("clusters", 1238, 65_536, [16, 0.02, 0.1, 0.9]), ljamd.make_params(0.005)
    if name == "uniform_3m":  # outputs beyond the L2s: the non-temporal store instance
        return ljamd.generate("uniform", 77, 3_000_000), ljamd.make_params(0.012)
    if name == "edges":
        g = np.array([0.0, 0.04, 0.08, np.nextafter(np.float32(1.0), np.float32(0)), 1.0, 1.3, -0.2, 0.5, 0.5],
                     np.float32)
        xs, ys, zs = np.meshgrid(g, g, g[:4], indexing="ij")
        return (xs.ravel(), ys.ravel(), zs.ravel()), ljamd.make_params(0.04)
    if name == "single":
        return (np.array([0.5], np.float32),) * 3, ljamd.make_params(0.1)
    if name == "empty":
        e = np.zeros(0, np.float32)
        return (e, e, e), ljamd.make_params(0.1)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["c1_uniform_100k", "c2_surface_200k", "c5_clusters_64k", "uniform_3m", "edges",
                                  "single", "empty"])
def test_compute_sorted_equals_compute_permuted(engine, name):
    (x, y, z), p = _case(name)
    n = x.size
    print(f"[sorted] {name} n={n}", flush=True)
    torch = engine.torch
    _, perm_ref, _ = po.bin(po.from_lj(p), x, y, z)
    dx, dy, dz = engine.to_device(x, y, z)
    ref = engine.compute(dx, dy, dz, p)
    sx, sy, sz = (t.clone() for t in (dx, dy, dz))
    out = engine.compute_sorted(sx, sy, sz, p)
    torch.cuda.synchronize()
    r0 = ljamd.read_results(ref["result"])[0]
    r1 = ljamd.read_results(out["result"])[0]
    assert r0.tobytes() == r1.tobytes()
    if n == 0:
        return
    perm = out["perm"][:n].cpu().numpy().view(np.uint32)
    assert np.array_equal(perm, perm_ref)
    for got, src in ((sx, x), (sy, y), (sz, z)):
        assert np.array_equal(got.cpu().numpy().view(np.uint32), src[perm].view(np.uint32))
    for k in ("fx", "fy", "fz", "e"):
        a = out[k][:n].cpu().numpy().view(np.uint32)
        b = ref[k][:n].cpu().numpy()[perm].view(np.uint32)
        assert np.array_equal(a, b), k
    # again on the sorted positions: the stable sort keeps them, the outputs repeat
    out2 = engine.compute_sorted(sx, sy, sz, p)
    torch.cuda.synchronize()
    assert np.array_equal(out2["perm"][:n].cpu().numpy(), np.arange(n, dtype=np.int32))
    for k in ("fx", "fy", "fz", "e"):
        assert torch.equal(out2[k][:n].view(torch.int32), out[k][:n].view(torch.int32)), k


@pytest.mark.parametrize("pred,depth", [("cudalj", 1), ("triangle", 8), ("custom", 16)])
def test_compute_sorted_other_predicates_and_queues(engine, pred, depth):
    """The CUDALJ mode (every candidate, itself skipped by slot), a ray model and the LDS
    queue of accepted partners give the same bits in sorted order as lj_compute's,
    permuted."""
    x, y, z = ljamd.generate("uniform", 91, 50_000)
    p = ljamd.make_params(0.05, predicate=pred, depth=depth, block=128)
    if pred == "cudalj":
        p.softening2, p.sigma = 1e-5, 1.0
    n = x.size
    print(f"[sorted] {pred} depth={depth} n={n}", flush=True)
    torch = engine.torch
    dx, dy, dz = engine.to_device(x, y, z)
    ref = engine.compute(dx, dy, dz, p)
    sx, sy, sz = (t.clone() for t in (dx, dy, dz))
    out = engine.compute_sorted(sx, sy, sz, p)
    torch.cuda.synchronize()
    assert ljamd.read_results(ref["result"])[0].tobytes() == ljamd.read_results(out["result"])[0].tobytes()
    perm = out["perm"][:n].cpu().numpy().view(np.uint32)
    for k in ("fx", "fy", "fz", "e"):
        a = out[k][:n].cpu().numpy().view(np.uint32)
        b = ref[k][:n].cpu().numpy()[perm].view(np.uint32)
        assert np.array_equal(a, b), k


def test_staged_outputs_without_energies_and_after_a_larger_call(engine):
    """Outputs beyond the L2s go through a staging array (one record per particle at its
    caller index, then split into fx, fy, fz, e).  Without the per-particle energies
    (e = NULL) the forces and the result record are those of the call with them, and a
    smaller call after a larger one on the same context gets its own outputs (every
    caller index of the smaller call is staged again): equal, permuted, to
    lj_compute_sorted's, whose stores go straight to the sorted slots.  The LBVH call
    shares the staging array and the split."""
    torch = engine.torch
    for n, seed in ((3_000_000, 91), (2_300_000, 92)):
        x, y, z = ljamd.generate("uniform", seed, n)
        p = ljamd.make_params(0.012)
        print(f"[staged] n={n}", flush=True)
        dx, dy, dz = engine.to_device(x, y, z)
        for lbvh in (False, True):
            a = engine.compute(dx, dy, dz, p, lbvh=lbvh)
            b = engine.compute(dx, dy, dz, p, energies=False, lbvh=lbvh)
            torch.cuda.synchronize()
            assert b["e"] is None
            assert a["result"].cpu().numpy().tobytes() == b["result"].cpu().numpy().tobytes()
            for k in ("fx", "fy", "fz"):
                assert np.array_equal(a[k].cpu().numpy().view(np.uint32), b[k].cpu().numpy().view(np.uint32))
            if not lbvh:
                sx, sy, sz = (t.clone() for t in (dx, dy, dz))
                s = engine.compute_sorted(sx, sy, sz, p)
                torch.cuda.synchronize()
                perm = s["perm"][:n].cpu().numpy().view(np.uint32)
                for k in ("fx", "fy", "fz", "e"):
                    assert np.array_equal(a[k].cpu().numpy()[perm].view(np.uint32), s[k].cpu().numpy().view(np.uint32))
        del dx, dy, dz
