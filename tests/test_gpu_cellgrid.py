"""GPU parity of the cell-grid path (binning, forces/energies, neighbour lists)
against the CPU oracle, through the C ABI.

Bars: bit-exact for cell offsets, permutation, candidate and pair counts and
neighbour lists; relative force error <= 1e-5 against the FP64 oracle for the
default (mixed) precision, the reference's own tolerance (Cuda-version/cudaLJ.cu:246-250).
"""
import numpy as np
import pytest

import ljamd
import pyoracle as po

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5  # relative, per particle, vs FP64


def _case(name):
    if name == "c1_uniform_100k":
        x, y, z = ljamd.generate("uniform", 1234, 100_000)
        return (x, y, z), ljamd.make_params(0.04)
    if name == "c2_surface_200k":
        x, y, z = ljamd.generate("surface", 1235, 200_000)
        return (x, y, z), ljamd.make_params(0.01, box_lo=(-0.5,) * 3, box_hi=(1.5,) * 3)
    if name == "c5_clusters_64k":
        x, y, z = ljamd.generate("clusters", 1238, 65_536, [16, 0.02, 0.1, 0.9])
        return (x, y, z), ljamd.make_params(0.005)
    if name == "tiny_7":
        x, y, z = ljamd.generate("uniform", 3, 7)
        return (x, y, z), ljamd.make_params(0.5)
    if name == "single":
        return (np.array([0.5], np.float32),) * 3, ljamd.make_params(0.1)
    if name == "empty":
        e = np.zeros(0, np.float32)
        return (e, e, e), ljamd.make_params(0.1)
    if name == "edges":
        # exactly on cell faces, the largest float below the box edge (rounds to G),
        # outside the box on both sides, duplicates (d2 = 0), NaN-free
        rc = 0.04
        vals = [0.0, 0.04, 0.08, 0.12, np.nextafter(np.float32(1.0), np.float32(0)), 1.0, 1.3, -0.2, 0.5, 0.5]
        g = np.array(vals, np.float32)
        xs, ys, zs = np.meshgrid(g, g, g[:4], indexing="ij")
        return (xs.ravel(), ys.ravel(), zs.ravel()), ljamd.make_params(rc)
    if name == "one_cell_crowd":
        x, y, z = ljamd.generate("uniform", 11, 3000, [0.30, 0.31])
        return (x, y, z), ljamd.make_params(0.05)
    if name == "gap_classes":
        # one row of 131,072 cells of edge 2^-17 (exact keys): runs of empty cells of
        # every length class of the offsets kernel (up to 8 cells: the lane; 9-256: the
        # wave; longer: 4096-cell pieces), each class boundary on both sides, repeated
        # cells, and a long empty stretch before the first and after the last key
        rc = 2.0**-17
        rng = np.random.default_rng(21)
        cls = [1, 2, 8, 9, 10, 64, 255, 256, 257, 4095, 4096, 4097, 9000]
        wts = np.array([8, 8, 8, 8, 8, 8, 8, 8, 8, 1, 1, 1, 1], float)
        gaps = rng.choice(cls, size=1000, p=wts / wts.sum())
        gaps[rng.random(gaps.size) < 0.2] = 0  # a second particle in the same cell
        cells = 5000 + np.cumsum(gaps)
        cells = cells[cells < (1 << 17) - 6000]
        x = ((cells + 0.5) * rc).astype(np.float32)
        y = np.full(x.size, 0.5 * rc, np.float32)
        return (x, y, y.copy()), ljamd.make_params(rc, box_lo=(0.0, 0.0, 0.0), box_hi=(1.0, rc, rc))
    if name == "sparse_25bit_keys":
        # 300^3 = 27,000,000 cells: a 25-bit key, sorted in passes of 8, 8 and 9 bits
        x, y, z = ljamd.generate("uniform", 13, 200_000)
        return (x, y, z), ljamd.make_params(1.0 / 300)
    if name == "sparse_26bit_keys":
        # 333^3 = 36,926,037 cells: a 26-bit key, sorted in passes of 8, 9 and 9 bits
        x, y, z = ljamd.generate("uniform", 12, 200_000)
        return (x, y, z), ljamd.make_params(0.003)
    if name == "sparse_27bit_keys":
        # 450^3 = 91,125,000 cells: a 27-bit key, sorted in three 9-bit passes
        x, y, z = ljamd.generate("uniform", 14, 200_000)
        return (x, y, z), ljamd.make_params(1.0 / 450)
    raise KeyError(name)


BIN_CASES = ["c1_uniform_100k", "c2_surface_200k", "c5_clusters_64k", "tiny_7", "single", "empty", "edges", "one_cell_crowd",
             "sparse_25bit_keys", "sparse_26bit_keys", "sparse_27bit_keys", "gap_classes"]


@pytest.mark.parametrize("name", BIN_CASES)
def test_bin_bit_exact(engine, name):
    (x, y, z), p = _case(name)
    print(f"[bin] {name} n={x.size}", flush=True)
    q = po.from_lj(p)
    off_ref, perm_ref, _ = po.bin(q, x, y, z)
    dx, dy, dz = engine.to_device(x, y, z)
    off, perm, cand = engine.bin(dx, dy, dz, p)
    engine.torch.cuda.synchronize()
    off = off.cpu().numpy().view(np.uint32)
    perm = perm.cpu().numpy().view(np.uint32)
    assert np.array_equal(off, off_ref)
    assert np.array_equal(perm, perm_ref)
    assert int(cand.cpu()[0]) == po.candidates(q, off_ref) == po.candidates_direct(q, x, y, z, off_ref)


def _forces(engine, x, y, z, p, energies=True):
    dx, dy, dz = engine.to_device(x, y, z)
    out = engine.compute(dx, dy, dz, p, energies=energies)
    engine.torch.cuda.synchronize()
    n = x.size
    res = ljamd.read_results(out["result"])[0]
    f = [out[k][:n].cpu().numpy() for k in ("fx", "fy", "fz")]
    e = out["e"][:n].cpu().numpy() if energies else None
    return f, e, res


FORCE_CASES = ["c1_uniform_100k", "c2_surface_200k", "c5_clusters_64k", "tiny_7", "single", "empty", "edges", "one_cell_crowd",
               "gap_classes"]


@pytest.mark.parametrize("name", ["c1_uniform_100k", "c2_surface_200k", "edges", "single", "empty"])
def test_reorder_matches_oracle(engine, name):
    """lj_reorder (the -us pre-sort of SDK/optixMultiSphereAABB/sort.cu:28-41): the
    positions come back in the oracle's stable cell order, xo = x[perm], with perm the
    oracle's permutation, bit for bit."""
    (x, y, z), p = _case(name)
    n = x.size
    print(f"[reorder] {name} n={n}", flush=True)
    _, perm_ref, _ = po.bin(po.from_lj(p), x, y, z)
    dx, dy, dz = engine.to_device(x, y, z)
    xo, yo, zo, perm = engine.reorder(dx, dy, dz, p)
    engine.torch.cuda.synchronize()
    assert np.array_equal(perm.cpu().numpy().view(np.uint32), perm_ref)
    for got, src in ((xo, x), (yo, y), (zo, z)):
        assert np.array_equal(got.cpu().numpy().view(np.uint32), src[perm_ref].view(np.uint32))


@pytest.mark.parametrize("name", FORCE_CASES)
def test_forces_mixed_vs_fp64(engine, name):
    (x, y, z), p = _case(name)
    print(f"[force] {name} n={x.size}", flush=True)
    q = po.from_lj(p)
    ref = po.forces_f64(q, x, y, z)
    f, e, res = _forces(engine, x, y, z, p)
    assert int(res["n_pairs"]) == ref["n_pairs"]
    assert int(res["n_candidates"]) == ref["n_candidates"]
    if x.size == 0:
        return
    fin = np.isfinite(ref["fx"]) & (np.abs(ref["fx"]) < 3e38) & (np.abs(ref["fy"]) < 3e38) & (np.abs(ref["fz"]) < 3e38)
    err = po.rel_errors([a[fin] for a in f], (ref["fx"][fin], ref["fy"][fin], ref["fz"][fin]))["rel"]
    assert err.max() <= FORCE_TOL, (err.max(), int(np.argmax(err)))
    efin = fin & (np.abs(ref["e"]) < 3e38)
    eerr = np.abs(e[efin] - ref["e"][efin]) / np.maximum(np.abs(ref["e"][efin]), 1e-30)
    assert eerr.max() <= FORCE_TOL
    if np.all(fin):
        assert res["e_total"] == pytest.approx(ref["e_total"], rel=1e-9, abs=1e-12)
        assert res["virial"] == pytest.approx(ref["virial"], rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("refine,refine_x", [(1, 1), (2, 2), (4, 4), (1, 4), (1, 8), (1, 16), (2, 8), (4, 16)])
@pytest.mark.parametrize("name", ["c1_uniform_100k", "c2_surface_200k", "c5_clusters_64k", "edges", "one_cell_crowd", "tiny_7",
                                  "gap_classes"])
def test_refined_grids_same_pairs_and_forces(engine, name, refine, refine_x):
    """The 27-cell grid and the half/quarter-edge grids with stencil trimming visit
    every accepted pair: counts bit-exact, forces within the bar."""
    (x, y, z), p = _case(name)
    p.refine = refine
    p.refine_x = refine_x
    print(f"[refine] {name} k={refine} kx={refine_x}", flush=True)
    ref = po.forces_f64(po.from_lj(p), x, y, z)
    f, e, res = _forces(engine, x, y, z, p)
    assert int(res["n_pairs"]) == ref["n_pairs"]
    assert int(res["n_candidates"]) == ref["n_candidates"]
    fin = (np.abs(ref["fx"]) < 3e38) & (np.abs(ref["fy"]) < 3e38) & (np.abs(ref["fz"]) < 3e38)
    err = po.rel_errors([a[fin] for a in f], (ref["fx"][fin], ref["fy"][fin], ref["fz"][fin]))["rel"]
    assert err.max() <= FORCE_TOL


@pytest.mark.parametrize("name", ["c1_uniform_100k", "c5_clusters_64k", "tiny_7"])
def test_kernel_shapes_same_results(engine, name):
    """Workgroup size and LDS queue depth change neither the pairs nor the order in
    which each particle's partners are summed: outputs are bit-identical across them
    at a fixed refinement."""
    (x, y, z), p = _case(name)
    p.refine = 2
    print(f"[shapes] {name}", flush=True)
    base = None
    for block in (64, 128, 256):
        for depth in (1, 8, 16, 32):
            p.block, p.depth = block, depth
            f, e, res = _forces(engine, x, y, z, p)
            cur = (np.stack(f).view(np.uint32), e.view(np.uint32), int(res["n_pairs"]), int(res["n_candidates"]),
                   float(res["e_total"]))
            if base is None:
                base = cur
                continue
            assert np.array_equal(cur[0], base[0]), (block, depth)
            assert np.array_equal(cur[1], base[1]), (block, depth)
            assert cur[2:4] == base[2:4], (block, depth)
            # the FP64 total is summed per workgroup, so it may differ in the last bits
            assert cur[4] == pytest.approx(base[4], rel=1e-12, abs=1e-300)


def test_autotune_picks_a_valid_shape(engine):
    """lj_autotune times the shapes, writes a valid one into the params, and the tuned
    params give the same pairs and forces within the bar."""
    (x, y, z), p = _case("c1_uniform_100k")
    dx, dy, dz = engine.to_device(x, y, z)
    print("[autotune] c1", flush=True)
    t = engine.autotune(dx, dy, dz, p, reps=2)
    assert t["refine"] in (1, 2, 4) and t["block"] in (64, 128, 256) and t["depth"] in (1, 8, 16, 32)
    assert t["refine_x"] in (1, 2, 4, 8, 16) and t["refine_x"] >= t["refine"]
    assert (p.refine, p.refine_x, p.block, p.depth) == (t["refine"], t["refine_x"], t["block"], t["depth"])
    assert t["n_tried"] == 12 + 11
    assert 0 < t["ms_best"] and 0 < t["ms_default"]
    ref = po.forces_f64(po.from_lj(p), x, y, z)
    f, e, res = _forces(engine, x, y, z, p)
    assert int(res["n_pairs"]) == ref["n_pairs"]
    assert int(res["n_candidates"]) == ref["n_candidates"]
    err = po.rel_errors(f, (ref["fx"], ref["fy"], ref["fz"]))["rel"]
    assert err.max() <= FORCE_TOL


def test_autotune_cudalj_keeps_the_27_cells(engine):
    (x, y, z), p = _case("c1_uniform_100k")
    q = ljamd.make_params(float(p.rc), sigma=1.0, softening2=1e-5, predicate="cudalj")
    dx, dy, dz = engine.to_device(x[:20000], y[:20000], z[:20000])
    t = engine.autotune(dx, dy, dz, q, reps=1)
    assert t["refine"] == 1 and q.refine == 1 and t["refine_x"] == 1
    assert t["n_tried"] == 1 + 11


def test_trimming_at_cell_faces(engine):
    """Pairs at exactly the cutoff across cell faces and corners, for every refinement."""
    rc = np.float32(0.0625)  # cell faces at exact binary fractions
    base = np.array([0.25, 0.25, 0.25], np.float32)
    pts = [base]
    for dxyz in ([rc, 0, 0], [0, rc, 0], [0, 0, -rc], [rc / 2, rc / 2, 0], [-rc * 0.57735, rc * 0.57735, rc * 0.57735],
                 [np.float32(0.0625 * 0.99999), 0, 0], [np.float32(0.0625 * 1.00001), 0, 0]):
        pts.append(base + np.array(dxyz, np.float32))
    pts = np.array(pts, np.float32)
    x, y, z = (np.ascontiguousarray(pts[:, d]) for d in range(3))
    for refine, refine_x in ((1, 1), (2, 2), (4, 4), (1, 8), (2, 16)):
        p = ljamd.make_params(float(rc), refine=refine, refine_x=refine_x)
        rows, cols = po.pairs_brute(po.from_lj(p), x, y, z)
        f, e, res = _forces(engine, x, y, z, p)
        assert int(res["n_pairs"]) == int(rows[-1]), refine


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-4), ("fp32_comp", 1e-4)])
def test_forces_fp32_modes(engine, prec, tol):
    (x, y, z), p = _case("c1_uniform_100k")
    p.precision = ljamd.PREC[prec]
    ref = po.forces_f64(po.from_lj(p), x, y, z)
    f, e, res = _forces(engine, x, y, z, p)
    assert int(res["n_pairs"]) == ref["n_pairs"]
    err = po.rel_errors(f, (ref["fx"], ref["fy"], ref["fz"]), ref["cond"])
    # conditioned error |F - F64| / sum_j |f_ij|: the FP32 pair chain rounds sigma^2,
    # r2 and 1/r2 (v_rcp_f32, 1 ulp), and f ~ r^-14 amplifies those ~7x, so about
    # 35 ulp = 2e-6 per pair; 4e-6 leaves room for the summation
    assert err["cond"].max() <= 4e-6
    assert np.quantile(err["rel"], 0.999) <= tol


@pytest.mark.parametrize("pred", ["custom", "sphere", "triangle"])
def test_compat_predicates(engine, pred):
    x, y, z = ljamd.generate("uniform", 21, 20_000)
    p = ljamd.make_params(0.06, predicate=pred)
    ref = po.forces_f64(po.from_lj(p), x, y, z)
    f, e, res = _forces(engine, x, y, z, p)
    assert int(res["n_pairs"]) == ref["n_pairs"]
    err = po.rel_errors(f, (ref["fx"], ref["fy"], ref["fz"]))["rel"]
    assert err.max() <= FORCE_TOL


def _ray_twins(seed, c):
    """Uniform particles plus partners placed where the ray models of the Sphere and
    Triangle programs count a pair 0, 2 or 4 times: within 2e-5 of a particle on one or
    two of y and z, within c 1e-5 in x, and near-coincident."""
    x, y, z = ljamd.generate("uniform", seed, 20_000, [0.05, 0.95])
    rng = np.random.default_rng(seed)
    base = rng.integers(0, x.size, 3000)
    off = rng.uniform(-0.9 * c, 0.9 * c, (3000, 3))
    kind = np.arange(3000) % 4
    off[kind == 0, 1] = rng.uniform(-2e-5, 2e-5, (kind == 0).sum())
    off[kind == 1, 1:] = rng.uniform(-2e-5, 2e-5, ((kind == 1).sum(), 2))
    off[kind == 2, 0] = rng.uniform(-2e-5 * c, 2e-5 * c, (kind == 2).sum())
    off[kind == 3] = rng.uniform(-3e-5, 3e-5, ((kind == 3).sum(), 3))
    X = np.concatenate([x, (x[base] + off[:, 0]).astype(np.float32)])
    Y = np.concatenate([y, (y[base] + off[:, 1]).astype(np.float32)])
    Z = np.concatenate([z, (z[base] + off[:, 2]).astype(np.float32)])
    return X, Y, Z


@pytest.mark.parametrize("pred", ["sphere", "triangle"])
def test_ray_models_with_degenerate_pairs(engine, pred):
    """The Sphere and Triangle programs' ray-hit rules (SDK/optixMultiSphere/optixLJ.cu:
    77-194, SDK/optixMultiTriangle/optixLJ.cu:82-217) on inputs full of the pairs they
    count other than once: summed multiplicities (n_pairs) equal the oracle's, the
    neighbour lists (partners counted at least once) equal the oracle's bit for bit on
    the cell grid, the refined grid and the LBVH, and forces are within 1e-5 of FP64."""
    c = 0.06
    x, y, z = _ray_twins(31, c)
    p = ljamd.make_params(c, sigma=2e-5, predicate=pred)
    q = po.from_lj(p)
    print(f"[ray model] {pred} n={x.size}", flush=True)
    ref = po.forces_f64(q, x, y, z)
    rows_o, cols_o = po.pairs_cellgrid(q, x, y, z)
    if pred == "triangle":
        assert ref["n_pairs"] > cols_o.size  # some pairs count twice or four times
    else:  # near-coincident pairs that the AABB test accepts are missed by the rays
        assert ref["n_pairs"] < po.forces_f64(po.params(c, sigma=2e-5), x, y, z)["n_pairs"]
    dx, dy, dz = engine.to_device(x, y, z)
    for lbvh, refine in ((False, 0), (False, 2), (True, 0)):
        p.refine = refine
        out = engine.compute(dx, dy, dz, p, lbvh=lbvh)
        engine.torch.cuda.synchronize()
        res = ljamd.read_results(out["result"])[0]
        assert int(res["n_pairs"]) == ref["n_pairs"], (lbvh, refine)
        f = [out[k][: x.size].cpu().numpy() for k in ("fx", "fy", "fz")]
        err = po.rel_errors(f, (ref["fx"], ref["fy"], ref["fz"]))["rel"]
        assert err.max() <= FORCE_TOL, (lbvh, refine, err.max())
        assert res["e_total"] == pytest.approx(ref["e_total"], rel=1e-9)
        rows, cols = engine.pairs(dx, dy, dz, p, lbvh=lbvh)
        engine.torch.cuda.synchronize()
        assert np.array_equal(rows.cpu().numpy().astype(np.uint64), rows_o)
        assert np.array_equal(cols.cpu().numpy().view(np.uint32), cols_o)
    p.refine = 0


def test_cudalj_energy_semantics(engine):
    """Reference kernel semantics: every 27-cell candidate, softened r, sigma=eps=1,
    energies compared with CheckEqual's 1e-5 relative bar (cudaLJ.cu:246-250)."""
    x, y, z = ljamd.generate("uniform", 1234, 4096)
    p = ljamd.make_params(0.125, sigma=1.0, epsilon=1.0, softening2=1e-5, predicate="cudalj")
    e_ref = po.cudalj_energy(po.from_lj(p), x, y, z)
    f, e, res = _forces(engine, x, y, z, p)
    assert int(res["n_pairs"]) == int(res["n_candidates"])
    rel = np.abs(e - e_ref) / np.abs(e_ref)
    assert rel.max() <= 1e-5


@pytest.mark.parametrize("name", ["c1_uniform_100k", "c5_clusters_64k", "tiny_7", "edges", "one_cell_crowd"])
def test_pairs_bit_exact(engine, name):
    (x, y, z), p = _case(name)
    q = po.from_lj(p)
    rows_ref, cols_ref = po.pairs_cellgrid(q, x, y, z)
    dx, dy, dz = engine.to_device(x, y, z)
    rows, cols = engine.pairs(dx, dy, dz, p)
    engine.torch.cuda.synchronize()
    assert np.array_equal(rows.cpu().numpy().astype(np.uint64), rows_ref)
    assert np.array_equal(cols.cpu().numpy().view(np.uint32), cols_ref)


def test_pairs_capacity_error(engine):
    (x, y, z), p = _case("tiny_7")
    dx, dy, dz = engine.to_device(x, y, z)
    rows_ref, _ = po.pairs_brute(po.from_lj(p), x, y, z)
    need = int(rows_ref[-1])
    assert need > 1
    with pytest.raises(ljamd.LJError) as ei:
        engine.pairs(dx, dy, dz, p, capacity=need - 1)
    assert ei.value.code == -3


def test_deterministic_bits(engine):
    (x, y, z), p = _case("c1_uniform_100k")
    a, ea, ra = _forces(engine, x, y, z, p)
    b, eb, rb = _forces(engine, x, y, z, p)
    for u, v in zip(a, b):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    assert ra.tobytes() == rb.tobytes()


@pytest.mark.parametrize("name", ["c1_uniform_100k", "c3_sub_big"])
def test_compute_graph_replays_match_eager(engine, name):
    """lj_compute captured in a graph replays to the eager bits, with unrelated eager
    work between the replays (small outputs stored directly; large ones staged and
    split)."""
    torch = engine.torch
    if name == "c3_sub_big":  # outputs beyond the L2s: staged, then split
        x, y, z = ljamd.generate("uniform", 1236, 2_500_000)
        p = ljamd.make_params(0.01)
    else:
        (x, y, z), p = _case(name)
    print(f"[graph] {name} n={x.size}", flush=True)
    dx, dy, dz = engine.to_device(x, y, z)
    out = engine.compute(dx, dy, dz, p)
    torch.cuda.synchronize()
    eager = {k: out[k].clone() for k in ("fx", "fy", "fz", "e", "result")}
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        engine.compute(dx, dy, dz, p, out=out)
    torch.cuda.current_stream().wait_stream(s)
    with torch.cuda.graph(g):
        engine.compute(dx, dy, dz, p, out=out)
    for rep in range(3):
        for k in ("fx", "fy", "fz", "e", "result"):
            out[k].zero_()
        engine.compute(dx[:500], dy[:500], dz[:500], ljamd.make_params(0.1))  # unrelated eager work
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in ("fx", "fy", "fz", "e", "result"):
            assert torch.equal(out[k], eager[k]), (rep, k)
    del g


def test_invalid_params_rejected(engine):
    x, y, z = ljamd.generate("uniform", 1, 100)
    dx, dy, dz = engine.to_device(x, y, z)
    for bad in (
        ljamd.make_params(-1.0),
        ljamd.make_params(0.1, cell_edge=0.05),
        ljamd.make_params(0.1, box_lo=(1, 0, 0), box_hi=(0, 1, 1)),
        ljamd.make_params(0.1, predicate=9),
        ljamd.make_params(1e-9),
        ljamd.make_params(0.1, refine=3),
        ljamd.make_params(0.1, block=96),
        ljamd.make_params(0.1, depth=12),
    ):
        with pytest.raises(ljamd.LJError):
            engine.compute(dx, dy, dz, bad)


def test_c3_precision_subsample(engine):
    """16.7M uniform particles, rc = 0.01: max/RMS force error vs FP64 on the
    65,536-particle subsample (every 256th index); full-size invariants."""
    n = 16_777_216
    x, y, z = ljamd.generate("uniform", 1236, n)
    p = ljamd.make_params(0.01)
    f, e, res = _forces(engine, x, y, z, p, energies=True)
    sub = np.arange(0, n, 256)
    ref = po.forces_f64(po.from_lj(p), x, y, z, subset=sub, nthreads=16)
    fin = (np.abs(ref["fx"]) < 3e38) & (np.abs(ref["fy"]) < 3e38) & (np.abs(ref["fz"]) < 3e38)
    err = po.rel_errors([a[sub][fin] for a in f], (ref["fx"][fin], ref["fy"][fin], ref["fz"][fin]))["rel"]
    print(f"[c3] max rel {err.max():.3e} rms {np.sqrt((err**2).mean()):.3e} nonfinite64 {(~fin).sum()}")
    assert err.max() <= FORCE_TOL
    # size-independent invariants: ordered pairs come in (i,j),(j,i) couples
    assert int(res["n_pairs"]) % 2 == 0
    assert int(res["n_pairs"]) == pytest.approx(70.3 * n, rel=0.02)
