"""lj_compute_noalloc, the entry point that allocates nothing, against lj_compute.

Same kernels and summation order, so the results must be the same bits: forces,
energies and every field of lj_result, on two calls in a row (the second smaller, so
that it runs on what the first left in the shared array), on a refined kernel shape,
and at the largest input the array takes.  Inputs beyond it are refused with
LJ_ERR_CAPACITY, and the context's own workspace never grows.
"""
import numpy as np
import pytest

import ljamd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = ljamd.Engine(0)
    yield e
    e.close()


def _outputs(eng, n):
    return {k: eng.empty_f32(n) for k in ("fx", "fy", "fz", "e")} | {"result": eng.empty_result()}


def _same(eng, x, y, z, p):
    dx, dy, dz = eng.to_device(x, y, z)
    n = x.size
    gen = eng.workspace_generation()
    a = eng.compute_noalloc(dx, dy, dz, p, _outputs(eng, n))
    eng.torch.cuda.synchronize()
    assert eng.workspace_generation() == gen  # nothing grew through the context
    b = eng.compute(dx, dy, dz, p, out=_outputs(eng, n))
    eng.torch.cuda.synchronize()
    for k in ("fx", "fy", "fz", "e"):
        assert eng.torch.equal(a[k][:n].view(eng.torch.int32), b[k][:n].view(eng.torch.int32)), k
    ra, rb = ljamd.read_results(a["result"])[0], ljamd.read_results(b["result"])[0]
    assert ra.tobytes() == rb.tobytes()
    assert int(ra["n_pairs"]) > 0
    return ra


def test_noalloc_two_calls_match_compute(eng):
    print("[noalloc] 100k then 20k", flush=True)
    x, y, z = ljamd.generate("uniform", 1234, 100_000)
    _same(eng, x, y, z, ljamd.make_params(0.04))
    x, y, z = ljamd.generate("uniform", 77, 20_000)
    _same(eng, x, y, z, ljamd.make_params(0.07))


def test_noalloc_refined_shape_and_surface(eng):
    print("[noalloc] refined, surface", flush=True)
    x, y, z = ljamd.generate("uniform", 5, 200_000)
    _same(eng, x, y, z, ljamd.make_params(0.04, refine=2, refine_x=4))
    x, y, z = ljamd.generate("surface", 6, 100_000)
    _same(eng, x, y, z, ljamd.make_params(0.02, box_lo=(-0.5,) * 3, box_hi=(1.5,) * 3))


def test_noalloc_at_capacity_then_small(eng):
    n = ljamd.NOALLOC_MAX_PARTICLES
    print(f"[noalloc] n={n}", flush=True)
    x, y, z = ljamd.generate("uniform", 9, n)
    p = ljamd.make_params(0.0125, refine=1, refine_x=2)  # 80 x 80 x 160 = 1,024,000 cells
    _same(eng, x, y, z, p)
    x, y, z = ljamd.generate("uniform", 10, 3000)
    _same(eng, x, y, z, ljamd.make_params(0.1))


def test_noalloc_rejects_oversize(eng):
    x, y, z = ljamd.generate("uniform", 11, ljamd.NOALLOC_MAX_PARTICLES + 1)
    dx, dy, dz = eng.to_device(x, y, z)
    with pytest.raises(ljamd.LJError) as ei:
        eng.compute_noalloc(dx, dy, dz, ljamd.make_params(0.02), _outputs(eng, x.size))
    assert ei.value.code == -3
    dx, dy, dz = dx[:1000], dy[:1000], dz[:1000]
    with pytest.raises(ljamd.LJError) as ei:  # 200^3 cells
        eng.compute_noalloc(dx, dy, dz, ljamd.make_params(0.005), _outputs(eng, 1000))
    assert ei.value.code == -3 and "cells" in str(ei.value)
