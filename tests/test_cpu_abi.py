"""CPU checks of the C ABI: the library loads, exports every function declared in
include/lj_amd.h, and its host-only helpers behave.  No GPU compute is called."""
import ctypes
import pathlib
import re

import numpy as np
import pytest

import ljamd
import pyoracle as po

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "lj_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lj_[a-z_0-9]+)\s*\(", text)))


def test_header_declarations_are_all_exported():
    lib = ljamd.load()
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(ljamd.EXPORTS)


def test_struct_layouts_match_header_sizes():
    # lj_params: 4 floats, 2x3 floats, 1 float, 5 x 32-bit ints = 64 bytes
    assert ctypes.sizeof(ljamd.LJParams) == 64
    assert ctypes.sizeof(ljamd.LJResult) == 32
    assert ctypes.sizeof(ljamd.LJSystem) == 16 + 64
    assert ctypes.sizeof(ljamd.LJTuning) == 32
    assert ljamd.LJParams.refine.offset == 56 and ljamd.LJParams.refine_x.offset == 58
    assert ljamd.LJParams.block.offset == 60 and ljamd.LJParams.depth.offset == 62
    # the oracle's params are the leading fields of the engine's (same offsets)
    for name, _ in po.OrcParams._fields_:
        assert getattr(po.OrcParams, name).offset == getattr(ljamd.LJParams, name).offset


def test_version_string():
    assert "gfx950" in ljamd.version()


@pytest.mark.parametrize("rc,lo,hi", [(0.04, 0.0, 1.0), (0.01, -0.5, 1.5), (0.005, 0.0, 1.0), (0.3, 0.2, 0.9)])
def test_grid_dims_agrees_with_oracle(rc, lo, hi):
    p = ljamd.make_params(rc, box_lo=(lo,) * 3, box_hi=(hi,) * 3)
    assert ljamd.grid_dims(p) == po.grid_dims(po.from_lj(p))


def test_grid_dims_rejects_bad_params():
    for p in (
        ljamd.make_params(0.0),
        ljamd.make_params(0.1, box_lo=(1, 1, 1), box_hi=(0, 0, 0)),
        ljamd.make_params(0.1, cell_edge=0.05),  # cell edge below rc would miss pairs
        ljamd.make_params(1e-9),  # more than 2^24 cells per axis
        ljamd.make_params(0.1, predicate=17),
        ljamd.make_params(0.1, precision=9),
        ljamd.make_params(0.1, refine=3),
        ljamd.make_params(0.1, block=96),
        ljamd.make_params(0.1, depth=12),
        ljamd.make_params(0.1, refine_x=3),
        ljamd.make_params(0.1, refine=4, refine_x=2),
    ):
        with pytest.raises(ljamd.LJError):
            ljamd.grid_dims(p)
    # the CUDALJ compatibility mode may use any cell edge (no cutoff)
    ljamd.grid_dims(ljamd.make_params(0.1, cell_edge=0.05, predicate="cudalj"))


def test_generate_rejects_bad_input():
    lib = ljamd.load()
    buf = np.zeros(4, np.float32)
    assert lib.lj_generate(9, 1, 0, 4, None, buf.ctypes.data, buf.ctypes.data, buf.ctypes.data) == -1
    assert lib.lj_generate(0, 1, 0, -1, None, buf.ctypes.data, buf.ctypes.data, buf.ctypes.data) == -1
    assert lib.lj_generate(0, 1, 0, 4, None, None, None, None) == -1


def test_generate_streams_are_independent_and_deterministic():
    a = ljamd.generate("uniform", 5, 1000, stream=1)
    b = ljamd.generate("uniform", 5, 1000, stream=2)
    c = ljamd.generate("uniform", 5, 1000, stream=1)
    assert not np.array_equal(a[0], b[0])
    assert all(np.array_equal(u, v) for u, v in zip(a, c))
    # a prefix of a longer draw is the shorter draw (counter-based)
    d = ljamd.generate("uniform", 5, 2000, stream=1)
    assert np.array_equal(d[0][:1000], a[0])


def test_context_without_gpu_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = ljamd.load()
    h = ctypes.c_void_p()
    rc = lib.lj_context_create(ctypes.byref(h), 0)
    assert rc == -6 and not h.value
    assert b"device" in lib.lj_last_error(None)
    # and the Python engine refuses to run rather than falling back to the CPU
    with pytest.raises(ljamd.LJError):
        ljamd.Engine(0)


def test_batch_workload_shapes():
    from ljamd import workloads

    systems, X, Y, Z = workloads.batch_c4(nsys=64, nmin=256, nmax=65536)
    sizes = np.array([s.n for s in systems])
    assert sizes.min() >= 256 and sizes.max() <= 65536 and X.size == sizes.sum()
    firsts = np.array([s.first for s in systems])
    assert np.array_equal(firsts, np.concatenate([[0], np.cumsum(sizes)[:-1]]))
