"""ljamd -- Python binding of the MI355X Lennard-Jones cutoff engine (liblj_amd.so).

The engine's surface mirrors the reference's host drivers: particle count, cutoff
radius, uniform or sphere-surface generator, per-particle energies
(``executeSimulation`` in Cuda-version/cudaLJ.cu:515-677 and ``core`` in
SDK/optixMultiSphereAABB/optixSphere.cpp:82-584), plus per-particle forces and FP64
totals.  Device buffers are torch tensors on ``cuda``; torch is only the allocator
and the stream provider.  Every compute call goes through the C ABI of
``include/lj_amd.h``; there is no CPU fallback: without the built library or a GPU
the calls raise.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE.parent / "lib" / "liblj_amd.so"

LJ_OK = 0
ERRORS = {
    -1: "LJ_ERR_INVALID_ARG",
    -2: "LJ_ERR_OOM",
    -3: "LJ_ERR_CAPACITY",
    -4: "LJ_ERR_HIP",
    -5: "LJ_ERR_OVERFLOW",
    -6: "LJ_ERR_NO_DEVICE",
}

PRED = {"aabb": 0, "custom": 1, "cudalj": 2, "sphere": 3, "triangle": 4}
PREC = {"mixed": 0, "fp32": 1, "fp32_comp": 2}
FLAG_NO_VIRIAL = 1
FLAG_WALK_GUARD_TEST = 2
NOALLOC_MAX_PARTICLES = 1 << 20  # lj_amd.h LJ_NOALLOC_MAX_PARTICLES
NOALLOC_MAX_CELLS = 1 << 21

GEN = {"uniform": 0, "surface": 1, "clusters": 2}

# Every symbol include/lj_amd.h declares.
EXPORTS = (
    "lj_version",
    "lj_context_create",
    "lj_context_destroy",
    "lj_last_error",
    "lj_grid_dims",
    "lj_generate",
    "lj_bin",
    "lj_reorder",
    "lj_compute",
    "lj_pairs",
    "lj_lbvh_compute",
    "lj_lbvh_pairs",
    "lj_compute_batched",
    "lj_profile_enable",
    "lj_profile_read",
    "lj_autotune",
    "lj_autotune_batched",
    "lj_compute_list",
    "lj_max_displacement",
    "lj_workspace_generation",
    "lj_compute_noalloc",
    "lj_compute_sorted",
)
PHASES = ("bin", "force", "other")


class LJParams(ctypes.Structure):
    _fields_ = [
        ("rc", ctypes.c_float),
        ("sigma", ctypes.c_float),
        ("epsilon", ctypes.c_float),
        ("softening2", ctypes.c_float),
        ("box_lo", ctypes.c_float * 3),
        ("box_hi", ctypes.c_float * 3),
        ("cell_edge", ctypes.c_float),
        ("predicate", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("refine", ctypes.c_int16),
        ("refine_x", ctypes.c_int16),
        ("block", ctypes.c_int16),
        ("depth", ctypes.c_int16),
    ]


class LJResult(ctypes.Structure):
    _fields_ = [
        ("e_total", ctypes.c_double),
        ("virial", ctypes.c_double),
        ("n_candidates", ctypes.c_ulonglong),
        ("n_pairs", ctypes.c_ulonglong),
    ]


class LJSystem(ctypes.Structure):
    _fields_ = [("first", ctypes.c_int64), ("n", ctypes.c_int64), ("params", LJParams)]


class LJTuning(ctypes.Structure):
    _fields_ = [
        ("refine", ctypes.c_int32),
        ("refine_x", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("depth", ctypes.c_int32),
        ("n_tried", ctypes.c_int32),
        ("ms_best", ctypes.c_float),
        ("ms_default", ctypes.c_float),
        ("reserved", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


RESULT_DTYPE = np.dtype(
    [("e_total", "<f8"), ("virial", "<f8"), ("n_candidates", "<u8"), ("n_pairs", "<u8")]
)


class LJError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load(path: Optional[os.PathLike] = None) -> ctypes.CDLL:
    """Loads liblj_amd.so (built by __graft_entry__.build()); raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} is not built; run __graft_entry__.build()")
    # torch ships its own libamdhip64.so.7.  Loading it first makes liblj_amd.so bind
    # to that same runtime (same SONAME); loading ours first would put two HIP
    # runtimes in the process, and the second one to initialise sees no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(str(p))
    vp, i64, u64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64
    P = ctypes.POINTER
    sigs = {
        "lj_version": ([], ctypes.c_char_p),
        "lj_last_error": ([vp], ctypes.c_char_p),
        "lj_context_create": ([P(vp), ctypes.c_int], ctypes.c_int),
        "lj_context_destroy": ([vp], ctypes.c_int),
        "lj_grid_dims": ([P(LJParams), P(ctypes.c_int32), P(i64)], ctypes.c_int),
        "lj_generate": ([ctypes.c_int, u64, u64, i64, P(ctypes.c_double), vp, vp, vp], ctypes.c_int),
        "lj_bin": ([vp, vp, vp, vp, i64, P(LJParams), vp, vp, vp, vp], ctypes.c_int),
        "lj_reorder": ([vp, vp, vp, vp, i64, P(LJParams), vp, vp, vp, vp, vp], ctypes.c_int),
        "lj_compute": ([vp, vp, vp, vp, i64, P(LJParams), vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "lj_lbvh_compute": ([vp, vp, vp, vp, i64, P(LJParams), vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "lj_pairs": ([vp, vp, vp, vp, i64, P(LJParams), vp, vp, ctypes.c_ulonglong, P(ctypes.c_ulonglong), vp], ctypes.c_int),
        "lj_lbvh_pairs": ([vp, vp, vp, vp, i64, P(LJParams), vp, vp, ctypes.c_ulonglong, P(ctypes.c_ulonglong), vp], ctypes.c_int),
        "lj_compute_batched": ([vp, P(LJSystem), ctypes.c_int32, vp, vp, vp, vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "lj_profile_enable": ([vp, ctypes.c_int], ctypes.c_int),
        "lj_profile_read": ([vp, P(ctypes.c_double), P(ctypes.c_ulonglong)], ctypes.c_int),
        "lj_autotune": ([vp, vp, vp, vp, i64, P(LJParams), ctypes.c_int32, P(LJTuning), vp], ctypes.c_int),
        "lj_autotune_batched": ([vp, P(LJSystem), ctypes.c_int32, vp, vp, vp, ctypes.c_int32, P(LJTuning), vp], ctypes.c_int),
        "lj_compute_list": ([vp, vp, vp, vp, i64, P(LJParams), vp, vp, vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "lj_max_displacement": ([vp, vp, vp, vp, vp, vp, vp, i64, vp, vp], ctypes.c_int),
        "lj_workspace_generation": ([vp, P(u64)], ctypes.c_int),
        "lj_compute_noalloc": ([vp, vp, vp, vp, i64, P(LJParams), vp, vp, vp, vp, vp, vp], ctypes.c_int),
        "lj_compute_sorted": ([vp, vp, vp, vp, i64, P(LJParams), vp, vp, vp, vp, vp, vp, vp], ctypes.c_int),
    }
    for name, (args, res) in sigs.items():
        if hasattr(lib, name):  # tests check that every declared symbol is exported
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
    if path is None:
        _lib = lib
    return lib


def make_params(
    rc: float,
    sigma: Optional[float] = None,
    epsilon: float = 1.0,
    softening2: float = 0.0,
    box_lo: Sequence[float] = (0.0, 0.0, 0.0),
    box_hi: Sequence[float] = (1.0, 1.0, 1.0),
    cell_edge: float = 0.0,
    predicate: str | int = "aabb",
    precision: str | int = "mixed",
    flags: int = 0,
    refine: int = 0,
    block: int = 0,
    depth: int = 0,
    refine_x: int = 0,
) -> LJParams:
    """Run parameters; sigma defaults to rc / 2.5 (the workloads' choice).  refine,
    refine_x, block and depth are the kernel shape (0 = default; see Engine.autotune)."""
    p = LJParams()
    p.rc = rc
    p.sigma = rc / 2.5 if sigma is None else sigma
    p.epsilon = epsilon
    p.softening2 = softening2
    for d in range(3):
        p.box_lo[d] = box_lo[d]
        p.box_hi[d] = box_hi[d]
    p.cell_edge = cell_edge
    p.predicate = PRED[predicate] if isinstance(predicate, str) else int(predicate)
    p.precision = PREC[precision] if isinstance(precision, str) else int(precision)
    p.flags = flags
    p.refine = refine
    p.refine_x = refine_x
    p.block = block
    p.depth = depth
    return p


def params_to_dict(p: LJParams) -> dict:
    return {
        "rc": p.rc,
        "sigma": p.sigma,
        "epsilon": p.epsilon,
        "softening2": p.softening2,
        "box_lo": tuple(p.box_lo),
        "box_hi": tuple(p.box_hi),
        "cell_edge": p.cell_edge,
        "predicate": p.predicate,
        "precision": p.precision,
        "flags": p.flags,
        "refine": p.refine,
        "refine_x": p.refine_x,
        "block": p.block,
        "depth": p.depth,
    }


def grid_dims(p: LJParams) -> tuple[tuple[int, int, int], int]:
    lib = load()
    dims = (ctypes.c_int32 * 3)()
    nc = ctypes.c_int64()
    rc = lib.lj_grid_dims(ctypes.byref(p), dims, ctypes.byref(nc))
    if rc != LJ_OK:
        raise LJError(rc, "invalid params")
    return (dims[0], dims[1], dims[2]), nc.value


def generate(kind: str | int, seed: int, n: int, gparams: Optional[Sequence[float]] = None, stream: int = 0):
    """Host generator (Philox4x64-10); returns float32 numpy arrays x, y, z."""
    lib = load()
    k = GEN[kind] if isinstance(kind, str) else int(kind)
    x = np.empty(n, np.float32)
    y = np.empty(n, np.float32)
    z = np.empty(n, np.float32)
    gp = None
    if gparams is not None:
        gp = (ctypes.c_double * len(gparams))(*gparams)
    rc = lib.lj_generate(k, seed, stream, n, gp, x.ctypes.data, y.ctypes.data, z.ctypes.data)
    if rc != LJ_OK:
        raise LJError(rc, "lj_generate failed")
    return x, y, z


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise LJError(-6, "no GPU visible to torch; the engine has no CPU path")
    return torch


def _ptr(t) -> Optional[int]:
    return None if t is None else t.data_ptr()


def read_results(t, count: int = 1) -> np.ndarray:
    """Device lj_result array (uint8 tensor) -> numpy structured array."""
    return np.frombuffer(t.cpu().numpy().tobytes(), dtype=RESULT_DTYPE, count=count)


class Engine:
    """One lj_context on one device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.torch = _torch()
        self.device = device
        h = ctypes.c_void_p()
        rc = self.lib.lj_context_create(ctypes.byref(h), device)
        if rc != LJ_OK:
            raise LJError(rc, self.lib.lj_last_error(None).decode())
        self.ctx = h

    def close(self):
        if self.ctx:
            self.lib.lj_context_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- helpers
    def _check(self, rc: int):
        if rc != LJ_OK:
            raise LJError(rc, self.lib.lj_last_error(self.ctx).decode())

    def _stream(self):
        return self.torch.cuda.current_stream(self.device).cuda_stream

    def to_device(self, *arrays):
        t = self.torch
        return [t.from_numpy(np.ascontiguousarray(a)).to(f"cuda:{self.device}") for a in arrays]

    def empty_f32(self, n):
        return self.torch.empty(max(n, 1), dtype=self.torch.float32, device=f"cuda:{self.device}")

    def empty_result(self, count: int = 1):
        return self.torch.zeros(32 * count, dtype=self.torch.uint8, device=f"cuda:{self.device}")

    # ------------------------------------------------------------- calls
    def bin(self, x, y, z, p: LJParams):
        t = self.torch
        n = x.numel()
        _, nc = grid_dims(p)
        off = t.empty(nc + 1, dtype=t.int32, device=x.device)
        perm = t.empty(max(n, 1), dtype=t.int32, device=x.device)
        cand = t.zeros(1, dtype=t.int64, device=x.device)
        self._check(
            self.lib.lj_bin(self.ctx, _ptr(x), _ptr(y), _ptr(z), n, ctypes.byref(p), _ptr(off), _ptr(perm), _ptr(cand), self._stream())
        )
        return off, perm[:n], cand

    def reorder(self, x, y, z, p: LJParams):
        n = x.numel()
        xo, yo, zo = self.empty_f32(n), self.empty_f32(n), self.empty_f32(n)
        perm = self.torch.empty(max(n, 1), dtype=self.torch.int32, device=x.device)
        self._check(
            self.lib.lj_reorder(self.ctx, _ptr(x), _ptr(y), _ptr(z), n, ctypes.byref(p), _ptr(xo), _ptr(yo), _ptr(zo), _ptr(perm), self._stream())
        )
        return xo[:n], yo[:n], zo[:n], perm[:n]

    def compute(self, x, y, z, p: LJParams, energies: bool = True, out=None, lbvh: bool = False):
        """Forces (and per-particle energies) of one system; returns a dict of tensors
        plus the device lj_result buffer under 'result'."""
        n = x.numel()
        if out is None:
            out = {
                "fx": self.empty_f32(n),
                "fy": self.empty_f32(n),
                "fz": self.empty_f32(n),
                "e": self.empty_f32(n) if energies else None,
                "result": self.empty_result(),
            }
        fn = self.lib.lj_lbvh_compute if lbvh else self.lib.lj_compute
        self._check(
            fn(self.ctx, _ptr(x), _ptr(y), _ptr(z), n, ctypes.byref(p), _ptr(out["fx"]), _ptr(out["fy"]), _ptr(out["fz"]),
               _ptr(out["e"]), _ptr(out["result"]), self._stream())
        )
        return out

    def compute_sorted(self, x, y, z, p: LJParams, energies: bool = True, out=None):
        """lj_compute_sorted: forces and energies with the particles left in cell order.
        x, y, z are overwritten with the sorted positions; the returned fx, fy, fz, e
        are in that order and 'perm' holds each particle's input index."""
        n = x.numel()
        if out is None:
            out = {
                "fx": self.empty_f32(n),
                "fy": self.empty_f32(n),
                "fz": self.empty_f32(n),
                "e": self.empty_f32(n) if energies else None,
                "perm": self.torch.empty(max(n, 1), dtype=self.torch.int32, device=x.device),
                "result": self.empty_result(),
            }
        self._check(
            self.lib.lj_compute_sorted(self.ctx, _ptr(x), _ptr(y), _ptr(z), n, ctypes.byref(p), _ptr(out["fx"]),
                                       _ptr(out["fy"]), _ptr(out["fz"]), _ptr(out["e"]), _ptr(out.get("perm")),
                                       _ptr(out["result"]), self._stream())
        )
        return out

    def compute_noalloc(self, x, y, z, p: LJParams, out):
        """lj_compute_noalloc: compute() with the workspace in the library's fixed device
        array (no allocation); `out` holds fx, fy, fz, e (nullable) and result."""
        self._check(
            self.lib.lj_compute_noalloc(self.ctx, _ptr(x), _ptr(y), _ptr(z), x.numel(), ctypes.byref(p), _ptr(out["fx"]),
                                        _ptr(out["fy"]), _ptr(out["fz"]), _ptr(out["e"]), _ptr(out["result"]),
                                        self._stream())
        )
        return out

    def pairs(self, x, y, z, p: LJParams, capacity: Optional[int] = None, lbvh: bool = False):
        """CSR neighbour lists (row offsets by original index, rows sorted)."""
        t = self.torch
        n = x.numel()
        rows = t.empty(n + 1, dtype=t.int64, device=x.device)
        needed = ctypes.c_ulonglong()
        fn = self.lib.lj_lbvh_pairs if lbvh else self.lib.lj_pairs
        if capacity is None:
            rc = fn(self.ctx, _ptr(x), _ptr(y), _ptr(z), n, ctypes.byref(p), _ptr(rows), None, 0, ctypes.byref(needed), self._stream())
            if rc not in (LJ_OK, -3):
                self._check(rc)
            capacity = needed.value
        cols = t.empty(max(capacity, 1), dtype=t.int32, device=x.device)
        self._check(
            fn(self.ctx, _ptr(x), _ptr(y), _ptr(z), n, ctypes.byref(p), _ptr(rows), _ptr(cols), capacity, ctypes.byref(needed), self._stream())
        )
        return rows, cols[: needed.value]

    def workspace_generation(self) -> int:
        """Changes whenever a call reallocates workspace (a captured graph's addresses
        are then stale): record it after capturing, replay only while it is unchanged."""
        g = ctypes.c_uint64()
        self._check(self.lib.lj_workspace_generation(self.ctx, ctypes.byref(g)))
        return g.value

    def profile(self, enable: bool = True):
        """Turns per-phase hipEvent timing on (and resets it) or off."""
        self._check(self.lib.lj_profile_enable(self.ctx, 1 if enable else 0))

    def profile_read(self) -> dict:
        """Summed device milliseconds and launch counts per phase since profile(True)."""
        ms = (ctypes.c_double * 3)()
        calls = (ctypes.c_ulonglong * 3)()
        self._check(self.lib.lj_profile_read(self.ctx, ms, calls))
        return {name: (ms[q], calls[q]) for q, name in enumerate(PHASES)}

    def compute_batched(self, systems: Sequence[LJSystem], x, y, z, energies: bool = True, out=None):
        n = x.numel()
        arr = (LJSystem * len(systems))(*systems)
        if out is None:
            out = {
                "fx": self.empty_f32(n),
                "fy": self.empty_f32(n),
                "fz": self.empty_f32(n),
                "e": self.empty_f32(n) if energies else None,
                "result": self.empty_result(len(systems)),
            }
        self._check(
            self.lib.lj_compute_batched(self.ctx, arr, len(systems), _ptr(x), _ptr(y), _ptr(z), _ptr(out["fx"]), _ptr(out["fy"]),
                                        _ptr(out["fz"]), _ptr(out["e"]), _ptr(out["result"]), self._stream())
        )
        return out

    def compute_list(self, x, y, z, p: LJParams, rows, cols, energies: bool = True, out=None):
        """Forces over a stored CSR neighbour list (see pairs()); same outputs as compute()."""
        n = x.numel()
        if out is None:
            out = {
                "fx": self.empty_f32(n),
                "fy": self.empty_f32(n),
                "fz": self.empty_f32(n),
                "e": self.empty_f32(n) if energies else None,
                "result": self.empty_result(),
            }
        self._check(
            self.lib.lj_compute_list(self.ctx, _ptr(x), _ptr(y), _ptr(z), n, ctypes.byref(p), _ptr(rows), _ptr(cols),
                                     _ptr(out["fx"]), _ptr(out["fy"]), _ptr(out["fz"]), _ptr(out["e"]),
                                     _ptr(out["result"]), self._stream())
        )
        return out

    def max_displacement2(self, x, y, z, x0, y0, z0, out=None):
        """Device float tensor holding max |p - p0|^2."""
        if out is None:
            out = self.torch.empty(1, dtype=self.torch.float32, device=x.device)
        self._check(
            self.lib.lj_max_displacement(self.ctx, _ptr(x), _ptr(y), _ptr(z), _ptr(x0), _ptr(y0), _ptr(z0), x.numel(),
                                         _ptr(out), self._stream())
        )
        return out

    def autotune(self, x, y, z, p: LJParams, reps: int = 3) -> dict:
        """Times every valid kernel shape on these positions and writes the fastest
        into p (refine, refine_x, block, depth); returns the lj_tuning record as a dict."""
        t = LJTuning()
        self._check(
            self.lib.lj_autotune(self.ctx, _ptr(x), _ptr(y), _ptr(z), x.numel(), ctypes.byref(p), reps, ctypes.byref(t), self._stream())
        )
        return t.as_dict()

    def autotune_batched(self, systems: Sequence[LJSystem], x, y, z, reps: int = 3) -> dict:
        """One shape for a whole batch, written into every system's params."""
        arr = (LJSystem * len(systems))(*systems)
        t = LJTuning()
        self._check(
            self.lib.lj_autotune_batched(self.ctx, arr, len(systems), _ptr(x), _ptr(y), _ptr(z), reps, ctypes.byref(t), self._stream())
        )
        for q, sy in enumerate(systems):
            sy.params.refine = arr[q].params.refine
            sy.params.refine_x = arr[q].params.refine_x
            sy.params.block = arr[q].params.block
            sy.params.depth = arr[q].params.depth
        return t.as_dict()


def version() -> str:
    return load().lj_version().decode()
