"""ctypes binding of the CPU oracle (oracle/liblj_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() (as the
checker) and by bench.py's cpu_baseline leg.  Never used by the product path.
"""
from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liblj_oracle.so"

PRED = {"aabb": 0, "custom": 1, "cudalj": 2, "sphere": 3, "triangle": 4}


class OrcParams(ctypes.Structure):
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
        ("flags", ctypes.c_int32),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        vp, i64, u64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64
        P = ctypes.POINTER(OrcParams)
        L.orc_grid_dims.argtypes = [P, vp, vp]
        L.orc_cell_key.argtypes = [P, vp, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.orc_cell_key.restype = ctypes.c_uint32
        L.orc_bin.argtypes = [P, vp, vp, vp, i64, vp, vp, vp]
        L.orc_candidates.argtypes = [vp, vp]
        L.orc_candidates.restype = u64
        L.orc_candidates_direct.argtypes = [P, vp, vp, vp, i64, vp]
        L.orc_candidates_direct.restype = u64
        L.orc_d2.argtypes = [ctypes.c_float] * 6
        L.orc_d2.restype = ctypes.c_float
        L.orc_accept.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, i64, i64]
        L.orc_pair_mult.argtypes = [P] + [ctypes.c_float] * 6 + [i64, i64]
        L.orc_pairs_brute.argtypes = [P, vp, vp, vp, i64, vp, vp, u64]
        L.orc_pairs_brute.restype = u64
        L.orc_pairs_cellgrid.argtypes = [P, vp, vp, vp, i64, vp, vp, u64]
        L.orc_pairs_cellgrid.restype = u64
        L.orc_pairs_rows.argtypes = [P, vp, vp, vp, i64, vp, i64, vp, vp, u64, ctypes.c_int]
        L.orc_pairs_rows.restype = u64
        L.orc_forces_f64.argtypes = [P, vp, vp, vp, i64, vp, i64, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int]
        L.orc_forces_f32_cellgrid.argtypes = [P, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp, ctypes.c_int]
        L.orc_cudalj_energy.argtypes = [P, vp, vp, vp, i64, vp]
        L.orc_philox4x64_10.argtypes = [vp, vp, vp]
        _lib = L
    return _lib


def params(rc, sigma=None, epsilon=1.0, softening2=0.0, box_lo=(0, 0, 0), box_hi=(1, 1, 1), cell_edge=0.0,
           predicate="aabb"):
    p = OrcParams()
    p.rc = rc
    p.sigma = rc / 2.5 if sigma is None else sigma
    p.epsilon = epsilon
    p.softening2 = softening2
    for d in range(3):
        p.box_lo[d] = box_lo[d]
        p.box_hi[d] = box_hi[d]
    p.cell_edge = cell_edge
    p.predicate = PRED[predicate] if isinstance(predicate, str) else int(predicate)
    p.precision = 0
    p.flags = 0
    return p


def from_lj(p) -> OrcParams:
    """Converts an ljamd.LJParams (same layout) into OrcParams."""
    q = OrcParams()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(q))
    return q


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def grid_dims(p):
    dims = np.zeros(3, np.int32)
    nc = np.zeros(1, np.int64)
    rc = lib().orc_grid_dims(ctypes.byref(p), dims.ctypes.data, nc.ctypes.data)
    if rc:
        raise ValueError(f"orc_grid_dims: {rc}")
    return tuple(int(d) for d in dims), int(nc[0])


def cell_key(p, dims, x, y, z):
    d = np.asarray(dims, np.int32)
    return lib().orc_cell_key(ctypes.byref(p), d.ctypes.data, x, y, z)


def bin(p, x, y, z):
    x, y, z = _f32(x), _f32(y), _f32(z)
    n = x.size
    _, nc = grid_dims(p)
    off = np.zeros(nc + 1, np.uint32)
    perm = np.zeros(max(n, 1), np.uint32)
    keys = np.zeros(max(n, 1), np.uint32)
    rc = lib().orc_bin(ctypes.byref(p), x.ctypes.data, y.ctypes.data, z.ctypes.data, n, off.ctypes.data,
                       perm.ctypes.data, keys.ctypes.data)
    if rc:
        raise ValueError(f"orc_bin: {rc}")
    return off, perm[:n], keys[:n]


def candidates(p, off):
    dims, _ = grid_dims(p)
    d = np.asarray(dims, np.int32)
    off = np.ascontiguousarray(off, np.uint32)
    return int(lib().orc_candidates(d.ctypes.data, off.ctypes.data))


def candidates_direct(p, x, y, z, off):
    x, y, z = _f32(x), _f32(y), _f32(z)
    off = np.ascontiguousarray(off, np.uint32)
    return int(lib().orc_candidates_direct(ctypes.byref(p), x.ctypes.data, y.ctypes.data, z.ctypes.data, x.size,
                                           off.ctypes.data))


def d2(pi, pj):
    return lib().orc_d2(float(pi[0]), float(pi[1]), float(pi[2]), float(pj[0]), float(pj[1]), float(pj[2]))


def _pairs(fn, p, x, y, z):
    x, y, z = _f32(x), _f32(y), _f32(z)
    n = x.size
    rows = np.zeros(n + 1, np.uint64)
    m = fn(ctypes.byref(p), x.ctypes.data, y.ctypes.data, z.ctypes.data, n, rows.ctypes.data, None, 0)
    cols = np.zeros(max(int(m), 1), np.uint32)
    fn(ctypes.byref(p), x.ctypes.data, y.ctypes.data, z.ctypes.data, n, rows.ctypes.data, cols.ctypes.data, m)
    return rows, cols[: int(m)]


def pairs_brute(p, x, y, z):
    return _pairs(lib().orc_pairs_brute, p, x, y, z)


def pairs_cellgrid(p, x, y, z):
    return _pairs(lib().orc_pairs_cellgrid, p, x, y, z)


def pairs_rows(p, x, y, z, subset, nthreads=0):
    """CSR rows (sorted) of the particles in subset, as orc_pairs_cellgrid gives them."""
    x, y, z = _f32(x), _f32(y), _f32(z)
    sub = np.ascontiguousarray(subset, np.int64)
    rows = np.zeros(sub.size + 1, np.uint64)
    args = (ctypes.byref(p), x.ctypes.data, y.ctypes.data, z.ctypes.data, x.size, sub.ctypes.data, sub.size)
    m = lib().orc_pairs_rows(*args, rows.ctypes.data, None, 0, nthreads)
    cols = np.zeros(max(int(m), 1), np.uint32)
    lib().orc_pairs_rows(*args, rows.ctypes.data, cols.ctypes.data, m, nthreads)
    return rows, cols[: int(m)]


def forces_f64(p, x, y, z, subset=None, nthreads=0):
    """FP64 oracle. Returns dict fx, fy, fz, e, cond (per row of subset or all),
    e_total, virial, n_candidates, n_pairs (over the rows computed)."""
    x, y, z = _f32(x), _f32(y), _f32(z)
    n = x.size
    sub = None if subset is None else np.ascontiguousarray(subset, np.int64)
    m = n if sub is None else sub.size
    out = {k: np.zeros(max(m, 1)) for k in ("fx", "fy", "fz", "e", "cond")}
    sums = np.zeros(2)
    counts = np.zeros(2, np.uint64)
    rc = lib().orc_forces_f64(ctypes.byref(p), x.ctypes.data, y.ctypes.data, z.ctypes.data, n,
                              None if sub is None else sub.ctypes.data, m,
                              out["fx"].ctypes.data, out["fy"].ctypes.data, out["fz"].ctypes.data,
                              out["e"].ctypes.data, out["cond"].ctypes.data, sums.ctypes.data,
                              counts.ctypes.data, nthreads)
    if rc:
        raise ValueError(f"orc_forces_f64: {rc}")
    res = {k: v[:m] for k, v in out.items()}
    res.update(e_total=float(sums[0]), virial=float(sums[1]), n_candidates=int(counts[0]), n_pairs=int(counts[1]))
    return res


def forces_f32(p, x, y, z, nthreads=0):
    """FP32 cell-grid restatement (the CPU baseline)."""
    x, y, z = _f32(x), _f32(y), _f32(z)
    n = x.size
    out = {k: np.zeros(max(n, 1), np.float32) for k in ("fx", "fy", "fz", "e")}
    sums = np.zeros(2)
    counts = np.zeros(2, np.uint64)
    rc = lib().orc_forces_f32_cellgrid(ctypes.byref(p), x.ctypes.data, y.ctypes.data, z.ctypes.data, n,
                                       out["fx"].ctypes.data, out["fy"].ctypes.data, out["fz"].ctypes.data,
                                       out["e"].ctypes.data, sums.ctypes.data, counts.ctypes.data, nthreads)
    if rc:
        raise ValueError(f"orc_forces_f32_cellgrid: {rc}")
    res = {k: v[:n] for k, v in out.items()}
    res.update(e_total=float(sums[0]), virial=float(sums[1]), n_candidates=int(counts[0]), n_pairs=int(counts[1]))
    return res


def cudalj_energy(p, x, y, z):
    x, y, z = _f32(x), _f32(y), _f32(z)
    e = np.zeros(max(x.size, 1))
    rc = lib().orc_cudalj_energy(ctypes.byref(p), x.ctypes.data, y.ctypes.data, z.ctypes.data, x.size, e.ctypes.data)
    if rc:
        raise ValueError(f"orc_cudalj_energy: {rc}")
    return e[: x.size]


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint64)
    k = np.ascontiguousarray(key, np.uint64)
    out = np.zeros(4, np.uint64)
    lib().orc_philox4x64_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def rel_errors(f32, f64, cond=None):
    """Per-particle relative force error |F - F64| / |F64| (rows with |F64| = 0 use
    the absolute error), plus the conditioned form |F - F64| / sum_j |f_ij|."""
    fx, fy, fz = (np.asarray(a, np.float64) for a in f32)
    gx, gy, gz = (np.asarray(a, np.float64) for a in f64)
    diff = np.sqrt((fx - gx) ** 2 + (fy - gy) ** 2 + (fz - gz) ** 2)
    mag = np.sqrt(gx**2 + gy**2 + gz**2)
    rel = np.where(mag > 0, diff / np.where(mag > 0, mag, 1.0), diff)
    out = {"rel": rel}
    if cond is not None:
        c = np.asarray(cond, np.float64)
        out["cond"] = np.where(c > 0, diff / np.where(c > 0, c, 1.0), diff)
    return out
