"""The benchmark workloads as concrete, seeded inputs.

The reference sweeps box divisions and particles per cell with positions from cuRAND
or drand48 (Cuda-version/cudaLJ.cu:701-792, SDK/optixMultiSphereAABB/optixSphere.cpp:
713-799).  These are the engine's fixed workloads instead:

  c1  N=100,000 uniform in [0,1)^3, rc=0.04                    (seed 1234)
  c2  N=1,000,000 on the unit sphere centred (0.5,0.5,0.5),
      rc=0.01, box [-0.5,1.5]^3                                (seed 1235)
  c3  N=16,777,216 uniform in [0,1)^3, rc=0.01                 (seed 1236)
  c4  4,096 systems, N_i log-uniform in [256, 65536], even ids
      cube / odd ids sphere surface, rc_i for a mean neighbour
      count k_i ~ U[20,80]                                     (seed 1237)
  c5  N=4,194,304 in 256 Gaussian clusters (std 0.02, centres
      U[0.1,0.9]^3), rc=0.005                                  (seed 1238)
sigma = rc / 2.5 and epsilon = 1 everywhere; the cell edge is rc.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import LJSystem, generate, make_params

UNIT = dict(box_lo=(0.0, 0.0, 0.0), box_hi=(1.0, 1.0, 1.0))
SPHERE_BOX = dict(box_lo=(-0.5, -0.5, -0.5), box_hi=(1.5, 1.5, 1.5))


@dataclass
class Workload:
    name: str
    desc: str
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    params: object


def single(name: str, n: int | None = None, precision: str = "mixed") -> Workload:
    if name == "c1":
        n = n or 100_000
        x, y, z = generate("uniform", 1234, n)
        p = make_params(0.04, precision=precision, **UNIT)
        desc = f"c1 uniform N={n} rc=0.04"
    elif name == "c2":
        n = n or 1_000_000
        x, y, z = generate("surface", 1235, n, [0.5, 0.5, 0.5, 1.0, 0.0])
        p = make_params(0.01, precision=precision, **SPHERE_BOX)
        desc = f"c2 sphere-surface N={n} rc=0.01"
    elif name == "c3":
        n = n or 16_777_216
        x, y, z = generate("uniform", 1236, n)
        p = make_params(0.01, precision=precision, **UNIT)
        desc = f"c3 uniform N={n} rc=0.01"
    elif name == "c5":
        n = n or 4_194_304
        x, y, z = generate("clusters", 1238, n, [256, 0.02, 0.1, 0.9])
        p = make_params(0.005, precision=precision, **UNIT)
        desc = f"c5 256 gaussian clusters N={n} rc=0.005"
    else:
        raise KeyError(name)
    return Workload(name, desc + f" sigma=rc/2.5 eps=1 {precision}", x, y, z, p)


def batch_c4(nsys: int = 4096, seed: int = 1237, nmin: int = 256, nmax: int = 65536, precision: str = "mixed"):
    """Returns (systems, x, y, z): packed positions and one LJSystem per system."""
    bg = np.random.Philox(key=np.array([seed, 0xC4], np.uint64))
    u = bg.random_raw(2 * nsys).reshape(nsys, 2)
    u = (u >> np.uint64(11)).astype(np.float64) * 2.0**-53
    sizes = np.rint(np.exp(math.log(nmin) + u[:, 0] * (math.log(nmax) - math.log(nmin)))).astype(np.int64)
    ks = 20.0 + 60.0 * u[:, 1]
    total = int(sizes.sum())
    X = np.empty(total, np.float32)
    Y = np.empty(total, np.float32)
    Z = np.empty(total, np.float32)
    systems = []
    first = 0
    for s in range(nsys):
        n, k = int(sizes[s]), float(ks[s])
        if s % 2 == 0:
            rc = (3.0 * k / (4.0 * math.pi * n)) ** (1.0 / 3.0)
            x, y, z = generate("uniform", seed, n, stream=s + 1)
            p = make_params(rc, precision=precision, **UNIT)
        else:
            rc = 2.0 * math.sqrt(k / (n - 1))
            x, y, z = generate("surface", seed, n, [0.5, 0.5, 0.5, 1.0, 0.0], stream=s + 1)
            p = make_params(rc, precision=precision, **SPHERE_BOX)
        X[first : first + n] = x
        Y[first : first + n] = y
        Z[first : first + n] = z
        systems.append(LJSystem(first, n, p))
        first += n
    return systems, X, Y, Z
