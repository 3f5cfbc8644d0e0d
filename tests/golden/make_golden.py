"""Writes the committed golden fixtures of tests/golden/.

  philox_kat.json   Philox4x64-10 known answers from numpy.random.Philox (numpy, not
                    the reference: the reference uses cuRAND, which is not restated).
                    numpy increments the counter before each block, so
                    Philox(key=k, counter=c).random_raw(4) == philox4x64_10(c + 1, k).
  small_systems.npz positions of small seeded systems (the engine's generator) with the
                    oracle's FP64 forces, energies, pair counts, cell offsets and
                    permutation: a regression pin for generator, binning and oracle.

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "particle-interaction-with-optix_amd"))
sys.path.insert(0, str(ROOT / "oracle"))


def philox_kat():
    cases = []
    rng = np.random.default_rng(2024)
    for _ in range(12):
        key = [int(v) for v in rng.integers(0, 2**63, size=2, dtype=np.int64)]
        ctr = [int(v) for v in rng.integers(1, 2**63, size=4, dtype=np.int64)]
        bg = np.random.Philox(key=np.array(key, np.uint64), counter=np.array([ctr[0] - 1] + ctr[1:], np.uint64))
        out = [int(v) for v in bg.random_raw(4)]
        cases.append({"ctr": ctr, "key": key, "out": out})
    for key, ctr in (([0, 0], [1, 0, 0, 0]), ([1234, 0], [1, 0, 0, 0]), ([1236, 0], [7, 0, 0, 0])):
        bg = np.random.Philox(key=np.array(key, np.uint64), counter=np.array([ctr[0] - 1] + ctr[1:], np.uint64))
        cases.append({"ctr": ctr, "key": key, "out": [int(v) for v in bg.random_raw(4)]})
    (HERE / "philox_kat.json").write_text(json.dumps(cases, indent=1) + "\n")


def small_systems():
    import ljamd
    import pyoracle as po

    specs = {
        "uniform": ("uniform", 1234, 3000, None, dict(rc=0.08)),
        "surface": ("surface", 1235, 3000, None, dict(rc=0.06, box_lo=(-0.5,) * 3, box_hi=(1.5,) * 3)),
        "clusters": ("clusters", 1238, 3000, [8, 0.02, 0.1, 0.9], dict(rc=0.02)),
    }
    arrays = {}
    for name, (kind, seed, n, gp, kw) in specs.items():
        x, y, z = ljamd.generate(kind, seed, n, gp)
        p = po.params(**kw)
        off, perm, _ = po.bin(p, x, y, z)
        r = po.forces_f64(p, x, y, z)
        arrays[f"{name}_pos"] = np.stack([x, y, z])
        arrays[f"{name}_params"] = np.array([p.rc, p.sigma, p.epsilon, *p.box_lo, *p.box_hi], np.float32)
        arrays[f"{name}_f64"] = np.stack([r["fx"], r["fy"], r["fz"], r["e"]])
        arrays[f"{name}_counts"] = np.array([r["n_candidates"], r["n_pairs"]], np.uint64)
        arrays[f"{name}_totals"] = np.array([r["e_total"], r["virial"]])
        arrays[f"{name}_offsets"] = off
        arrays[f"{name}_perm"] = perm
    np.savez_compressed(HERE / "small_systems.npz", **arrays)


if __name__ == "__main__":
    philox_kat()
    small_systems()
    print("wrote", sorted(p.name for p in HERE.iterdir() if p.suffix in (".json", ".npz")))
