"""The command-line driver (bin/ljbench) run as a child process and checked against the
oracle on the same positions.

ljbench restates the reference's drivers on the C ABI: the cudaLJ sweep of
executeSimulation (Cuda-version/cudaLJ.cu:515-677, 701-792), the AABB driver's
checked single run and its sweep (SDK/optixMultiSphereAABB/optixSphere.cpp:550-575,
622-830), and an engine mode.  Each run here is checked for its pair and candidate
counts, energies and CSV columns.  The driver is started before this process touches
the GPU (the module runs first among the GPU tests), so no process that has
initialised the GPU starts another program; the positions are regenerated here with
numpy's Philox (the generator the CPU tests pin lj_generate to), so this process
never loads the engine library.
"""
import csv
import math
import pathlib
import re
import subprocess

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parents[1]
LJBENCH = ROOT / "particle-interaction-with-optix_amd" / "bin" / "ljbench"


@pytest.fixture(autouse=True)
def _fresh_process():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_initialized():
        pytest.skip("this process already uses the GPU; ljbench is started only from a process that does not")
    if not LJBENCH.exists():
        pytest.fail(f"{LJBENCH} is not built; run __graft_entry__.build()")


def _uniform(seed, stream, n, lo=0.0, hi=1.0):
    """lj_generate(LJ_GEN_UNIFORM): Philox4x64-10 block (i, 0, 0, 0) under key (seed,
    stream), top 24 bits of its first three words scaled to [lo, hi) in FP64 and
    rounded to FP32 (kept below hi)."""
    bg = np.random.Philox(key=np.array([seed, stream], np.uint64), counter=np.full(4, 2**64 - 1, np.uint64))
    w = bg.random_raw(4 * n).reshape(n, 4)[:, :3]
    v = (lo + (w >> np.uint64(40)).astype(np.float64) * 2.0**-24 * (hi - lo)).astype(np.float32)
    v = np.where(v >= np.float32(hi), np.nextafter(np.float32(hi), np.float32(lo)), v)
    return v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy()


def _run(args, cwd, timeout=120):
    print(f"[ljbench] {' '.join(args)}", flush=True)
    r = subprocess.run([str(LJBENCH), *args], cwd=cwd, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def test_ljbench_engine_mode_matches_oracle(tmp_path):
    """`ljbench lj --n 100000 --rc 0.04` (the C1 workload): pairs, 27-cell candidates,
    energy and virial totals equal the FP64 oracle's on the same Philox positions."""
    out = _run(["lj", "--n", "100000", "--rc", "0.04", "--loops", "1"], tmp_path)
    m = re.search(r"pairs=(\d+) candidates=(\d+) e_total=(\S+) virial=(\S+)", out)
    assert m, out
    x, y, z = _uniform(1234, 0, 100_000)
    ref = po.forces_f64(po.params(0.04), x, y, z, nthreads=8)
    assert int(m.group(1)) == ref["n_pairs"]
    assert int(m.group(2)) == ref["n_candidates"]
    assert float(m.group(3)) == pytest.approx(ref["e_total"], rel=1e-9)
    assert float(m.group(4)) == pytest.approx(ref["virial"], rel=1e-9)


def test_ljbench_cudalj_sweep_matches_oracle(tmp_path):
    """`ljbench cudalj --max-boxdiv 4`: the reference's uniform sweep for boxDiv 2 and 4
    (N = boxDiv^3 * 2^k, k = 0..5, five loops each, cell edge 1/boxDiv, positions
    regenerated every loop).  Every loop's 27-cell candidate count equals the oracle's
    and its energy total equals the oracle's restatement of cudaLJ's softened,
    cutoff-free energy (cudaLJ.cu:28-56, 299-341); the CSV has the reference's columns
    and one row per loop, NbInteractions the sum over the point's loops."""
    out = _run(["cudalj", "--max-boxdiv", "4"], tmp_path)
    blocks = re.split(r"^NbParticles: ", out, flags=re.M)[1:]  # one block per sweep point
    assert len(blocks) == 12, out[-2000:]
    points = []
    for b in blocks:
        n = int(b.split("\n", 1)[0])
        box_div = int(re.search(r"BoxDiv: (\d+)", b).group(1))
        loops = [(int(a), int(c), float(e)) for a, c, e in re.findall(r" - loop (\d+): candidates (\d+) e_total (\S+)", b)]
        total = int(re.search(r" - Nb interactions: (\d+)", b).group(1))
        assert [l for l, _, _ in loops] == [0, 1, 2, 3, 4]
        assert total == sum(c for _, c, _ in loops)
        cell = 1.0 / box_div
        p = po.params(np.float32(cell), sigma=1.0, epsilon=1.0, softening2=1e-5, predicate="cudalj")
        for l, cand, e_tot in loops:
            x, y, z = _uniform(1234, l, n)
            off, _, _ = po.bin(p, x, y, z)
            assert cand == po.candidates(p, off), (n, box_div, l)
            e_ref = 0.5 * po.cudalj_energy(p, x, y, z).sum()
            assert e_tot == pytest.approx(e_ref, rel=1e-9), (n, box_div, l)
        points.append((n, box_div, total))
    files = list(tmp_path.glob("results-*.csv"))
    assert len(files) == 1
    with open(files[0]) as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["NbParticles", "NbInteractions", "NbLoops", "boxDiv", "nbCells", "partspercell",
                       "interactionsperparticle", "timeinit", "timecompute", "timetotal"]
    assert len(rows) == 1 + 5 * len(points)
    for q, (n, box_div, total) in enumerate(points):
        for r in rows[1 + 5 * q: 6 + 5 * q]:
            assert (int(r[0]), int(r[1]), int(r[2]), int(r[3]), int(r[4])) == (n, total, 5, box_div, box_div**3)
            # the reference writes these with the stream's default 6 significant digits
            assert float(r[5]) == pytest.approx(n / box_div**3, rel=1e-5)
            assert float(r[6]) == pytest.approx(total / n, rel=1e-5)
            assert float(r[9]) == pytest.approx(float(r[7]) + float(r[8]), rel=1e-5)


def test_ljbench_aabb_checked_run_matches_oracle(tmp_path):
    """`ljbench aabb -ns 200 -bd 8`: the AABB driver's checked run (N = 200, rc = 1/8,
    particles pre-sorted by cell, sigma = eps = 1).  Each of the 100 printed energies
    agrees with the driver's own FP64 brute force and with the FP64 oracle at the same
    position (the printed positions are exact FP32 values)."""
    out = _run(["aabb", "-ns", "200", "-bd", "8"], tmp_path)
    lines = re.findall(r"Energy for particle (\d+) position (\S+) (\S+) (\S+) is (\S+) it has been computed as (\S+)", out)
    assert len(lines) == 100
    x, y, z = _uniform(48, 0, 200)
    p = po.params(np.float32(1.0 / 8), sigma=1.0, epsilon=1.0)
    ref = po.forces_f64(p, x, y, z)
    where = {(a, b, c): i for i, (a, b, c) in enumerate(zip(x.tolist(), y.tolist(), z.tolist()))}
    for k, px, py, pz, brute, gpu in lines:
        i = where[(float(np.float32(px)), float(np.float32(py)), float(np.float32(pz)))]
        assert float(gpu) == pytest.approx(ref["e"][i], rel=2e-6, abs=1e-30), k
        assert float(gpu) == pytest.approx(float(brute), rel=2e-6, abs=1e-30), k


def test_ljbench_aabb_sweep_csv(tmp_path):
    """`ljbench aabb -rb -us --max-boxdiv 2`: the AABB sweep with the cell-order pre-sort
    for boxDiv 2 (N = 8..256, five loops each) writes the reference's CSV
    (results-spherecustomsorted-*.csv) with NbInteractions = N (N / nbBoxes) 27
    (optixSphere.cpp:741)."""
    _run(["aabb", "-rb", "-us", "--max-boxdiv", "2"], tmp_path)
    files = list(tmp_path.glob("results-spherecustomsorted-*.csv"))
    assert len(files) == 1
    with open(files[0]) as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["NbParticles", "NbInteractions", "NbLoops", "boxDiv", "nbCells", "partspercell", "timeinit",
                       "timecompute", "timetotal"]
    ns = [8 * 2**k for k in range(6)]
    assert [int(r[0]) for r in rows[1:]] == [n for n in ns for _ in range(5)]
    for r in rows[1:]:
        n = int(r[0])
        assert int(r[1]) == n * (n // 8) * 27 and int(r[3]) == 2 and int(r[4]) == 8
        assert all(math.isfinite(float(v)) and float(v) >= 0 for v in r[6:])


@pytest.mark.parametrize("prog,count_flag", [("custom", "-np"), ("sphere", "-ns"), ("triangle", "-np")])
def test_ljbench_program_checked_runs_match_oracle(tmp_path, prog, count_flag):
    """`ljbench custom|sphere|triangle -np/-ns 300 -bd 8`: the checked runs of the
    Custom, Sphere and Triangle drivers (SDK/optixMultiCustom/optixMultiCustom.cpp:
    593-616 and the Sphere / Triangle copies) with each program's pair rule.  The 100
    printed energies equal the FP64 oracle's under the same rule at the same positions,
    and the driver's host brute force (dist < c, no ray geometry), which the rules agree
    with on these generic positions."""
    out = _run([prog, count_flag, "300", "-bd", "8"], tmp_path)
    lines = re.findall(r"Energy for particle (\d+) position (\S+) (\S+) (\S+) is (\S+) it has been computed as (\S+)", out)
    assert len(lines) == 100
    x, y, z = _uniform(48, 0, 300)
    p = po.params(np.float32(1.0 / 8), sigma=1.0, epsilon=1.0, predicate=prog)
    ref = po.forces_f64(p, x, y, z)
    for k, px, py, pz, brute, gpu in lines:
        i = int(k)  # no pre-sort in these drivers: particle k is input k
        assert (np.float32(px), np.float32(py), np.float32(pz)) == (x[i], y[i], z[i])
        assert float(gpu) == pytest.approx(ref["e"][i], rel=2e-6, abs=1e-30), k
        assert float(gpu) == pytest.approx(float(brute), rel=2e-6, abs=1e-30), k


def test_ljbench_triangle_sweep_csv(tmp_path):
    """`ljbench triangle -rb --max-boxdiv 2` writes results-triangle-*.csv with the
    reference's columns and one row per loop."""
    _run(["triangle", "-rb", "--max-boxdiv", "2"], tmp_path)
    files = list(tmp_path.glob("results-triangle-*.csv"))
    assert len(files) == 1
    with open(files[0]) as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["NbParticles", "NbInteractions", "NbLoops", "boxDiv", "nbCells", "partspercell", "timeinit",
                       "timecompute", "timetotal"]
    assert [int(r[0]) for r in rows[1:]] == [n for n in (8 * 2**k for k in range(6)) for _ in range(5)]
