"""Neighbour-list reuse (Verlet skin): forces over a stored list against the cell grid
and the FP64 oracle, the displacement test, and a short velocity-Verlet run in which the
reused list must find exactly the cell grid's pairs at every step."""
import numpy as np
import pytest

import ljamd
import pyoracle as po

pytestmark = pytest.mark.gpu


def _list_params(p, skin):
    q = ljamd.make_params(float(p.rc) + skin, sigma=float(p.sigma), box_lo=tuple(p.box_lo), box_hi=tuple(p.box_hi))
    return q


@pytest.mark.parametrize("lbvh_list", [False, True])
def test_list_forces_match_cellgrid_after_moves(engine, lbvh_list):
    rng = np.random.default_rng(11)
    n, rc, skin = 100_000, 0.04, 0.008
    x, y, z = ljamd.generate("uniform", 4321, n)
    p = ljamd.make_params(rc)
    dx, dy, dz = engine.to_device(x, y, z)
    rows, cols = engine.pairs(dx, dy, dz, _list_params(p, skin), lbvh=lbvh_list)
    # every particle moves by less than skin / 2
    step = skin / 2 / np.sqrt(3) * 0.999
    x1, y1, z1 = (np.clip(a + rng.uniform(-step, step, n).astype(np.float32), 0.0, np.float32(0.99999994))
                  .astype(np.float32) for a in (x, y, z))
    ex, ey, ez = engine.to_device(x1, y1, z1)
    print(f"[list] lbvh={lbvh_list} entries={cols.numel()}", flush=True)
    out = engine.compute_list(ex, ey, ez, p, rows, cols)
    ref_gpu = engine.compute(ex, ey, ez, p)
    engine.torch.cuda.synchronize()
    res = ljamd.read_results(out["result"])[0]
    res_c = ljamd.read_results(ref_gpu["result"])[0]
    assert int(res["n_pairs"]) == int(res_c["n_pairs"])
    assert int(res["n_candidates"]) == cols.numel()
    ref = po.forces_f64(po.from_lj(p), x1, y1, z1)
    assert int(res["n_pairs"]) == ref["n_pairs"]
    f = [out[k][:n].cpu().numpy() for k in ("fx", "fy", "fz")]
    err = po.rel_errors(f, (ref["fx"], ref["fy"], ref["fz"]))["rel"]
    assert err.max() <= 1e-5
    assert res["e_total"] == pytest.approx(ref["e_total"], rel=1e-9)
    assert res["virial"] == pytest.approx(ref["virial"], rel=1e-9)


def test_max_displacement(engine):
    rng = np.random.default_rng(5)
    n = 50_001
    x0, y0, z0 = (rng.uniform(0, 1, n).astype(np.float32) for _ in range(3))
    d = rng.normal(0, 1e-3, (3, n)).astype(np.float32)
    x, y, z = x0 + d[0], y0 + d[1], z0 + d[2]
    t = engine.to_device(x, y, z, x0, y0, z0)
    m = engine.max_displacement2(*t)
    engine.torch.cuda.synchronize()
    ddx, ddy, ddz = (a.astype(np.float64) - b.astype(np.float64) for a, b in ((x, x0), (y, y0), (z, z0)))
    want = float(np.max(ddx * ddx + ddy * ddy + ddz * ddz))
    assert float(m.cpu()[0]) == pytest.approx(want, rel=1e-5)
    # no displacement: exactly zero
    m0 = engine.max_displacement2(t[3], t[4], t[5], t[3], t[4], t[5])
    engine.torch.cuda.synchronize()
    assert float(m0.cpu()[0]) == 0.0


def _md(engine, use_list, steps=200):
    """Velocity Verlet, sigma = epsilon = mass = 1: a 16^3 simple-cubic lattice at
    spacing 1.12 with small thermal velocities (T ~ 0.05), rc = 2.37.  With use_list the
    forces come from a list built at rc + skin and rebuilt only when some particle has
    moved more than skin / 2; its pairs must equal the cell grid's at every step.
    (The potential is truncated without a shift, as in the reference, so the energy
    jumps whenever a pair crosses rc; the free crystal relaxes and such crossings occur,
    which is why the check is against the cell-grid trajectory, not energy
    conservation.)"""
    torch = engine.torch
    rng = np.random.default_rng(3)
    m_side, a = 16, 1.12
    g = (np.arange(m_side) * a + 10.0).astype(np.float32)
    X, Y, Z = (v.ravel().astype(np.float32) for v in np.meshgrid(g, g, g, indexing="ij"))
    n = X.size
    rc, skin, dt = 2.37, 0.05, 0.002
    p = ljamd.make_params(rc, sigma=1.0, box_lo=(0.0, 0.0, 0.0), box_hi=(40.0, 40.0, 40.0))
    pl = _list_params(p, skin)
    x, y, z = (t.clone() for t in engine.to_device(X, Y, Z))
    v = [torch.from_numpy(rng.normal(0, np.sqrt(0.05), n).astype(np.float32)).cuda() for _ in range(3)]
    for c in range(3):
        v[c] -= v[c].mean()
    state = {"rebuilds": 0}

    def build():
        rows, cols = engine.pairs(x, y, z, pl)
        state.update(rows=rows, cols=cols, x0=x.clone(), y0=y.clone(), z0=z.clone())

    def forces():
        if use_list:
            m2 = float(engine.max_displacement2(x, y, z, state["x0"], state["y0"], state["z0"]).cpu()[0])
            if m2 > (skin / 2) ** 2:
                build()
                state["rebuilds"] += 1
            out = engine.compute_list(x, y, z, p, state["rows"], state["cols"])
            chk = engine.compute(x, y, z, p)
            torch.cuda.synchronize()
            assert int(ljamd.read_results(out["result"])[0]["n_pairs"]) == int(ljamd.read_results(chk["result"])[0]["n_pairs"])
        else:
            out = engine.compute(x, y, z, p)
            torch.cuda.synchronize()
        return [out["fx"][:n], out["fy"][:n], out["fz"][:n]], float(ljamd.read_results(out["result"])[0]["e_total"])

    if use_list:
        build()
    f, epot = forces()
    energies = []
    for _ in range(steps):
        for c, pos in enumerate((x, y, z)):
            v[c] += 0.5 * dt * f[c]
            pos += dt * v[c]
        f, epot = forces()
        for c in range(3):
            v[c] += 0.5 * dt * f[c]
        energies.append(0.5 * float(sum((vc.double() ** 2).sum() for vc in v)) + epot)
    return np.array(energies), state["rebuilds"], torch.stack([x, y, z]).cpu().numpy()


def test_velocity_verlet_with_list_reuse(engine):
    e_list, rebuilds, pos_list = _md(engine, True)
    e_grid, _, pos_grid = _md(engine, False)
    print(f"[md] rebuilds={rebuilds} E0={e_list[0]:.6f} drift={np.max(np.abs(e_list - e_list[0])):.3e} "
          f"max|x_list - x_grid|={np.max(np.abs(pos_list - pos_grid)):.3e}", flush=True)
    assert 1 <= rebuilds < len(e_list)
    # same dynamics as the cell grid every step: the pair sets are equal, only the
    # summation order of each force differs
    assert np.max(np.abs(pos_list - pos_grid)) <= 1e-4
    assert np.max(np.abs(e_list - e_grid)) <= 1e-6 * abs(e_grid[0])
    # over the first steps, before the relaxing crystal brings pairs across rc, the
    # integrator conserves the energy
    assert np.max(np.abs(e_list[:5] - e_list[0])) <= 1e-7 * abs(e_list[0])
