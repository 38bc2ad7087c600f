"""The reference's own forward example (examples/forward: SH, 201 x 201
grid of 2.4 km cells, nt 4800, 11 sources fired together, 132 stations,
field exports every 800 steps), from the fixture tests/golden/forward_example.npz
(made by tests/golden/make_forward_example.py from the reference's input
files), run unchanged through the INI runner and the fd2d plugin and
checked against the oracle."""
import os

import numpy as np
import pytest

from oracle import oracle as O

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "forward_example.npz")


def _materialize(root):
    d = np.load(FIX)
    wd = os.path.join(root, "forward")
    for sub in ("model_true", "model_init", "output"):
        os.makedirs(os.path.join(wd, sub))
    npt = np.array([d["vp"].size], np.int32)
    for name in ("x", "z", "vp", "vs", "rho"):
        for sub in ("model_true", "model_init"):
            with open(os.path.join(wd, sub, "proc000000_%s.bin" % name), "wb") as f:
                npt.tofile(f)
                d[name].astype(np.float32).tofile(f)
    for k, fn in (("config_ini", "config.ini"), ("sources_dat", "sources.dat"), ("stations_dat", "stations.dat")):
        open(os.path.join(wd, fn), "w").write(str(d[k]))
    return wd, d


def test_fixture_holds_the_reference_example():
    d = np.load(FIX)
    nx, nz, dx, dz = O.grid_shape(d["x"], d["z"], np.array([d["x"].size], np.int32))
    assert (nx, nz, float(dx), float(dz)) == (201, 201, 2400.0, 2400.0)
    cfg = str(d["config_ini"])
    assert "nt = 4800" in cfg and "save_snapshot = 800" in cfg and "combine_sources = yes" in cfg
    assert len(str(d["sources_dat"]).split("\n")) >= 11 and len(str(d["stations_dat"]).strip().split("\n")) == 132


@pytest.mark.gpu
def test_reference_forward_example_runs_unchanged(tmp_path):
    from seispie_amd import sprun
    wd, d = _materialize(str(tmp_path))
    print("case: reference examples/forward through sprun", flush=True)
    wf = sprun.main(["--workdir", wd])
    s = wf.solver
    assert (s.nx, s.nz, s.nt) == (201, 201, 4800)
    lam, mu = O.vps2lm(d["vp"], d["vs"], d["rho"])
    bound = O.set_bound(s.nx, s.nz, 20, 0.015, 1, 1, 1, 0)
    src = np.loadtxt(os.path.join(wd, "sources.dat"), ndmin=2)
    stf = O.make_stf(src, s.nt, s.dt)
    m = {"rho": d["rho"], "lam": lam, "mu": mu}
    sim = O.Sim(O.SH, s.nx, s.nz, s.dx, s.dz, s.dt, s.nt, m, bound, s.src_id, s.rec_id, -1)
    ref = sim.forward(stf)[0]["y"]
    got = s.obs_y.reshape(s.nrec, s.nt)
    assert np.abs(ref).max() > 0
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("records rel L2 %.2e" % err)
    assert err <= 1e-5
    # save_snapshot = 800: vy after step 800 (and 1600 ... 4000) exported
    for it in range(800, 4800, 800):
        assert os.path.exists(os.path.join(wd, "output", "proc%06d_vy.bin" % it))
    vy = np.fromfile(os.path.join(wd, "output", "proc000800_vy.bin"), np.float32)[1:]
    short = O.Sim(O.SH, s.nx, s.nz, s.dx, s.dz, s.dt, 801, m, bound, s.src_id, s.rec_id, -1)
    _, snaps = short.forward([x.reshape(-1, s.nt)[:, :801].reshape(-1) for x in stf], sae=801)
    want = np.asarray(snaps).reshape(2, -1)[1]
    assert np.abs(want).max() > 0
    assert np.linalg.norm(vy - want) <= 1e-5 * np.linalg.norm(want)
