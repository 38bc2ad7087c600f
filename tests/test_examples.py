"""The reference's adjoint and inversion examples through the INI runner and
the fd2d plugin, checked against the oracle.

* examples/adjoint (inherits examples/forward: SH, 201 x 201 cells of
  2.4 km, nt 4800, dt 0.06 s, 12 sources, 132 stations; combine_sources no,
  adjoint_interval 10, smooth 5): traces of the true model are generated by
  the plugin (import_traces without a traces path), the gradient at the
  initial model is written as output/proc000000_kmu.bin and compared with the
  oracle's run_kernel (solver.py:332-404), smoothing included.
* examples/inverse-checker (inherits examples/adjoint; L-BFGS with the
  example's line-search settings): the observed traces are read from raw
  float32 vy_%06d.npy files (fd2d.py:363-373), here written from the oracle's
  forward runs of the true model, because the files the reference ships were
  made with another model or geometry (tests/test_oracle_pins.py records
  that) and are 30 MB.  Two iterations run through the plugin on the GPU
  (compute_gradient, compute_misfit in every line-search trial, update_model)
  and, separately, with the same optimizer driving the oracle; the misfit
  histories and the final mu models must agree.

Inputs come from tests/golden/forward_example.npz (the reference's example
files, see tests/golden/make_forward_example.py); niter is the only value
changed (10 -> 2) to keep the oracle's half of the test short.
"""
import io
import os

import numpy as np
import pytest

from oracle import oracle as O

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "forward_example.npz")


def _write_bin(path, npt, a):
    with open(path, "wb") as f:
        np.array([npt], np.int32).tofile(f)
        np.asarray(a, np.float32).tofile(f)


def materialize(root, niter=None):
    """examples/forward, examples/adjoint and examples/inverse-checker under
    root, from the fixture; returns (dirs dict, fixture)."""
    d = np.load(FIX)
    npt = d["vp"].size
    fw = os.path.join(root, "forward")
    for sub, pre in (("model_true", ""), ("model_init", "init_")):
        os.makedirs(os.path.join(fw, sub))
        for name in ("vp", "vs", "rho"):
            _write_bin(os.path.join(fw, sub, "proc000000_%s.bin" % name), npt, d[pre + name])
        for name in ("x", "z"):
            _write_bin(os.path.join(fw, sub, "proc000000_%s.bin" % name), npt, d[name])
    for k, fn in (("config_ini", "config.ini"), ("sources_dat", "sources.dat"), ("stations_dat", "stations.dat")):
        open(os.path.join(fw, fn), "w").write(str(d[k]))
    dirs = {"forward": fw}
    for ex in ("adjoint", "inverse-checker"):
        wd = os.path.join(root, ex)
        os.makedirs(wd)
        cfg = str(d[ex.replace("-", "_") + "_config_ini"])
        if ex == "inverse-checker" and niter is not None:
            assert "niter = 10" in cfg
            cfg = cfg.replace("niter = 10", "niter = %d" % niter)
        open(os.path.join(wd, "config.ini"), "w").write(cfg)
        dirs[ex] = wd
    return dirs, d


class Example:
    """The forward example's grid, geometry and models in the oracle's
    terms (fd2d.py:122-252 restated by oracle.py)."""

    def __init__(self, d, nt=4800, dt=0.06):
        self.nx, self.nz, self.dx, self.dz = O.grid_shape(d["x"], d["z"], np.array([d["x"].size], np.int32))
        self.nt, self.dt = nt, dt
        src = np.loadtxt(io.StringIO(str(d["sources_dat"])), ndmin=2)
        rec = np.loadtxt(io.StringIO(str(d["stations_dat"])), ndmin=2)
        self.nsrc = src.shape[0]
        self.src_id = O.grid_index(src, self.dx, self.dz, self.nz)
        self.rec_id = O.grid_index(rec, self.dx, self.dz, self.nz)
        self.stf = O.make_stf(src, nt, dt)
        self.bound = O.set_bound(self.nx, self.nz, 20, 0.015, 1, 1, 1, 0)
        self.true = self.model(d["vp"], d["vs"], d["rho"])
        self.init = self.model(d["init_vp"], d["init_vs"], d["init_rho"])

    @staticmethod
    def model(vp, vs, rho):
        lam, mu = O.vps2lm(vp, vs, rho)
        return {"rho": np.asarray(rho, np.float32), "lam": lam, "mu": mu}

    def sim(self, model, isrc=0):
        return O.Sim(O.SH, self.nx, self.nz, self.dx, self.dz, self.dt, self.nt, model, self.bound,
                     self.src_id, self.rec_id, isrc)

    def traces(self):
        """Records of the true model, one source at a time."""
        sim = self.sim(self.true)
        out = []
        for i in range(self.nsrc):
            sim.isrc = i
            out.append({"y": sim.forward(self.stf)[0]["y"]})
        return out


class _HostArray:
    def __init__(self, owner):
        self._owner = owner

    def copy_to_host(self, ary=None, stream=None):
        out = self._owner.model["mu"].copy()
        if ary is None:
            return out
        ary[...] = out
        return ary


class _NoStream:
    def synchronize(self):
        pass


class OracleSolver:
    """The fd2d plugin calls the optimizer makes (compute_gradient,
    compute_misfit, update_model, mu, stream, export_field), answered by
    the oracle's run_kernel on the CPU."""

    def __init__(self, ex, data, sae=10, sigma=5):
        self.ex, self.data, self.sae, self.sigma = ex, data, sae, sigma
        self.nx, self.nz = ex.nx, ex.nz
        self.model = {k: v.copy() for k, v in ex.init.items()}
        self.mu = _HostArray(self)
        self.stream = _NoStream()
        self.exported = {}

    def _run(self, adj):
        sim = self.ex.sim(self.model)
        return O.run_kernel(sim, self.ex.stf, self.data, self.sae, adj=adj, sigma=self.sigma if adj else 0)

    def compute_gradient(self):
        misfit, K = self._run(True)
        return K["mu"].copy(), np.float32(misfit), self.model["mu"].copy()

    def compute_misfit(self):
        return np.float32(self._run(False)[0])

    def update_model(self, mu):
        self.model["mu"] = np.ascontiguousarray(mu, np.float32).copy()

    def export_field(self, field, name, it=0):
        self.exported[(name, it)] = np.array(field, np.float32)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_fixture_holds_the_adjoint_and_inversion_examples():
    d = np.load(FIX)
    adj = str(d["adjoint_config_ini"])
    inv = str(d["inverse_checker_config_ini"])
    assert "inherit = forward" in adj and "adjoint_interval = 10" in adj and "smooth = 5" in adj
    assert "combine_sources = no" in adj
    assert "inherit = adjoint" in inv and "method = lbfgs" in inv and "traces = traces" in inv
    assert d["init_vs"].size == d["vs"].size and np.ptp(d["init_vs"]) == 0 and np.ptp(d["vs"]) > 0


def test_inverse_checker_config_resolves_through_two_inherits(tmp_path):
    from seispie_amd import sprun
    dirs, _ = materialize(str(tmp_path))
    config, _ = sprun.parse_config(dirs["inverse-checker"], "config.ini")
    assert config["workflow"]["mode"] == "inversion"
    assert config["solver"]["adjoint_interval"] == "10" and config["solver"]["sh"] == "yes"
    assert config["optimize"]["lbfgs_mem"] == "5"
    assert config["path"]["traces"] == "traces"
    assert config["path"]["sources"] == "../adjoint/../forward/sources.dat"


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_reference_adjoint_example_matches_the_oracle(tmp_path):
    from seispie_amd import sprun
    dirs, d = materialize(str(tmp_path))
    print("case: reference examples/adjoint through sprun (12 shots, 201^2, nt 4800)", flush=True)
    wf = sprun.main(["--workdir", dirs["adjoint"]])
    s = wf.solver
    assert (s.nx, s.nz, s.nt, s.nsrc, s.sae, s.sigma) == (201, 201, 4800, 12, 10, 5)
    kmu = np.fromfile(os.path.join(dirs["adjoint"], "output", "proc000000_kmu.bin"), np.float32)
    assert kmu[:1].view(np.int32)[0] == 201 * 201
    # generated traces: raw float32, one file per source (fd2d.py:393-409)
    ex = Example(d)
    data = ex.traces()
    for i in (0, 11):
        got = np.fromfile(os.path.join(dirs["adjoint"], "output", "traces", "vy_%06d.npy" % i), np.float32)
        np.testing.assert_array_equal(got, data[i]["y"].reshape(-1))
    print("traces equal the oracle's; running the oracle's run_kernel", flush=True)
    misfit, K = O.run_kernel(ex.sim(ex.init), ex.stf, data, 10, adj=True, sigma=5)
    assert np.abs(K["mu"]).max() > 0 and misfit > 0
    err = _rel(kmu[1:], K["mu"])
    print("kmu rel L2 %.2e" % err)
    assert err <= 1e-4


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_reference_inversion_from_trace_files_matches_the_oracle(tmp_path):
    from seispie_amd import sprun
    dirs, d = materialize(str(tmp_path), niter=2)
    ex = Example(d)
    data = ex.traces()
    tr = os.path.join(dirs["inverse-checker"], "traces")
    os.makedirs(tr)
    for i, rec in enumerate(data):
        rec["y"].astype(np.float32).reshape(-1).tofile(os.path.join(tr, "vy_%06d.npy" % i))
    print("case: reference examples/inverse-checker, 2 L-BFGS iterations from trace files", flush=True)
    wf = sprun.main(["--workdir", dirs["inverse-checker"]])
    gpu_hist = [m for _, m in wf.optimize.history]
    s = wf.solver
    assert len(s.syn_y) == ex.nsrc and not os.path.exists(os.path.join(dirs["inverse-checker"], "output", "traces"))
    out = os.path.join(dirs["inverse-checker"], "output")
    mu_gpu = np.fromfile(os.path.join(out, "proc000002_mu.bin"), np.float32)[1:]
    print("GPU misfits", gpu_hist, flush=True)

    from seispie_amd.optimize.lbfgs import lbfgs
    from seispie_amd.sprun import parse_config

    class _Wf:
        pass

    config, _ = parse_config(dirs["inverse-checker"], "config.ini")
    solver = OracleSolver(ex, data, sae=10, sigma=5)
    wfo = _Wf()
    wfo.solver = solver
    opt = lbfgs()
    opt.config = config["optimize"]
    opt.mpi = None
    opt.quiet = True
    opt.setup(wfo)
    opt.run()
    cpu_hist = [m for _, m in opt.history]
    print("oracle misfits", cpu_hist, flush=True)
    assert len(gpu_hist) == len(cpu_hist) == 3
    assert gpu_hist[-1] < gpu_hist[0]
    for a, b in zip(gpu_hist, cpu_hist):
        assert abs(a - b) <= 1e-4 * abs(b), (a, b)
    mu_cpu = solver.exported[("mu", 2)]
    err = _rel(mu_gpu, mu_cpu)
    print("final mu rel L2 %.2e; update rel L2 %.2e" % (err, _rel(mu_gpu - ex.init["mu"], mu_cpu - ex.init["mu"])))
    assert err <= 1e-6
    assert _rel(mu_gpu - ex.init["mu"], mu_cpu - ex.init["mu"]) <= 1e-3
