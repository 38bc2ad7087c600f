"""Benchmark: batched P-SV forward + adjoint gradient on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 3a|2|1|3b|4|5]

One "step" is one gradient pass of the workload: forward run of every shot
(with adjoint snapshots), misfit + adjoint sources, adjoint run with kernel
accumulation.  Inputs (model, sources, observed data) are resident on the
GPU before the timed region.  Throughput = sum over shots of nx*nz*nt*2
cell-updates / wall time (forward and adjoint both counted, nothing
recomputed).

The JSON line adds
  roofline:     the dominant kernel (the step kernel, which advances two
                time steps per launch where it can) against the 8 TB/s HBM
                peak: algorithmic bytes per launch (per time step, 8 B per
                state field per shot cell + 4 B per material field per grid
                cell, times the steps per launch) over its mean launch time
                from HIP events on the library's stream, in a forward pass
                timed on its own; `traffic` is the measured HBM bytes per
                launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/);
  cpu_baseline: the oracle's unfused reference kernel sequence (one shot,
                a fixed window of steps on the same grid) on this host's cores;
  time_to_misfit: the second half of the metric, on the learned-conditioner
                FWI workload (P-SV 500 x 500, 64 shots, nt 3000, inverting
                vp): plain L-BFGS for 20 iterations, then L-BFGS with the
                learned step conditioner (trained inside its run on the
                first gradient pass); the wall time the conditioned run needs
                to reach the plain run's final misfit over the plain run's
                wall time (--no-fwi skips it).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12


def build_problem(name):
    from seispie_amd import configs
    if name == "3a":
        true = configs.psv_config3()
        init = configs.psv_config3()
        init.grids[0].fields = [np.full_like(f, f.mean()) for f in true.grids[0].fields]
    elif name == "3b":
        true = configs.psv_config3_mixed()
        init = configs.psv_config3_mixed()
        for g in init.grids:
            g.fields = [np.full_like(f, f.mean()) for f in g.fields]
    elif name == "2":
        true = configs.psv_config2(true_model=True)
        init = configs.psv_config2(true_model=False)
    elif name == "1":
        true = configs.sh_config1()
        init = configs.sh_config1()
        init.grids[0].fields = [np.full_like(f, f.max()) for f in true.grids[0].fields]
    elif name == "4":
        true = configs.spin_config4()
        init = configs.spin_config4()
        g = init.grids[0]
        g.fields = [np.full_like(f, f.mean()) for f in g.fields]
    elif name == "5":
        true = configs.fwi_config5(true_model=True)
        init = configs.fwi_config5(true_model=False)
    else:
        raise SystemExit("unknown config %s" % name)
    return true, init


def observed_data(true, chunk=0, order=0, fast=False):
    """Records of the true model (same work decomposition and arithmetic as
    the timed runs, so a kernel trace of the command sees one setting)."""
    from seispie_amd import _lib, configs
    ctx = _lib.Context(0)
    if chunk:
        ctx.set_option(_lib.OPT_CHUNK, chunk)
        ctx.set_option(_lib.OPT_ORDER, order)
    if fast:
        ctx.set_option(_lib.OPT_EXACT, 0)
    ids = configs.load_problem(ctx, true)
    ctx.forward(0)
    data = [ctx.records(i) for i in ids]
    ctx.close()
    return data


def host_cpu():
    """This host's CPUs: nproc, the CPUs this process may run on (affinity
    and the cgroup's CPU quota), and the lscpu model name."""
    import math
    import subprocess
    info = {"nproc": os.cpu_count() or 1}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["affinity"] = info["nproc"]
    usable = info["affinity"]
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            info["cgroup_cpus"] = round(int(quota) / int(period), 2)
            usable = min(usable, max(1, math.floor(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    info["usable"] = usable
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        model = next((ln.split(":", 1)[1].strip() for ln in out.splitlines() if ln.startswith("Model name")), None)
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            model = next(ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name"))
        except (OSError, StopIteration):
            model = "unknown"
    info["model"] = model
    return info


def baseline_build():
    """The CPU baseline build: oracle.c with gcc -O3 -march=native -fopenmp,
    compiled for this host into a temporary directory (a few seconds); the
    prebuilt x86-64-v3 build if that fails."""
    import subprocess
    import tempfile
    out = tempfile.mkdtemp(prefix="sp_baseline_")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "baseline", "BASELINE_OUT=" + out],
                       check=True, capture_output=True, timeout=300)
        return os.path.join(out, "libbaseline_native.so"), "gcc -O3 -march=native -fopenmp, built on this host"
    except (OSError, subprocess.SubprocessError):
        path = os.path.join(ROOT, "oracle", "_build", "libbaseline_v3.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "portable"], check=True)
        return path, "gcc -O3 -march=x86-64-v3 -fopenmp (prebuilt; the native build failed)"


def cpu_baseline(prob, warm=20, nsteps=200):
    """The reference's unfused kernel sequence (oracle.c, solver.py:205-234
    and :290-321) on this host's cores, one shot of the workload's first
    grid: a forward window and an adjoint window of nsteps steps after warm
    steps each, on every usable core and on one core.  Gcell-upd/s counts
    both windows (P = 2), as the GPU figure does."""
    from oracle import oracle as O
    from tests.helpers import oracle_sim, oracle_stf
    g = prob.grids[0]
    shot = g.shots[0]
    sim = oracle_sim(prob, g, shot)
    stf = oracle_stf(prob, shot)
    path, how = baseline_build()
    L = O.baseline_lib(path)
    cpu = host_cpu()
    # adjoint sources: the shot's source time function at every receiver
    nrec = shot.rec_k.size
    ad = [np.tile(np.asarray(c).reshape(-1, prob.nt)[:1], (nrec, 1)).reshape(-1) for c in stf]
    out = {}
    for nth in sorted({cpu["usable"], 1}, reverse=True):
        tf = sim.cpu_bench(stf, warm, nsteps, nth, L)
        ta = sim.cpu_bench_adjoint(ad, prob.sae, warm, nsteps, nth, L)
        cells = g.nx * g.nz * nsteps
        out[nth] = {"fwd_adj": 2 * cells / (tf + ta) / 1e9, "fwd": cells / tf / 1e9, "adj": cells / ta / 1e9,
                    "seconds": tf + ta}
    return cpu, how, out


def autograd_bench(true, init, args):
    """Forward + backward of the waveform misfit through
    seispie_amd.autograd (checkpointed adjoint with recomputation)."""
    import torch
    from seispie_amd import autograd
    if len(init.grids) != 1 or init.physics == _lib_spin():
        raise SystemExit("--autograd needs a one-grid SH or P-SV workload")
    g = init.grids[0]
    data = observed_data(true)
    adj = "discrete" if args.discrete else "reference"
    sim = autograd.Simulation(init.physics, g.nx, g.nz, g.dx, g.dz, init.nt, init.dt,
                              [(s.src_k, s.stf, s.rec_k) for s in g.shots], sponge=g.sponge,
                              sae=1 if args.discrete else init.sae, adjoint=adj)
    dev = sim.device
    d = torch.tensor(np.stack([x[:sim.nmis] for x in data]), device=dev)
    m = [torch.tensor(f, device=dev, requires_grad=True) for f in g.fields]

    def step():
        for t in m:
            t.grad = None
        loss = autograd.waveform_loss(sim(*m), d, init.dt)
        loss.backward()
        return loss

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss = step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = init.cell_updates(passes=2)
    print(json.dumps({
        "metric": "Gcell-upd/s (autograd forward + backward)", "value": round(cells * args.steps / wall / 1e9, 3),
        "unit": "Gcell-upd/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2), "higher_is_better": True, "dtype": "f32 storage, f64 arithmetic",
        "data": "synthetic", "loss": float(loss.detach()),
        "config": {"workload": init.name + " waveform misfit, loss.backward()", "shots": init.nshots, "nt": init.nt,
                   "adjoint": adj, "adjoint_interval": sim.sae, "checkpoint_interval": sim.interval,
                   "note": "cell updates count forward and adjoint once; the forward recomputed from "
                           "checkpoints in backward is not counted"}}))
    sim.close()
    return 0


def _lib_spin():
    from seispie_amd import _lib
    return _lib.SPIN


def fwi_time_to_misfit(niter, surface=False):
    """Time-to-misfit of the learned conditioner against plain L-BFGS; a
    failure is reported in the line instead of ending the bench.  surface:
    the free-surface acquisition (sources at 10 m, receivers at 2 dz, top
    free) with the discrete adjoint and a 20-row gradient mute."""
    from seispie_amd import configs, fwi
    t0 = time.time()
    try:
        true = configs.fwi_config5(true_model=True, surface=surface)
        init = configs.fwi_config5(true_model=False, surface=surface)
        kw = {"adjoint": "discrete", "mute_top": 20} if surface else {}
        r = fwi.time_to_misfit(true, init, niter=niter,
                               log=lambda s: print("# " + s, file=sys.stderr, flush=True), **kw)
    except Exception as e:  # noqa: BLE001 -- reported, not hidden
        return {"error": "%s: %s" % (type(e).__name__, e)}
    g = init.grids[0]
    src_z = float(g.shots[0].src_k[0] % g.nz) * g.dz
    rec_z = float(g.shots[0].rec_k[0] % g.nz) * g.dz
    if surface:
        geometry = ("sources at %g m, receivers at %g m depth, top free, sponges on three sides; discrete adjoint "
                    "(the exact transpose of the time stepping), gradient muted in the top 20 rows with a cosine "
                    "ramp over the next 20" % (src_z, rec_z))
    else:
        geometry = ("sources at %g m, receivers at %g m depth, sponges on all four sides (width %d): the "
                    "reference-style adjoint (forward operator reused, solver.py:308-321) is the misfit's "
                    "gradient only away from the free surface's edge bands, so the acquisition sits below "
                    "the top sponge instead of at 10 m under a free top" % (src_z, rec_z, g.sponge[0]))
    out = {"workload": init.name + " vp inversion, L-BFGS (lbfgs_mem 5, backtracking)", "iterations": niter,
           "geometry": geometry,
           "plain_history": [round(f, 4) for _, f in r["plain_history"]],
           "conditioned_history": [round(f, 4) for _, f in r["learned_history"]],
           "value": None if r["time_to_misfit"] is None else round(r["time_to_misfit"], 4),
           "unit": "fraction of plain L-BFGS wall time, conditioner training included", "target": 0.6,
           "without_training": (None if r["time_to_misfit_without_training"] is None
                                else round(r["time_to_misfit_without_training"], 4)),
           "plain_seconds": round(r["plain_seconds"], 2), "plain_final_misfit": r["plain_final_misfit"],
           "conditioned_seconds": round(r["learned_seconds"], 2),
           "conditioned_final_misfit": r["learned_final_misfit"],
           "conditioner_prep_seconds": round(r["prep_seconds"], 3),
           "conditioner_train_seconds": round(r["train_seconds"], 3),
           "wall_seconds": round(time.time() - t0, 1)}
    return out


CONFIG_BLOCK = ("1", "2", "3b", "4", "5")
SMOOTH_SIGMA = 5  # smooth of the reference's examples/adjoint config


def _context(args):
    from seispie_amd import _lib
    ctx = _lib.Context(0)
    if args.no_fuse:
        ctx.set_option(_lib.OPT_FUSE, 0)
    if args.fuse >= 0:
        ctx.set_option(_lib.OPT_FUSE, args.fuse)
    if args.group:
        ctx.set_option(_lib.OPT_GROUP, args.group)
    if args.fast:
        ctx.set_option(_lib.OPT_EXACT, 0)
    if args.graph:
        ctx.set_option(_lib.OPT_GRAPH, 1)
    if args.chunk:
        ctx.set_option(_lib.OPT_CHUNK, args.chunk)
    ctx.set_option(_lib.OPT_ORDER, args.order)
    return ctx


def forward_only(true, init, args):
    """Forward passes only (the step kernel), for counter runs."""
    from seispie_amd import configs
    ctx = _context(args)
    configs.load_problem(ctx, init)
    for _ in range(args.warmup + args.steps):
        ctx.forward(0)
    ctx.sync()
    t = ctx.timing()
    print(json.dumps({"forward_launch_ms": t["total_ms"] / t["step_launches"],
                      "alg_bytes_per_launch": t["alg_bytes"] / t["step_launches"],
                      "steps_per_launch": round(t["alg_bytes"] / t["step_launches"] / init.alg_bytes_per_step(), 3),
                      "workload": init.name}))
    ctx.close()
    return 0


def time_gradients(true, init, steps, warmup, args, quiet=False):
    """Observed data from the true model, then `steps` timed gradient passes
    (compute_gradient's work: forward with snapshots, misfit, adjoint with
    kernel accumulation, k_mu smoothing) at the initial model after
    `warmup` untimed ones; then one forward pass alone, timed per launch
    with HIP events, for the dominant step kernel's roofline."""
    from seispie_amd import _lib, configs
    say = (lambda m: None) if quiet else (lambda m: print(m, file=sys.stderr, flush=True))
    t0 = time.time()
    data = observed_data(true, args.chunk, args.order, args.fast)
    say("# observed data for %s in %.1fs" % (true.name, time.time() - t0))
    ctx = _context(args)
    ids = configs.load_problem(ctx, init)
    for i, d in zip(ids, data):
        ctx.set_data(i, d)
    del data
    sae = init.sae
    tuned = None
    if not args.chunk:
        ch, od = ctx.autotune(20)
        tuned = {"chunk": ch, "order": od, "steps_per_launch_option": ctx.get_option(_lib.OPT_FUSE),
                 "group": ctx.get_option(_lib.OPT_GROUP)}
        say("# autotuned %s" % tuned)

    def gradient():
        norms = ctx.gradient(sae)
        for gi in range(len(init.grids)):
            ctx.smooth_kernel(gi, _lib.K_MU, SMOOTH_SIGMA)
        return norms

    for w in range(warmup):
        gradient()
        ctx.sync()
        say("# warmup %d done" % w)
    ctx.sync()
    t0 = time.perf_counter()
    for s in range(steps):
        norms = gradient()
        say("# step %d misfit %.6e" % (s, norms.sum()))
    ctx.sync()
    wall = time.perf_counter() - t0

    ctx.forward(0)
    ctx.sync()
    ft = ctx.timing()
    launch_ms = ft["total_ms"] / ft["step_launches"]
    bytes_per_launch = ft["alg_bytes"] / ft["step_launches"]
    steps_per_launch = bytes_per_launch / init.alg_bytes_per_step()
    mode = step_mode(ctx.get_option(_lib.OPT_FUSE), init.physics, [g.nz for g in init.grids])
    wide = mode in (2, 4)
    if steps_per_launch > 1.5:
        kname = {1: "k_step2<%s, W=1> (two time steps per launch)",
                 2: "k_step2<%s, W=2> (two time steps per launch, two cells per lane)",
                 3: "k_step2g<%s, W=1> (two time steps per launch, column groups of %d wavefronts)",
                 4: "k_step2g<%s, W=2> (two time steps per launch, column groups of %d wavefronts, "
                    "two cells per lane)"}[mode]
    else:
        kname = "k_step3<%s>"
    pname = {1: "SH", 2: "P-SV", 3: "spin"}[init.physics]
    kname = kname % ((pname, ctx.get_option(_lib.OPT_GROUP) or 4) if "%d" in kname else pname)
    ctx.close()
    cells = init.cell_updates(passes=2)
    return {"wall": wall, "value": cells * steps / wall / 1e9, "cells": cells, "launch_ms": launch_ms,
            "bytes_per_launch": bytes_per_launch, "achieved": bytes_per_launch / (launch_ms * 1e-3),
            "steps_per_launch": steps_per_launch, "wide": wide, "mode": mode, "kname": kname, "tuned": tuned,
            "pass_frac": round(init.alg_bytes_per_step() * init.nt * 2 * steps / wall / HBM_PEAK, 4)}


def step_mode(fuse, physics, nzs):
    """The two-step launch a batch gets for an SP_OPT_FUSE value: 1, 2
    strips of one or two cells per lane, 3, 4 column groups of one or two
    cells per lane, falling back as the library does."""
    from seispie_amd import _lib
    if fuse == 3:
        return 3
    if fuse == 4 and physics != _lib.SPIN and all(nz % 2 == 0 for nz in nzs):
        return 4
    if fuse >= 2 and physics != _lib.SPIN and all(nz % 2 == 0 for nz in nzs):
        return 2
    return 1 if fuse else 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="3a")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--order", type=int, default=0, help="work order when --chunk is given")
    ap.add_argument("--nt", type=int, default=0, help="override the number of time steps")
    ap.add_argument("--forward-only", action="store_true",
                    help="time forward passes (the step kernel) instead of gradients; for counter runs")
    ap.add_argument("--no-fwi", action="store_true", help="skip the time-to-misfit measurement")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the per-configuration block (configs 1, 2, 3b, 4, 5: three gradients each)")
    ap.add_argument("--fwi-iters", type=int, default=20)
    ap.add_argument("--fwi-surface", action="store_true",
                    help="time-to-misfit at the free-surface acquisition (discrete adjoint, 20-row gradient mute)")
    ap.add_argument("--fwi-only", action="store_true", help="only the time-to-misfit measurement")
    ap.add_argument("--autograd", action="store_true",
                    help="time loss.backward() through the torch.autograd.Function (checkpointed adjoint) instead")
    ap.add_argument("--discrete", action="store_true",
                    help="with --autograd: the discrete adjoint (exact transpose of the time stepping, every step)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the forward and adjoint runs from captured hipGraphs")
    ap.add_argument("--no-fuse", action="store_true",
                    help="one time step per launch (the single-step kernel) instead of two where allowed")
    ap.add_argument("--fuse", type=int, default=-1,
                    help="steps per launch option: 0 one step, 1 two steps (k_step2), 2 two steps with two "
                         "cells per lane (k_step2 with W = 2; SH and P-SV with even nz), 3, 4 two steps in column "
                         "groups with one or two cells per lane (k_step2g; nz <= 512 or 1024)")
    ap.add_argument("--group", type=int, default=0,
                    help="column groups (--fuse 3, 4): wavefronts per block (default 4)")
    ap.add_argument("--fast", action="store_true",
                    help="step kernels in fast mode (reciprocal multiplications and fused multiply-adds) instead "
                         "of the default exact mode, whose records equal the oracle bit for bit")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a plain 256 MiB elementwise kernel whose HBM bytes are known, to scale "
                         "FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(shots would be split across GPUs with one kernel all-reduce per gradient).", file=sys.stderr)
        return 2

    from seispie_amd import _lib, configs
    if args.calibrate:
        import torch
        a = torch.full((64 << 20,), 1.0, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 2.0, out=b)  # reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        del a, b
    true, init = build_problem(args.config)
    if args.nt:
        for p in (true, init):
            p.nt = args.nt
            for g in p.grids:
                for s in g.shots:
                    s.stf = np.ascontiguousarray(s.stf[..., :args.nt])
    if args.autograd:
        return autograd_bench(true, init, args)
    if args.forward_only:
        return forward_only(true, init, args)
    if args.fwi_only:
        print(json.dumps({"time_to_misfit": fwi_time_to_misfit(args.fwi_iters, args.fwi_surface)}))
        return 0
    r = time_gradients(true, init, args.steps, args.warmup, args)
    wall, value, cells, sae = r["wall"], r["value"], r["cells"], init.sae

    traffic = None
    tf = os.path.join(ROOT, "profiles", "step_kernel_traffic.json")
    if os.path.exists(tf):
        t = json.load(open(tf)).get(init.name.rsplit("_nt", 1)[0]
                                    + ({1: "_fused", 2: "_wide", 3: "_group", 4: "_group_wide"}[r["mode"]]
                                       if r["steps_per_launch"] > 1.5 else "")
                                    + ("_fast" if args.fast else ""))
        if t:
            traffic = t["hbm_bytes_per_launch"]

    others = None
    if not args.no_configs and args.config == "3a" and not args.nt:
        others = []
        for name in CONFIG_BLOCK:
            t0 = time.time()
            try:
                ct, ci = build_problem(name)
                c = time_gradients(ct, ci, 3, 1, args, quiet=True)
                others.append({"config": name, "workload": ci.name + " forward+adjoint gradient",
                               "gcell_upd_s": round(c["value"], 3), "ms_per_gradient": round(c["wall"] / 3 * 1e3, 2),
                               "gradients_timed": 3, "shots": ci.nshots, "nt": ci.nt,
                               "grids": [[g.nx, g.nz] for g in ci.grids],
                               "hbm_roofline_frac_of_whole_pass": c["pass_frac"],
                               "tuned": c["tuned"],
                               "dominant_kernel": {"kernel": c["kname"], "launch_ms": round(c["launch_ms"], 4),
                                                   "achieved_GBps": round(c["achieved"] / 1e9, 1),
                                                   "frac": round(c["achieved"] / HBM_PEAK, 4)}})
            except Exception as e:  # noqa: BLE001 -- reported in the line, not hidden
                others.append({"config": name, "error": "%s: %s" % (type(e).__name__, e)})
            print("# config %s done in %.1fs" % (name, time.time() - t0), file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        host, how, cb = cpu_baseline(init)
        n = host["usable"]
        g = init.grids[0]
        cpu = {"value": round(cb[n]["fwd_adj"], 4), "unit": "Gcell-upd/s", "cores": n, "kind": "port",
               "sample": "oracle.c (the reference's unfused kernel sequence, float32 storage / float64 math; %s) "
                         "on %dx%d, one shot: 200 forward steps and 200 adjoint steps (interaction every %d), "
                         "each after 20 warm-up steps, on %d threads; forward %.4f, adjoint %.4f Gcell-upd/s; "
                         "1 core: %.4f forward+adjoint" % (how, g.nx, g.nz, init.sae, n, cb[n]["fwd"], cb[n]["adj"],
                                                           cb[1]["fwd_adj"]),
               "one_core": round(cb[1]["fwd_adj"], 5), "nproc": host["nproc"], "affinity": host["affinity"],
               "cgroup_cpus": host.get("cgroup_cpus"), "cpu_model": host["model"]}

    ttm = None
    if not args.no_fwi and not args.forward_only:
        ttm = fwi_time_to_misfit(args.fwi_iters, args.fwi_surface)

    line = {
        "metric": "Gcell-upd/s (% HBM peak); time-to-misfit, learned vs plain precond",
        "value": round(value, 3),
        "unit": "Gcell-upd/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 storage, f64 arithmetic",
        "data": "synthetic",
        "config": {"workload": init.name + " forward+adjoint gradient",
                   "grids": [[g.nx, g.nz] for g in init.grids], "shots": init.nshots, "nt": init.nt,
                   "adjoint_interval": sae, "cell_updates_per_step": cells,
                   "arithmetic": "fast" if args.fast else "exact (the reference's float64 operation order)",
                   "step": "forward with snapshots, misfit and adjoint sources, adjoint with kernel "
                           "accumulation, smoothing of k_mu (sigma %d), as compute_gradient" % SMOOTH_SIGMA,
                   "tuned": r["tuned"],
                   "hbm_roofline_frac_of_whole_pass": r["pass_frac"]},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved"] / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(r["achieved"] / HBM_PEAK, 4), "traffic": traffic,
                     "kernel": r["kname"], "launch_ms": round(r["launch_ms"], 4),
                     "alg_bytes_per_launch": r["bytes_per_launch"], "steps_per_launch": round(r["steps_per_launch"], 3),
                     "traffic_rate_GBps": round(traffic / (r["launch_ms"] * 1e-3) / 1e9, 1) if traffic else None,
                     "note": ("a launch advances steps_per_launch time steps and is credited their algorithmic "
                              "bytes while it reads and writes the state once; traffic is the HBM bytes it moves "
                              "(rocprofv3 FETCH_SIZE x2 + WRITE_SIZE, profiles/step_kernel_traffic.json)")},
        "cpu_baseline": cpu,
        "time_to_misfit": ttm,
        "configs": others,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
