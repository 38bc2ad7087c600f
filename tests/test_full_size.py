"""Checks at the benchmark's full sizes, where the oracle is too slow:
size-independent properties that hold bit for bit.

* Batching: every shot of the 128-shot 1000^2 batch (and of the mixed
  64-shot batch on three grids) records exactly what it records alone;
  the batch runs the two-step kernel with two cells per lane, the single
  shots the one with one cell per lane.
* Linearity: scaling every source time function by 2 scales every record
  by 2.  A power-of-two scale commutes with every float32 and float64
  rounding except in the subnormal range: the wavefront's leading edge
  starts at subnormal amplitudes, where float32 rounds to a fixed grid, and
  the relative error made there (~1e-7) travels with the front and grows
  with multiple scattering (the CPU oracle shows the same: its float64
  replay is exactly linear, its float32 mode is not; 5e-7 after 400 steps
  on 100^2).  After 4000 steps the worst of the 128 shots differs by
  1.1e-5 rel L2, so the test bounds it by 1e-4: a wrong source scaling or
  a non-linear term would be off by O(1).
"""
import numpy as np
import pytest

from seispie_amd import _lib, configs

pytestmark = pytest.mark.gpu


def _records(ctx_factory, prob, scale=1.0, **opts):
    ctx = ctx_factory()
    for k, v in opts.items():
        ctx.set_option(getattr(_lib, "OPT_" + k.upper()), v)
    for g in prob.grids:
        for s in g.shots:
            s.stf = (s.stf * np.float32(scale)).astype(np.float32)
    ids = configs.load_problem(ctx, prob)
    ctx.forward(0)
    out = [ctx.records(i) for i in ids]
    ctx.close()
    return out


def _single(prob, g, shot):
    return configs.Problem(prob.physics, prob.nt, prob.dt,
                           [configs.Grid(g.nx, g.nz, g.dx, g.dz, g.fields, g.param, g.sponge, [shot])], sae=prob.sae)


@pytest.mark.parametrize("name,fuse,chunk", [("3a", 2, 128), ("3b", 2, 128), ("3a", 4, 64), ("3b", 4, 48)])
def test_batched_shots_equal_single_runs_at_full_size(ctx_factory, name, fuse, chunk):
    """fuse 2: strips of two cells per lane; fuse 4: column groups of two
    cells per lane (3b: five wavefronts per block, the 500- and 400-deep
    grids leaving one idle)."""
    prob = configs.psv_config3() if name == "3a" else configs.psv_config3_mixed()
    print("case: full-size batch", prob.name, prob.nshots, "shots, fuse", fuse, flush=True)
    batch = _records(ctx_factory, prob, fuse=fuse, chunk=chunk)
    flat = [(g, s) for g in prob.grids for s in g.shots]
    for k in (0, len(flat) // 2 + 1, len(flat) - 1):
        g, s = flat[k]
        alone = _records(ctx_factory, _single(prob, g, s))[0]
        assert np.abs(alone).max() > 0
        np.testing.assert_array_equal(batch[k], alone, err_msg="shot %d" % k)


def test_records_are_linear_in_the_source_at_full_size_within_1e4(ctx_factory):
    base = configs.psv_config3()
    print("case: full-size linearity", base.name, flush=True)
    r1 = _records(ctx_factory, base)
    r2 = _records(ctx_factory, configs.psv_config3(), scale=2.0)
    worst = 0.0
    for a, b in zip(r1, r2):
        a2 = 2.0 * a.astype(np.float64)
        worst = max(worst, np.linalg.norm(b - a2) / np.linalg.norm(a2))
    print("worst rel L2 of records(2 stf) - 2 records(stf): %.2e" % worst)
    assert worst <= 1e-4


def test_one_shot_of_the_benchmark_matches_the_oracle(ctx_factory):
    """Shot 0 of the 128-shot 1000^2 workload over all 4000 steps against
    the oracle (float32 storage, float64 arithmetic; ~15 s of host time).
    The default exact mode must equal it bit for bit, with one or two cells
    per lane, in strips and in column groups.  The fast mode (reciprocal multiplications and fused
    multiply-adds) is OUTSIDE the 1e-5 seismogram tolerance at this length
    and is held to 2e-5 only to catch a regression: its float32 rounding
    differences grow chaotically with the number of steps, to 1.1e-5 here,
    the size of the gap between the oracle's own float32 and float64 modes
    (5.3e-6).  Nothing that must meet the tolerance (the bench, the plugin,
    autograd) runs it."""
    from tests.helpers import oracle_records, rel_l2
    prob = configs.psv_config3()
    g = prob.grids[0]
    print("case: full-size shot 0 vs oracle", flush=True)
    want = oracle_records(prob, g, g.shots[0])[0]
    assert np.abs(want).max() > 0
    exact = _records(ctx_factory, _single(prob, g, g.shots[0]))[0]
    np.testing.assert_array_equal(exact, want)
    wide = _records(ctx_factory, _single(prob, g, g.shots[0]), fuse=2, chunk=128)[0]
    np.testing.assert_array_equal(wide, want)
    group = _records(ctx_factory, _single(prob, g, g.shots[0]), fuse=4, chunk=64)[0]
    np.testing.assert_array_equal(group, want)
    fast = _records(ctx_factory, _single(prob, g, g.shots[0]), exact=0)[0]
    err = rel_l2(fast, want)
    print("shot 0, fast mode: rel L2 %.2e" % err)
    assert err <= 2e-5


@pytest.mark.parametrize("name", ["2", "3b", "4", "5"])
def test_one_shot_per_grid_of_each_workload_equals_the_oracle(ctx_factory, name):
    """At the other benchmark workloads' full sizes and lengths (P-SV 400^2
    x 3000 steps; the mixed 300x600 / 500x500 / 800x400 grids x 4000; spin
    600^2 x 3000; the FWI model 500^2 x 3000), the first shot of every grid
    run alone records exactly what the oracle records."""
    from tests.helpers import oracle_records
    prob = {"2": configs.psv_config2, "3b": configs.psv_config3_mixed, "4": configs.spin_config4,
            "5": configs.fwi_config5}[name]()
    for g in prob.grids:
        print("case: workload", name, "grid %dx%d" % (g.nx, g.nz), "nt", prob.nt, flush=True)
        want = oracle_records(prob, g, g.shots[0])[0]
        assert np.abs(want).max() > 0
        got = _records(ctx_factory, _single(prob, g, g.shots[0]))[0]
        np.testing.assert_array_equal(got, want)
