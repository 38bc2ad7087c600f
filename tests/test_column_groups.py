"""Column groups (SP_OPT_FUSE 3 and 4: k_step2g, one or two z cells per
lane) against the strip kernels and the oracle.

A column group is one block whose wavefronts (SP_OPT_GROUP of them) march
the same rows of one shot and cover a span of its z column; z neighbours
across wavefront edges come through LDS, and spans that end inside the
column keep 6 halo cells there.  Each cell's operations are those of the
strip kernels, so records, misfits and kernels must be equal bit for bit,
and records equal the oracle's.  The batches are ragged: with 1 or 2
wavefronts per block the 250-deep grid takes three spans (inner halos at
z = 110..121 and 226..237 with one cell per lane); with 8 the shallower
grids leave wavefronts past the column's end that only take part in the
barriers.  Receivers and sources sit on wavefront edges (z = 62..65,
126..129, 190..193), on span edges and on the grid's corners.
"""
import numpy as np
import pytest

from seispie_amd import _lib, configs
from tests.helpers import oracle_records
from tests.test_gpu_parity import _data_for, _run_forward, _small

pytestmark = pytest.mark.gpu

EDGE_Z = (0, 1, 2, 62, 63, 64, 65, 110, 115, 116, 121, 126, 127, 128, 129, 190, 191, 192, 193, 226, 231, 232, 237)


def _edge_geometry(prob, nt, seed):
    """Receivers on every row (cycling through the wavefront-edge cells) and
    at the column's last cells; sources at corners and on edges."""
    rng = np.random.default_rng(seed)
    for g in prob.grids:
        nz = g.nz
        zs = [z for z in EDGE_Z if z < nz] + [nz - 2, nz - 1]
        rec = [i * nz + zs[i % len(zs)] for i in range(g.nx)]
        rec += [(g.nx // 2) * nz + z for z in zs]
        src = sorted({0, g.nx * nz - 1, 3 * nz + min(63, nz - 1), 7 * nz + min(128, nz - 1),
                      int(rng.integers(4, g.nx - 4)) * nz + int(rng.integers(0, nz))})
        comps = configs.ricker_stf(nt, prob.dt, 30.0, ang=30.0)
        stf = configs.shot_stf(prob.physics, comps)
        stf_many = np.concatenate([stf * (1 + 0.1 * q) for q in range(len(src))], axis=1)
        g.shots = [configs.Shot(np.array(src, np.int32), stf_many * (1 + 0.05 * s), np.array(rec, np.int32))
                   for s in range(len(g.shots))]
    return prob


def _ragged(physics, nt):
    if physics == _lib.SPIN:
        # 150 x 150 (three wavefronts) and 60 x 60 (one) in one batch
        grids = [configs.spin_config4(scale=0.25, nt=nt, nshots=2).grids[0],
                 configs.spin_config4(scale=0.1, nt=nt, nshots=1).grids[0]]
        for g in grids:
            g.sponge = (12, 0.015, 1, 1, 1, 0)
    else:
        # nz 250 (W = 1: four wavefronts, W = 2: two), 60 and 130
        grids = [_small(physics, 67, 250, nt, 2, seed=6).grids[0], _small(physics, 40, 60, nt, 1, seed=7).grids[0],
                 _small(physics, 31, 130, nt, 1, seed=8).grids[0]]
    prob = configs.Problem(physics, nt, 5e-4, grids, sae=10)
    return _edge_geometry(prob, nt, seed=int(physics))


def _run(ctx_factory, prob, data, fuse, chunk, order, sae, group=0):
    ctx = ctx_factory()
    ctx.set_option(_lib.OPT_FUSE, fuse)
    ctx.set_option(_lib.OPT_GROUP, group)
    ctx.set_option(_lib.OPT_CHUNK, chunk)
    ctx.set_option(_lib.OPT_ORDER, order)
    ids = configs.load_problem(ctx, prob)
    ctx.forward(0)
    recs = [ctx.records(i) for i in ids]
    for i, d in zip(ids, data):
        ctx.set_data(i, d)
    norms = ctx.gradient(sae)
    kern = [[ctx.kernel(g, w) for w in (_lib.K_LAM, _lib.K_MU, _lib.K_RHO, _lib.K_HESS)]
            for g in range(len(prob.grids))]
    launches = ctx.timing()["step_launches"]
    assert ctx.get_option(_lib.OPT_FUSE) == fuse
    ctx.close()
    return recs, norms, kern, launches


@pytest.mark.parametrize("physics,fuse,group", [(_lib.SH, 3, 4), (_lib.SH, 4, 2), (_lib.SH, 3, 1), (_lib.PSV, 3, 2),
                                                (_lib.PSV, 3, 8), (_lib.PSV, 4, 4), (_lib.PSV, 4, 1),
                                                (_lib.SPIN, 3, 4), (_lib.SPIN, 3, 2)])
@pytest.mark.parametrize("chunk,order,sae", [(16, 0, 10), (4, 1, 5), (36, 1, 1)])
def test_column_groups_equal_strips(ctx_factory, physics, fuse, group, chunk, order, sae):
    nt = 120
    prob = _ragged(physics, nt)
    prob.sae = sae
    data = _data_for(prob)
    print("case: column groups", physics, "fuse", fuse, "group", group, chunk, order, "sae", sae, flush=True)
    r1, n1, k1, l1 = _run(ctx_factory, prob, data, 1, chunk, order, sae)
    rg, ng, kg, lg = _run(ctx_factory, prob, data, fuse, chunk, order, sae, group)
    assert lg == l1  # two steps per launch in both
    for a, b in zip(r1, rg):
        assert np.abs(a).max() > 0
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(n1, ng)
    for ga, gb in zip(k1, kg):
        for a, b in zip(ga, gb):
            np.testing.assert_array_equal(a, b)
    s = 0
    for g in prob.grids:
        for shot in g.shots:
            ref = oracle_records(prob, g, shot)[0]
            np.testing.assert_array_equal(rg[s], ref, err_msg="shot %d" % s)
            s += 1


@pytest.mark.parametrize("physics,fuse,nz", [(_lib.SH, 4, 61), (_lib.PSV, 4, 1031), (_lib.PSV, 3, 1040),
                                             (_lib.PSV, 4, 1040)])
def test_deep_columns_and_odd_nz(ctx_factory, physics, fuse, nz):
    """Columns of several spans (1040 cells: five spans of 256 one-cell or
    two spans of 512 two-cell cells) equal the strips; a batch that cannot
    hold two cells per lane (odd nz) runs the one-cell strips.  The records
    equal those of SP_OPT_FUSE 1, the option unchanged."""
    prob = _small(physics, 20, nz, 80, 1)
    g = prob.grids[0]
    # sources and receivers around the span boundaries (one cell per lane:
    # spans start at 244, 488, 732, 976; two cells: at 500)
    zs = [z for z0 in (244, 250, 488, 500, 732, 976) for z in range(z0 - 8, z0 + 9, 2) if z < nz]
    zs = zs or list(range(1, nz - 1, 3))
    rec = [i * nz + zs[(5 * i + k) % len(zs)] for i in range(2, g.nx - 2) for k in range(4)]
    src = sorted({8 * nz + min(250, nz - 3), 10 * nz + min(500, nz - 5), 12 * nz + min(735, nz - 7),
                  9 * nz + min(980, nz - 9)})
    comps = configs.ricker_stf(80, prob.dt, 30.0, ang=30.0)
    stf = configs.shot_stf(physics, comps)
    g.shots = [configs.Shot(np.array(src, np.int32), np.concatenate([stf * (1 + 0.1 * q) for q in range(len(src))],
                                                                    axis=1), np.array(rec, np.int32))]
    print("case: deep column / odd nz", physics, fuse, nz, flush=True)
    _, one = _run_forward(ctx_factory, prob, fuse=1)
    ctx, grp = _run_forward(ctx_factory, prob, fuse=fuse)
    assert ctx.get_option(_lib.OPT_FUSE) == fuse
    for a, b in zip(one, grp):
        assert np.abs(a).max() > 0
        np.testing.assert_array_equal(a, b)


def test_spin_two_cells_falls_back(ctx_factory):
    prob = _small(_lib.SPIN, nt=80, nshots=1)
    print("case: spin fuse 4 fallback", flush=True)
    _, one = _run_forward(ctx_factory, prob, fuse=1)
    _, grp = _run_forward(ctx_factory, prob, fuse=4)
    np.testing.assert_array_equal(one[0], grp[0])


@pytest.mark.parametrize("fuse", [3, 4])
def test_column_groups_on_tiny_grids(ctx_factory, fuse):
    """Grids smaller than the stencil bands and than one wavefront."""
    from tests.test_gpu_parity import _check_forward
    for nx, nz in ((5, 4), (3, 70), (9, 10), (12, 130)):
        prob = _small(_lib.SH, nx, nz, 50, 1, nrec=2, sponge=(3, 0.1, 1, 1, 1, 1))
        print("case: tiny", nx, nz, "fuse", fuse, flush=True)
        _, recs = _run_forward(ctx_factory, prob, fuse=fuse)
        _check_forward(prob, recs, nonzero=False)


@pytest.mark.parametrize("physics,fuse", [(_lib.PSV, 3), (_lib.SH, 4)])
def test_column_groups_in_replayed_graphs(ctx_factory, physics, fuse):
    """Forward runs and a gradient replayed from captured graphs with column
    groups equal the eager strip runs bit for bit, on repeated runs."""
    prob = _small(physics, 70, 130, 150, 2)
    data = _data_for(prob)
    print("case: graph", physics, "fuse", fuse, flush=True)
    ctx = ctx_factory()
    ctx.set_option(_lib.OPT_FUSE, 1)
    ids = configs.load_problem(ctx, prob)
    ctx.forward(0)
    eager = [ctx.records(i) for i in ids]
    for i, d in zip(ids, data):
        ctx.set_data(i, d)
    n_e = ctx.gradient(prob.sae)
    k_e = ctx.kernel(0, _lib.K_MU)
    ctx.close()
    ctx = ctx_factory()
    ctx.set_option(_lib.OPT_GRAPH, 1)
    ctx.set_option(_lib.OPT_FUSE, fuse)
    ids = configs.load_problem(ctx, prob)
    for rep in range(3):
        ctx.forward(0)
        for i in ids:
            np.testing.assert_array_equal(ctx.records(i), eager[i], err_msg="replay run %d" % rep)
        for i, d in zip(ids, data):
            ctx.set_data(i, d)
        np.testing.assert_array_equal(ctx.gradient(prob.sae), n_e)
        np.testing.assert_array_equal(ctx.kernel(0, _lib.K_MU), k_e)


@pytest.mark.parametrize("physics,fuse", [(_lib.SH, 4), (_lib.PSV, 3), (_lib.PSV, 4), (_lib.SPIN, 3)])
@pytest.mark.parametrize("mode,sae", [(0, 5), (1, 1)])
def test_checkpointed_adjoints_with_column_groups(ctx_factory, physics, fuse, mode, sae):
    """sp_forward_ckpt / sp_adjoint_ckpt, reference-style (mode 0) and
    discrete (mode 1): segments recomputed from checkpoints by the two-step
    kernel, which writes the snapshots of both of its steps, in column
    groups of 2 wavefronts (three spans of a 250-cell column) against the
    strips, beside a free surface; kernels and records bit for bit."""
    from tests.test_discrete_adjoint import _spin_surface_problem, _surface_problem
    nt = 200
    prob = (_spin_surface_problem(70, 250, nt) if physics == _lib.SPIN else _surface_problem(physics, 70, 250, nt))
    g = prob.grids[0]
    ncomp = 1 if physics == _lib.SH else 2
    rng = np.random.default_rng(3)
    W = [rng.standard_normal((ncomp, s.rec_k.size, nt)).astype(np.float32) for s in g.shots]
    kids = [_lib.K_LAM, _lib.K_MU, _lib.K_RHO, _lib.K_HESS]
    if physics == _lib.SPIN:
        kids += [_lib.K_NU, _lib.K_J, _lib.K_MUC]
    print("case: checkpointed adjoint", physics, "fuse", fuse, "mode", mode, "sae", sae, flush=True)
    out = []
    for f in (1, fuse):
        ctx = ctx_factory()
        ctx.set_option(_lib.OPT_FUSE, f)
        ctx.set_option(_lib.OPT_GROUP, 2)
        ctx.set_option(_lib.OPT_CHUNK, 24)
        ctx.set_option(_lib.OPT_ADJOINT, mode)
        ids = configs.load_problem(ctx, prob)
        ctx.forward_ckpt(20)
        recs = [ctx.records(i) for i in ids]
        for i, w in zip(ids, W):
            ctx.set_adjoint_source(i, w)
        ctx.adjoint_ckpt(sae)
        out.append((recs, [ctx.kernel(0, k) for k in kids]))
        ctx.close()
    (r1, k1), (rg, kg) = out
    for a, b in zip(r1, rg):
        assert np.abs(a).max() > 0
        np.testing.assert_array_equal(a, b)
    for a, b in zip(k1, kg):
        np.testing.assert_array_equal(a, b)
    assert np.abs(k1[1]).max() > 0
