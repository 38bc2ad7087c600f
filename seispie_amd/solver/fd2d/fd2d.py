"""fd2d solver plugin on the MI355X library.

Drop-in for seispie/solver/fd2d/fd2d.py (class fd2d, :110-443): the same
methods, attributes, config keys, file formats and return values, with the
time loops of seispie/solver/fd2d/solver.py replaced by the batched HIP
kernels of libseispie_hip (one launch per time step for all shots).

Differences a caller can see:
  * run_kernel runs every shot of the loop in one batched pass instead of
    one shot after another; misfit and kernels are the same sums.
  * run_kernel ignores self.mpi (every shot runs in this process), so the
    reference's shared-file gradient sum (solver.py:375-393), and the
    double count of a rank's own shot in it, are not reproduced.
  * the P-SV / spin adjoint works (the reference's raises NameError at
    solver.py:326-327); compute_gradient returns their mu kernel, and
    compute_kernels returns every kernel.
  * a missing `spin` key reads as 'no' (the reference raises KeyError,
    e.g. for examples/forward/config.ini).
"""
from os import path, makedirs
from time import time

import numpy as np

from seispie_amd import _lib
from seispie_amd.solver.base import base
from seispie_amd.solver.source.ricker import ricker


class _Stream:
    """Stands in for the numba stream the callers synchronise
    (optimize/base.py:60-61)."""

    def __init__(self, solver):
        self._solver = solver

    def synchronize(self):
        ctx = self._solver._ctx
        if ctx is not None:
            ctx.sync()


class _DeviceField:
    """A model or kernel array on the device with numba's copy_to_host
    (optimize/base.py:60)."""

    def __init__(self, solver, kind, which):
        self._solver, self._kind, self._which = solver, kind, which

    def copy_to_host(self, ary=None, stream=None):
        s = self._solver
        if self._kind == 'field':
            out = s._ctx.get_field(s._grid, self._which)
        else:
            out = s._ctx.kernel(s._grid, self._which)
        if ary is None:
            return out
        ary[...] = out.reshape(ary.shape)
        return ary

    @property
    def size(self):
        return self._solver.nx * self._solver.nz


class fd2d(base):
    # ------------------------------------------------------------ setup
    def setup(self, workflow):
        """fd2d.py:111-120"""
        self.stream = _Stream(self)
        self.nt = int(self.config['nt'])
        self.dt = float(self.config['dt'])
        self.sh = 1 if self.config['sh'] == 'yes' else 0
        self.psv = 1 if self.config['psv'] == 'yes' else 0
        self.spin = 1 if self.config.get('spin', 'no') == 'yes' else 0
        self.sae = 0
        self.nsa = 0
        self.device = int(self.config.get('device', 0))
        self._ctx = None
        self._ctx_psv = None
        self._grid = 0
        self.taskid = 0
        self._obs = {}

    @property
    def _phys(self):
        return _lib.SPIN if self.spin else _lib.PSV

    def import_sources(self):
        """fd2d.py:122-146 (Ricker time functions on the host, float64 -> float32)."""
        src = np.loadtxt(self.path['sources'], ndmin=2)
        nsrc = self.nsrc = src.shape[0]
        src_id = np.zeros(nsrc, dtype='int32')
        stf_x = np.zeros(nsrc * self.nt, dtype='float32')
        stf_y = np.zeros(nsrc * self.nt, dtype='float32')
        stf_z = np.zeros(nsrc * self.nt, dtype='float32')
        t = np.arange(self.nt) * self.dt
        for isrc in range(nsrc):
            src_x = int(np.round(src[isrc][0] / self.dx))
            src_z = int(np.round(src[isrc][1] / self.dz))
            src_id[isrc] = src_x * self.nz + src_z
            sx, sy, sz = ricker(t, *src[isrc][3:])
            stf_x[isrc * self.nt:(isrc + 1) * self.nt] = sx
            stf_y[isrc * self.nt:(isrc + 1) * self.nt] = sy
            stf_z[isrc * self.nt:(isrc + 1) * self.nt] = sz
        self.src_id = src_id
        self.stf_x, self.stf_y, self.stf_z = stf_x, stf_y, stf_z

    def import_stations(self):
        """fd2d.py:148-165"""
        rec = np.loadtxt(self.path['stations'], ndmin=2)
        nrec = self.nrec = rec.shape[0]
        rec_id = np.zeros(nrec, dtype='int32')
        for irec in range(nrec):
            rec_x = int(np.round(rec[irec][0] / self.dx))
            rec_z = int(np.round(rec[irec][1] / self.dz))
            rec_id[irec] = rec_x * self.nz + rec_z
        self.rec_id = rec_id

    def import_model(self, model_true):
        """fd2d.py:167-252: read proc000000_<name>.bin (int32 npt + float32),
        derive the grid, convert vp/vs to lambda/mu on the device and set the
        sponge."""
        model = dict()
        model_dir = self.path['model_true'] if model_true else self.path['model_init']
        model_params = ['x', 'z', 'lambda', 'mu', 'nu', 'j', 'lambda_c', 'mu_c', 'nu_c', 'rho'] \
            if self.spin else ['x', 'z', 'vp', 'vs', 'rho']
        for name in model_params:
            filename = path.join(model_dir, 'proc000000_' + name + '.bin')
            with open(filename, 'rb') as f:
                if not hasattr(self, 'npt'):
                    f.seek(0)
                    self.npt = np.fromfile(f, dtype='int32', count=1)
                f.seek(4)
                model[name] = np.fromfile(f, dtype='float32')

        npt = self.npt[0]
        self.threads_per_block = int(self.config['threads_per_block'])

        x = model['x']
        z = model['z']
        lx = x.max() - x.min()
        lz = z.max() - z.min()
        nx = self.nx = int(np.rint(np.sqrt(npt * lx / lz)))
        nz = self.nz = int(np.rint(np.sqrt(npt * lz / lx)))
        self.dx = lx / (nx - 1)
        self.dz = lz / (nz - 1)

        if self.spin:
            fields = [model[k] for k in ['lambda', 'mu', 'nu', 'j', 'lambda_c', 'mu_c', 'nu_c', 'rho']]
            self._build(fields, 1)
        else:
            self._build([model['vp'], model['vs'], model['rho']], 0)

        if self.config['save_coordinates']:
            self.export_field(x, 'x')
            self.export_field(z, 'z')

    def _sponge_args(self):
        return (int(self.config['abs_width']), float(self.config['abs_alpha']),
                1 if self.config['abs_left'] == 'yes' else 0,
                1 if self.config['abs_right'] == 'yes' else 0,
                1 if self.config['abs_bottom'] == 'yes' else 0,
                1 if self.config['abs_top'] == 'yes' else 0)

    def _new_ctx(self, physics, fields, param):
        ctx = _lib.Context(self.device)
        ctx.set_physics(physics, self.nt, self.dt)
        g = ctx.add_grid(self.nx, self.nz, self.dx, self.dz, fields, param)
        ctx.set_sponge(g, *self._sponge_args())
        return ctx

    def _build(self, fields, param):
        for c in (self._ctx, self._ctx_psv):
            if c is not None:
                c.close()
        self._ctx = self._ctx_psv = None
        if self.sh:
            self._ctx = self._new_ctx(_lib.SH, fields[:3] if param == 0 else fields, param)
        if self.psv:
            c = self._new_ctx(self._phys, fields, param)
            if self._ctx is None:
                self._ctx = c
            else:
                self._ctx_psv = c
        if self._ctx is None:
            raise ValueError('solver: one of sh or psv must be yes')
        self._grid = 0
        self.rho = _DeviceField(self, 'field', _lib.F_RHO)
        self.lam = _DeviceField(self, 'field', _lib.F_LAM)
        self.mu = _DeviceField(self, 'field', _lib.F_MU)
        self.k_lam = _DeviceField(self, 'kernel', _lib.K_LAM)
        self.k_mu = _DeviceField(self, 'kernel', _lib.K_MU)
        self.k_rho = _DeviceField(self, 'kernel', _lib.K_RHO)

    def _contexts(self):
        """(physics, context) pairs: SH first, then P-SV / spin."""
        out = []
        if self.sh:
            out.append((_lib.SH, self._ctx))
        if self.psv:
            out.append((self._phys, self._ctx if not self.sh else self._ctx_psv))
        return out

    def export_field(self, field, name, it=0):
        """fd2d.py:254-261"""
        name = 'proc%06d_%s' % (it, name)
        with open(self.path['output'] + '/' + name + '.bin', 'wb') as f:
            self.npt.tofile(f)
            np.ascontiguousarray(field, dtype='float32').tofile(f)

    def setup_adjoint(self):
        """fd2d.py:263-301 (buffers live in the library)."""
        self.sae = int(self.config['adjoint_interval'])
        self.nsa = int(self.nt / self.sae)
        self.sigma = int(self.config['smooth'])

    # ---------------------------------------------------------- shots
    def _stf(self, physics, ids):
        nt = self.nt
        pick = lambda a: np.stack([a[i * nt:(i + 1) * nt] for i in ids]) if len(ids) else np.zeros((0, nt), 'float32')
        if physics == _lib.SH:
            return pick(self.stf_y)[None]
        return np.stack([pick(self.stf_x), pick(self.stf_z)])

    def _load_shots(self, physics, ctx, shots):
        """shots: list of source-index lists, one simulation each.  The work
        decomposition is autotuned once per batch shape (the library caches
        the choice per process; results do not depend on it)."""
        ctx.clear_shots()
        for ids in shots:
            ctx.add_shot(self._grid, self.src_id[list(ids)], self._stf(physics, ids), self.rec_id)
        ctx.autotune(20)

    def _isrc(self):
        return -1 if self.config['combine_sources'] == 'yes' else self.taskid

    def _sources_of(self, i):
        """Sources that fire in task i (solver.py:180-183, stf_dsy :78-84)."""
        return list(range(self.nsrc)) if self.config['combine_sources'] == 'yes' else [i]

    # ------------------------------------------------------------- runs
    def run_forward(self):
        """solver.py:166-287 for the current taskid (or all sources with
        combine_sources = yes); records land in self._obs."""
        isrc = self._isrc()
        ids = list(range(self.nsrc)) if isrc < 0 else [isrc]
        self._obs = {}
        sfe = int(self.config.get('save_snapshot', 0) or 0)
        for physics, ctx in self._contexts():
            self._load_shots(physics, ctx, [ids])
            if sfe > 0:
                self._forward_exporting(physics, ctx, sfe)
            else:
                ctx.forward(0)
            rec = ctx.records(0)
            names = ['y'] if physics == _lib.SH else (['x', 'z', 'y'] if physics == _lib.SPIN else ['x', 'z'])
            for c, nm in enumerate(names):
                self._obs[nm] = rec[c].reshape(-1)

        if 'output_traces' in self.path:
            tracedir = self.path['output_traces']
            nrec, nt = self.nrec, self.nt
            i = isrc if isrc >= 0 else 0
            if self.sh:
                np.save('%s/uy_%06d.npy' % (tracedir, i), self._obs['y'].reshape([nrec, nt]))
            if self.psv:
                np.save('%s/ux_%06d.npy' % (tracedir, i), self._obs['x'].reshape([nrec, nt]))
                np.save('%s/uz_%06d.npy' % (tracedir, i), self._obs['z'].reshape([nrec, nt]))
                if self.spin:
                    np.save('%s/ry_%06d.npy' % (tracedir, i), self._obs['y'].reshape([nrec, nt]))

    def _forward_exporting(self, physics, ctx, sfe):
        """solver.py:244-261: after every step it > 0 with it % sfe == 0 the
        velocities are written as proc%06d_<name>.bin (it in the name):
        vy (SH); vx, vz (P-SV); and ry, the spin velocity, with spin."""
        names = ['vy'] if physics == _lib.SH else (['vx', 'vz', 'ry'] if physics == _lib.SPIN else ['vx', 'vz'])
        start = 0
        for it in range(sfe, self.nt, sfe):
            ctx.forward_steps(start, it + 1)
            start = it + 1
            v = ctx.wavefield(0, _lib.WF_V)
            for c, nm in enumerate(names):
                self.export_field(v[c], nm, it)
        if start < self.nt:
            ctx.forward_steps(start, self.nt)

    @property
    def obs_x(self):
        return self._obs.get('x')

    @property
    def obs_y(self):
        return self._obs.get('y')

    @property
    def obs_z(self):
        return self._obs.get('z')

    def run_adjoint(self):
        """solver.py:290-329 for the current task: the forward pass that
        produces its snapshots is rerun first, so this is self-contained."""
        self.setup_adjoint()
        ids = self._sources_of(self.taskid)
        for physics, ctx in self._contexts():
            self._load_shots(physics, ctx, [ids])
            self._set_shot_data(physics, ctx, 0, self.taskid)
            ctx.gradient(self.sae)

    def import_traces(self):
        """fd2d.py:354-423: raw float32 v?_%06d.npy files, or a forward run
        of every shot with the current (true) model, written to
        output/traces."""
        nsrc = self.nsrc
        self.syn_x, self.syn_y, self.syn_z = [], [], []
        if 'traces' in self.path:
            tracedir = self.path['traces']
            for i in range(nsrc):
                if self.sh:
                    self.syn_y.append(np.fromfile('%s/vy_%06d.npy' % (tracedir, i), dtype='float32'))
                if self.psv:
                    self.syn_x.append(np.fromfile('%s/vx_%06d.npy' % (tracedir, i), dtype='float32'))
                    self.syn_z.append(np.fromfile('%s/vz_%06d.npy' % (tracedir, i), dtype='float32'))
            return

        tracedir = self.path['output'] + '/traces'
        print('Generating traces')
        if not path.exists(tracedir):
            makedirs(tracedir)
        start = time()
        for physics, ctx in self._contexts():
            self._load_shots(physics, ctx, [self._sources_of(i) for i in range(nsrc)])
            ctx.forward(0)
            for i in range(nsrc):
                rec = ctx.records(i)
                if physics == _lib.SH:
                    out = rec[0].reshape(-1).copy()
                    self.syn_y.append(out)
                    out.tofile('%s/vy_%06d.npy' % (tracedir, i))
                else:
                    ox, oz = rec[0].reshape(-1).copy(), rec[1].reshape(-1).copy()
                    self.syn_x.append(ox)
                    self.syn_z.append(oz)
                    ox.tofile('%s/vx_%06d.npy' % (tracedir, i))
                    oz.tofile('%s/vz_%06d.npy' % (tracedir, i))
        print('Elapsed time: %.2fs' % (time() - start))
        print('')

    def _set_shot_data(self, physics, ctx, shot, i):
        nt, nrec = self.nt, self.nrec
        if physics == _lib.SH:
            d = np.asarray(self.syn_y[i], dtype='float32').reshape(1, nrec, nt)
        else:
            d = np.stack([np.asarray(self.syn_x[i], 'float32').reshape(nrec, nt),
                          np.asarray(self.syn_z[i], 'float32').reshape(nrec, nt)])
            if physics == _lib.SPIN:
                d = np.concatenate([d, np.zeros((1, nrec, nt), 'float32')])
        ctx.set_data(shot, d)

    def run_kernel(self, adj):
        """solver.py:332-404: misfit (and kernels) summed over every shot,
        all shots in one batched pass; k_mu smoothed when adj."""
        print('Computing kernels' if adj else 'Computing misfit')
        self.setup_adjoint()
        misfit = 0.0
        for physics, ctx in self._contexts():
            self._load_shots(physics, ctx, [self._sources_of(i) for i in range(self.nsrc)])
            for i in range(self.nsrc):
                self._set_shot_data(physics, ctx, i, i)
            if adj:
                norms = ctx.gradient(self.sae)
            else:
                ctx.forward(0)
                norms = ctx.misfit()
            misfit += float(norms.sum())
            if adj:
                if physics != _lib.SH:
                    # vp / vs / rho' kernels from the raw lam / mu / rho
                    # kernels and the model of this gradient, before k_mu is
                    # smoothed in place as solver.py:395-396 does
                    ctx.kernels_vps(self._grid)
                if self.sigma > 0:
                    ctx.smooth_kernel(self._grid, _lib.K_MU, self.sigma)
        misfit = np.float32(misfit)
        print('  misfit = %.2f' % misfit)
        if adj:
            return misfit, self.k_mu, self.mu
        return misfit

    def compute_misfit(self):
        """fd2d.py:425-426"""
        return self.run_kernel(0)

    def compute_gradient(self):
        """fd2d.py:428-437: (smoothed k_mu, misfit, mu) as host float32."""
        misfit, kernel, model = self.run_kernel(1)
        return kernel.copy_to_host(), misfit, model.copy_to_host()

    def compute_kernels(self, smooth=True):
        """Every kernel of the last gradient as host float32 arrays (P-SV /
        spin: lam, mu, rho, vp, vs, rho at fixed vp/vs; spin adds nu, j,
        muc).  'mu' is k_mu as run_kernel left it (smoothed once); the
        others are derived from the unsmoothed kernels and each smoothed
        once into a new array, so repeated calls return the same values."""
        out = {}
        for physics, ctx in self._contexts():
            if physics == _lib.SH:
                out['mu'] = ctx.kernel(self._grid, _lib.K_MU)
                continue
            names = {'lam': _lib.K_LAM, 'mu': _lib.K_MU, 'rho': _lib.K_RHO, 'vp': _lib.K_VP,
                     'vs': _lib.K_VS, 'rhop': _lib.K_RHOP}
            if physics == _lib.SPIN:
                names.update(nu=_lib.K_NU, j=_lib.K_J, muc=_lib.K_MUC)
            for nm, w in names.items():
                k = ctx.kernel(self._grid, w)
                if smooth and nm != 'mu' and self.sigma > 0:
                    k = ctx.smooth(k, self.nx, self.nz, self.sigma)
                out[nm] = k
        return out

    def update_model(self, mu):
        """fd2d.py:442-443: replaces mu only."""
        for _, ctx in self._contexts():
            ctx.set_field(self._grid, _lib.F_MU, np.ascontiguousarray(mu, dtype='float32'))

    def smooth(self, data):
        """fd2d.py:349-352, in place on a host float32 array of nx*nz."""
        out = self._ctx.smooth(data, self.nx, self.nz, self.sigma)
        data[...] = out.reshape(np.shape(data))

    def clear_kernels(self):
        """fd2d.py:311-316 (the library clears kernels at each gradient)."""

    def clear_wavefields(self):
        """fd2d.py:318-341 (the library clears wavefields at each run)."""
