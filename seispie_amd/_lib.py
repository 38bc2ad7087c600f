"""ctypes binding of libseispie_hip.so (include/seispie_hip.h).

The library is the only compute path: if it is missing or no HIP device is
visible, these functions raise instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

SH, PSV, SPIN = 1, 2, 3
F_RHO, F_LAM, F_MU, F_NU, F_J, F_LAMC, F_MUC, F_NUC = range(8)
K_LAM, K_MU, K_RHO, K_NU, K_J, K_MUC, K_VP, K_VS, K_RHOP, K_HESS = range(10)
OPT_CHUNK, OPT_ORDER, OPT_GRAPH, OPT_SNAP_MB, OPT_KERNEL, OPT_STRIP, OPT_FUSE, OPT_EXACT, OPT_ADJOINT, OPT_GROUP = \
    range(1, 11)
WF_V, WF_U, WF_S = 0, 1, 2

N_REC = {SH: 1, PSV: 2, SPIN: 3}   # record components: SH y; P-SV x, z; spin x, z, y
N_INJ = {SH: 1, PSV: 2, SPIN: 2}   # source components: SH y; P-SV/spin x, z
N_MIS = {SH: 1, PSV: 2, SPIN: 2}   # misfit components: SH y; P-SV/spin x, z

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SEISPIE_HIP_LIB") or os.path.join(_HERE, "lib", "libseispie_hip.so")
_lib = None


class SeispieError(RuntimeError):
    pass


def _share_torch_runtime():
    """One HIP runtime per process.  PyTorch-ROCm ships its own
    libamdhip64 (same soname, libamdhip64.so.7, as /opt/rocm's), which its
    libraries find by RUNPATH.  If this library were loaded first, its
    dependency would resolve to /opt/rocm's copy and a later `import torch`
    would map torch's copy as a second runtime, which then finds no GPU, and
    device pointers and streams could not pass between the two.  Mapping
    torch's copy first (without importing torch) makes this library's
    dependency resolve to it by soname and torch's to it by file identity.
    SEISPIE_HIP_SYSTEM_RUNTIME=1 keeps /opt/rocm's runtime (for processes
    that never import torch)."""
    import importlib.util
    if os.environ.get("SEISPIE_HIP_SYSTEM_RUNTIME") == "1":
        return
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.submodule_search_locations:
        return
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def load():
    """Load the HIP library (raises ImportError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libseispie_hip.so is not built (%s); run __graft_entry__.build() or "
                          "python -m seispie_amd.build" % LIB_PATH)
    _share_torch_runtime()
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
    sig = {
        "sp_create": ([i, ctypes.POINTER(vp)], i),
        "sp_destroy": ([vp], i),
        "sp_last_error": ([vp], ctypes.c_char_p),
        "sp_set_stream": ([vp, vp], i),
        "sp_sync": ([vp], i),
        "sp_set_option": ([vp, i, ll], i),
        "sp_set_physics": ([vp, i, i, d], i),
        "sp_add_grid": ([vp, i, i, ctypes.c_float, ctypes.c_float, i, vp, ctypes.POINTER(i)], i),
        "sp_set_sponge": ([vp, i, i, d, i, i, i, i], i),
        "sp_set_field": ([vp, i, i, vp, i], i),
        "sp_get_field": ([vp, i, i, vp, i], i),
        "sp_add_shot": ([vp, i, i, vp, vp, i, vp, ctypes.POINTER(i)], i),
        "sp_clear_shots": ([vp], i),
        "sp_num_components": ([vp], i),
        "sp_forward": ([vp, i], i),
        "sp_get_records": ([vp, i, vp, i], i),
        "sp_set_data": ([vp, i, vp, i], i),
        "sp_misfit": ([vp, vp], i),
        "sp_adjoint": ([vp, i], i),
        "sp_gradient": ([vp, i, vp], i),
        "sp_get_kernel": ([vp, i, i, vp, i], i),
        "sp_kernels_vps": ([vp, i], i),
        "sp_smooth_kernel": ([vp, i, i, i], i),
        "sp_smooth": ([vp, vp, i, i, i, i], i),
        "sp_smooth_noalloc": ([vp, vp, i, i, i], i),
        "sp_last_timing": ([vp, ctypes.POINTER(ll), ctypes.POINTER(d), ctypes.POINTER(d), ctypes.POINTER(d)], i),
        "sp_autotune": ([vp, i, ctypes.POINTER(i), ctypes.POINTER(i)], i),
        "sp_forward_ckpt": ([vp, i], i),
        "sp_adjoint_ckpt": ([vp, i], i),
        "sp_set_adjoint_source": ([vp, i, vp, i], i),
        "sp_set_model_vps": ([vp, i, vp, vp, vp, i], i),
        "sp_forward_steps": ([vp, i, i], i),
        "sp_get_option": ([vp, i, ctypes.POINTER(ll)], i),
        "sp_get_wavefield": ([vp, i, i, vp, i], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def exported_symbols():
    """Names of the C entry points the binding expects (all in the header)."""
    return ["sp_create", "sp_destroy", "sp_last_error", "sp_set_stream", "sp_sync", "sp_set_option",
            "sp_set_physics", "sp_add_grid", "sp_set_sponge", "sp_set_field", "sp_get_field",
            "sp_add_shot", "sp_clear_shots", "sp_num_components", "sp_forward", "sp_get_records",
            "sp_set_data", "sp_misfit", "sp_adjoint", "sp_gradient", "sp_get_kernel", "sp_kernels_vps",
            "sp_smooth_kernel", "sp_smooth", "sp_last_timing", "sp_autotune", "sp_forward_ckpt",
            "sp_adjoint_ckpt", "sp_set_adjoint_source", "sp_set_model_vps", "sp_forward_steps", "sp_get_wavefield", "sp_get_option", "sp_smooth_noalloc"]


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None and a.size != n:
        raise ValueError("expected %d values, got %d" % (n, a.size))
    return a


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _devptr(t, n, device):
    """Device pointer of a torch tensor that must be float32, contiguous,
    hold exactly n values and live on GPU `device`.  The C entry points take
    no length, so this check is what keeps a copy inside the allocation."""
    if not t.is_cuda or not t.is_contiguous() or str(t.dtype) != "torch.float32":
        raise ValueError("expected a contiguous float32 tensor on the GPU")
    if t.device.index != device:
        raise ValueError("tensor is on %s, the context on cuda:%d" % (t.device, device))
    if t.numel() != n:
        raise ValueError("expected a tensor of %d values, got %d" % (n, t.numel()))
    return ctypes.c_void_p(t.data_ptr())


def _is_tensor(a):
    return type(a).__module__.startswith("torch")


class Context:
    """One device context: grids, shots and the batched time loop."""

    def __init__(self, device=0):
        L = load()
        p = ctypes.c_void_p()
        rc = L.sp_create(int(device), ctypes.byref(p))
        if rc != 0:
            raise SeispieError("sp_create(device=%d) failed with code %d (no HIP device visible?)" % (device, rc))
        self._p = p
        self.L = L
        self.device = int(device)
        self.physics = None
        self.nt = None
        self.grids = []   # (nx, nz)
        self.shots = []   # (grid, nsrc, nrec)

    def close(self):
        if getattr(self, "_p", None):
            self.L.sp_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            msg = self.L.sp_last_error(self._p)
            raise SeispieError((msg or b"error %d" % rc).decode())

    # ------------------------------------------------------------ setup
    def set_stream(self, stream_handle):
        """Order the context's work on a hipStream_t (an int handle, e.g.
        torch.cuda.current_stream().cuda_stream; 0 is the null stream)."""
        self._chk(self.L.sp_set_stream(self._p, ctypes.c_void_p(stream_handle)))

    def sync(self):
        self._chk(self.L.sp_sync(self._p))

    def set_option(self, key, value):
        self._chk(self.L.sp_set_option(self._p, int(key), int(value)))

    def get_option(self, key):
        v = ctypes.c_longlong()
        self._chk(self.L.sp_get_option(self._p, int(key), ctypes.byref(v)))
        return v.value

    def set_physics(self, physics, nt, dt):
        self._chk(self.L.sp_set_physics(self._p, int(physics), int(nt), float(dt)))
        self.physics, self.nt, self.dt = int(physics), int(nt), float(dt)

    def add_grid(self, nx, nz, dx, dz, fields, param):
        """fields: (vp, vs, rho) for param 0, the 8 spin fields (lam, mu, nu,
        j, lam_c, mu_c, nu_c, rho) for param 1, (lam, mu, rho) for param 2."""
        need = {0: 3, 1: 8, 2: 3}.get(int(param))
        if need is not None and len(fields) != need:
            raise ValueError("param %d takes %d fields, got %d" % (param, need, len(fields)))
        n = nx * nz
        arrs = [_f32(f, n) for f in fields]
        ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        g = ctypes.c_int()
        self._chk(self.L.sp_add_grid(self._p, int(nx), int(nz), float(dx), float(dz), int(param), ptrs, ctypes.byref(g)))
        self.grids.append((nx, nz))
        return g.value

    def set_sponge(self, grid, width, alpha, left, right, bottom, top):
        self._chk(self.L.sp_set_sponge(self._p, grid, int(width), float(alpha), int(left), int(right),
                                       int(bottom), int(top)))

    def set_field(self, grid, field, values):
        nx, nz = self.grids[grid]
        if _is_tensor(values):
            self._chk(self.L.sp_set_field(self._p, grid, field, _devptr(values, nx * nz, self.device), 1))
        else:
            a = _f32(values, nx * nz)
            self._chk(self.L.sp_set_field(self._p, grid, field, _ptr(a), 0))

    def get_field(self, grid, field, out=None):
        nx, nz = self.grids[grid]
        if out is not None and _is_tensor(out):
            self._chk(self.L.sp_get_field(self._p, grid, field, _devptr(out, nx * nz, self.device), 1))
            return out
        a = np.empty(nx * nz, dtype=np.float32)
        self._chk(self.L.sp_get_field(self._p, grid, field, _ptr(a), 0))
        return a

    def add_shot(self, grid, src_k, stf, rec_k):
        """src_k [nsrc] grid indices; stf [n_inj, nsrc, nt]; rec_k [nrec]."""
        src_k = np.ascontiguousarray(src_k, dtype=np.int32).reshape(-1)
        rec_k = np.ascontiguousarray(rec_k, dtype=np.int32).reshape(-1)
        stf = _f32(stf, N_INJ[self.physics] * src_k.size * self.nt)
        s = ctypes.c_int()
        self._chk(self.L.sp_add_shot(self._p, grid, src_k.size, _ptr(src_k), _ptr(stf), rec_k.size,
                                     _ptr(rec_k), ctypes.byref(s)))
        self.shots.append((grid, src_k.size, rec_k.size))
        return s.value

    def clear_shots(self):
        self._chk(self.L.sp_clear_shots(self._p))
        self.shots = []

    # ------------------------------------------------------------- runs
    def forward(self, sae=0):
        self._chk(self.L.sp_forward(self._p, int(sae)))

    def forward_steps(self, it0, it1):
        """Steps [it0, it1) of the forward run of every shot (it0 = 0 starts one)."""
        self._chk(self.L.sp_forward_steps(self._p, int(it0), int(it1)))

    def wavefield(self, shot, kind=WF_V, out=None):
        """Current velocities (WF_V), displacements (WF_U) or stresses (WF_S)
        of a shot, [ncomp][nx*nz]."""
        nx, nz = self.grids[self.shots[shot][0]]
        nc = {WF_V: N_REC[self.physics], WF_U: N_REC[self.physics],
              WF_S: {SH: 2, PSV: 3, SPIN: 6}[self.physics]}[kind]
        if out is not None and _is_tensor(out):
            self._chk(self.L.sp_get_wavefield(self._p, shot, kind, _devptr(out, nc * nx * nz, self.device), 1))
            return out
        a = np.empty((nc, nx * nz), dtype=np.float32)
        self._chk(self.L.sp_get_wavefield(self._p, shot, kind, _ptr(a), 0))
        return a

    def records(self, shot, out=None):
        nrec = self.shots[shot][2]
        if out is not None and _is_tensor(out):
            self._chk(self.L.sp_get_records(self._p, shot, _devptr(out, N_REC[self.physics] * nrec * self.nt, self.device), 1))
            return out
        a = np.empty((N_REC[self.physics], nrec, self.nt), dtype=np.float32)
        self._chk(self.L.sp_get_records(self._p, shot, _ptr(a), 0))
        return a

    def set_data(self, shot, data):
        nrec = self.shots[shot][2]
        if _is_tensor(data):
            self._chk(self.L.sp_set_data(self._p, shot, _devptr(data, N_REC[self.physics] * nrec * self.nt, self.device), 1))
        else:
            a = _f32(data, N_REC[self.physics] * nrec * self.nt)
            self._chk(self.L.sp_set_data(self._p, shot, _ptr(a), 0))

    def misfit(self):
        out = np.zeros(len(self.shots) * N_MIS[self.physics], dtype=np.float64)
        self._chk(self.L.sp_misfit(self._p, _ptr(out)))
        return out.reshape(len(self.shots), N_MIS[self.physics])

    def forward_ckpt(self, interval):
        """Forward run of every shot keeping a full state every `interval` steps."""
        self._chk(self.L.sp_forward_ckpt(self._p, int(interval)))

    def adjoint_ckpt(self, sae):
        """Adjoint run recomputing the forward snapshots from the checkpoints."""
        self._chk(self.L.sp_adjoint_ckpt(self._p, int(sae)))

    def set_adjoint_source(self, shot, src):
        """[n_mis, nrec, nt], time-reversed; numpy or a device tensor."""
        nrec = self.shots[shot][2]
        if _is_tensor(src):
            self._chk(self.L.sp_set_adjoint_source(self._p, shot, _devptr(src, N_MIS[self.physics] * nrec * self.nt, self.device), 1))
        else:
            a = _f32(src, N_MIS[self.physics] * nrec * self.nt)
            self._chk(self.L.sp_set_adjoint_source(self._p, shot, _ptr(a), 0))

    def set_model_vps(self, grid, vp, vs, rho):
        nx, nz = self.grids[grid]
        if _is_tensor(vp):
            p = [_devptr(x, nx * nz, self.device) for x in (vp, vs, rho)]
            self._chk(self.L.sp_set_model_vps(self._p, grid, p[0], p[1], p[2], 1))
        else:
            a = [_f32(x, nx * nz) for x in (vp, vs, rho)]
            self._chk(self.L.sp_set_model_vps(self._p, grid, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), 0))

    def adjoint(self, sae):
        self._chk(self.L.sp_adjoint(self._p, int(sae)))

    def gradient(self, sae):
        out = np.zeros(len(self.shots) * N_MIS[self.physics], dtype=np.float64)
        self._chk(self.L.sp_gradient(self._p, int(sae), _ptr(out)))
        return out.reshape(len(self.shots), N_MIS[self.physics])

    def kernel(self, grid, which, out=None):
        nx, nz = self.grids[grid]
        if out is not None and _is_tensor(out):
            self._chk(self.L.sp_get_kernel(self._p, grid, which, _devptr(out, nx * nz, self.device), 1))
            return out
        a = np.empty(nx * nz, dtype=np.float32)
        self._chk(self.L.sp_get_kernel(self._p, grid, which, _ptr(a), 0))
        return a

    def kernels_vps(self, grid):
        self._chk(self.L.sp_kernels_vps(self._p, grid))

    def smooth_kernel(self, grid, which, sigma):
        self._chk(self.L.sp_smooth_kernel(self._p, grid, which, int(sigma)))

    def smooth(self, field, nx, nz, sigma):
        """Gaussian smoothing of a host array (returns a new array) or, for a
        torch tensor on the device, in place."""
        if _is_tensor(field):
            self._chk(self.L.sp_smooth(self._p, _devptr(field, int(nx) * int(nz), self.device), int(nx), int(nz), int(sigma), 1))
            return field
        a = _f32(field, nx * nz).copy()
        self._chk(self.L.sp_smooth(self._p, _ptr(a), int(nx), int(nz), int(sigma), 0))
        return a

    def smooth_noalloc(self, field, nx, nz, sigma):
        """The smoothing of smooth() in place on a device tensor, through the
        allocation-free entry point (scratch in a fixed device array; fields
        up to 8192 cells a side and 4,194,304 cells)."""
        self._chk(self.L.sp_smooth_noalloc(self._p, _devptr(field, int(nx) * int(nz), self.device),
                                           int(nx), int(nz), int(sigma)))
        return field

    def autotune(self, steps=20):
        """Pick the fastest rows-per-wavefront chunk and work order for the
        current shots; returns (chunk, order)."""
        ch, od = ctypes.c_int(), ctypes.c_int()
        self._chk(self.L.sp_autotune(self._p, int(steps), ctypes.byref(ch), ctypes.byref(od)))
        return ch.value, od.value

    def timing(self):
        n = ctypes.c_longlong()
        s, t, b = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._chk(self.L.sp_last_timing(self._p, ctypes.byref(n), ctypes.byref(s), ctypes.byref(t), ctypes.byref(b)))
        return {"step_launches": n.value, "step_ms": s.value, "total_ms": t.value, "alg_bytes": b.value}
