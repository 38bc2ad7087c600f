"""ctypes binding of the step conditioner's network (include/seispie_cnn.h).

The network runs only through the HIP kernels of libseispie_hip.so; there
is no CPU path.  Arrays are float32 [C][H][W]; numpy arrays are copied
through the host, torch tensors on the GPU are passed as device pointers.
"""
import ctypes

import numpy as np

from . import _lib


def _load():
    L = _lib.load()
    if getattr(L, "_cnn_bound", False):
        return L
    vp, i, d, f, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_longlong
    sig = {
        "sp_cnn_create": ([i, i, vp, ctypes.POINTER(vp)], i),
        "sp_cnn_destroy": ([vp], i),
        "sp_cnn_last_error": ([vp], ctypes.c_char_p),
        "sp_cnn_set_stream": ([vp, vp], i),
        "sp_cnn_sync": ([vp], i),
        "sp_cnn_num_params": ([vp], ll),
        "sp_cnn_set_params": ([vp, vp, i], i),
        "sp_cnn_get_params": ([vp, vp, i], i),
        "sp_cnn_forward": ([vp, vp, vp, i, i, i], i),
        "sp_cnn_loss_grad": ([vp, vp, vp, vp, i, i, i, ctypes.POINTER(d)], i),
        "sp_cnn_get_grad": ([vp, vp, i], i),
        "sp_cnn_adam_step": ([vp, f, f, f, f], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    L._cnn_bound = True
    return L


def exported_symbols():
    return ["sp_cnn_create", "sp_cnn_destroy", "sp_cnn_last_error", "sp_cnn_set_stream", "sp_cnn_sync",
            "sp_cnn_num_params", "sp_cnn_set_params", "sp_cnn_get_params", "sp_cnn_forward", "sp_cnn_loss_grad",
            "sp_cnn_get_grad", "sp_cnn_adam_step"]


class Net:
    """3x3 convolutions with ReLU between layers; channels = [c_in, c_1, ..., c_out]."""

    def __init__(self, channels, device=0, params=None, seed=0):
        self.L = _load()
        self.channels = [int(c) for c in channels]
        arr = (ctypes.c_int * len(self.channels))(*self.channels)
        p = ctypes.c_void_p()
        rc = self.L.sp_cnn_create(int(device), len(self.channels) - 1, arr, ctypes.byref(p))
        if rc != 0:
            raise _lib.SeispieError("sp_cnn_create failed with code %d (channels %s)" % (rc, self.channels))
        self._p = p
        self.nparam = int(self.L.sp_cnn_num_params(p))
        if params is None:
            params = init_params(self.channels, seed)
        self.set_params(params)

    def _chk(self, rc):
        if rc != 0:
            msg = self.L.sp_cnn_last_error(self._p)
            raise _lib.SeispieError((msg or b"error %d" % rc).decode())

    def close(self):
        if getattr(self, "_p", None):
            self.L.sp_cnn_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, handle):
        self._chk(self.L.sp_cnn_set_stream(self._p, ctypes.c_void_p(handle)))

    def sync(self):
        self._chk(self.L.sp_cnn_sync(self._p))

    def set_params(self, params):
        a = np.ascontiguousarray(params, np.float32)
        if a.size != self.nparam:
            raise ValueError("expected %d parameters, got %d" % (self.nparam, a.size))
        self._chk(self.L.sp_cnn_set_params(self._p, a.ctypes.data, 0))

    def params(self):
        a = np.empty(self.nparam, np.float32)
        self._chk(self.L.sp_cnn_get_params(self._p, a.ctypes.data, 0))
        return a

    def grad(self):
        a = np.empty(self.nparam, np.float32)
        self._chk(self.L.sp_cnn_get_grad(self._p, a.ctypes.data, 0))
        return a

    def forward(self, x):
        """x [c_in][H][W] -> y [c_out][H][W] (numpy in, numpy out)."""
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim != 3 or x.shape[0] != self.channels[0]:
            raise ValueError("input must be [%d][H][W]" % self.channels[0])
        _, H, W = x.shape
        y = np.empty((self.channels[-1], H, W), np.float32)
        self._chk(self.L.sp_cnn_forward(self._p, x.ctypes.data, y.ctypes.data, H, W, 0))
        return y

    def loss_grad(self, x, target, mask=None):
        """Masked MSE of the one-channel output against target [H][W];
        returns the loss, keeps the gradient for adam_step / grad()."""
        x = np.ascontiguousarray(x, np.float32)
        _, H, W = x.shape
        t = np.ascontiguousarray(target, np.float32).reshape(H, W)
        m = None if mask is None else np.ascontiguousarray(mask, np.float32).reshape(H, W)
        loss = ctypes.c_double()
        self._chk(self.L.sp_cnn_loss_grad(self._p, x.ctypes.data, t.ctypes.data,
                                          None if m is None else m.ctypes.data, H, W, 0, ctypes.byref(loss)))
        return loss.value

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self._chk(self.L.sp_cnn_adam_step(self._p, lr, beta1, beta2, eps))


def init_params(channels, seed=0):
    """He-normal weights, zero biases, in the layout of include/seispie_cnn.h."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in zip(channels[:-1], channels[1:]):
        parts.append(rng.standard_normal(ci * 9 * co) * np.sqrt(2.0 / (ci * 9)))
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)
