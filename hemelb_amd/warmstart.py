"""Learned warm start: a 3D U-Net, trained and run by the library's HIP
kernels (``include/hemelb_unet.h``), that maps a signed-distance wall field
plus iolet-parameter channels to the periodic-state velocity and pressure
of a vessel; the LB run then starts from the equilibrium of that field.

Input channels (per voxel of a cubic crop): signed distance to the wall in
voxels (positive in the fluid) divided by 16, inlet pressure amplitude
(mmHg), period / 1000 steps, tau, and cos / sin of the phase of the cycle.
Output channels: the velocity in units of 0.01 and (density - 1) / 0.01.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _native as N

CHANNELS = (16, 32, 64, 128)
IN_CHANNELS = 6
OUT_CHANNELS = 4
U_SCALE = 0.01
RHO_SCALE = 0.01

_SIGS = {
    "hlbn_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                              C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_void_p)]),
    "hlbn_destroy": (None, [C.c_void_p]),
    "hlbn_param_count": (C.c_int64, [C.c_void_p]),
    "hlbn_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hlbn_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hlbn_get_grads": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hlbn_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    "hlbn_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                C.POINTER(C.c_double), C.c_int32]),
    "hlbn_adam": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float,
                            C.c_float]),
    "hlbn_device_bytes": (C.c_int64, [C.c_void_p]),
    "hlbn_last_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "hlbn_stream": (C.c_void_p, [C.c_void_p]),
    "hlbn_last_error": (C.c_char_p, []),
}
_bound = False


def _lib():
    global _bound
    lib = N.lib()
    if not _bound:
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _bound = True
    return lib


def _check(rc):
    if rc != 0:
        raise N.HemeLBError(rc, (_lib().hlbn_last_error() or b"").decode())


def _ptr(a):
    if hasattr(a, "data_ptr"):        # a device tensor
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


LAYERS = ["e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "e4a", "e4b",
          "d3a", "d3b", "d2a", "d2b", "d1a", "d1b"]


def layer_shapes(chans=CHANNELS, cin_padded=16):
    c1, c2, c3, c4 = chans
    ci = [cin_padded, c1, c1, c2, c2, c3, c3, c4, c4 + c3, c3, c3 + c2, c2, c2 + c1, c1]
    co = [c1, c1, c2, c2, c3, c3, c4, c4, c3, c3, c2, c2, c1, c1]
    return list(zip(LAYERS, ci, co))


def pack_params(p: dict, chans=CHANNELS, cout=OUT_CHANNELS) -> np.ndarray:
    """Flat parameter vector (hemelb_unet.h order) from a dict of arrays
    (name.w [27,Co,Ci], name.b, head.w [Cout,C1], head.b); the first layer's
    weights are zero-padded to 16 input channels."""
    parts = []
    for name, ci, co in layer_shapes(chans):
        w = np.asarray(p[name + ".w"], np.float32)
        if w.shape[2] < ci:
            w = np.concatenate([w, np.zeros((27, co, ci - w.shape[2]), np.float32)], 2)
        parts += [w.reshape(-1), np.asarray(p[name + ".b"], np.float32)]
    parts += [np.asarray(p["head.w"], np.float32).reshape(-1),
              np.asarray(p["head.b"], np.float32)]
    return np.concatenate(parts)


def unpack(v: np.ndarray, chans=CHANNELS, cout=OUT_CHANNELS, cin=None) -> dict:
    out, k = {}, 0
    for name, ci, co in layer_shapes(chans):
        n = 27 * co * ci
        w = v[k:k + n].reshape(27, co, ci)
        out[name + ".w"] = w[:, :, :cin] if (cin is not None and name == "e1a") else w
        k += n
        out[name + ".b"] = v[k:k + co]
        k += co
    out["head.w"] = v[k:k + cout * chans[0]].reshape(cout, chans[0])
    k += cout * chans[0]
    out["head.b"] = v[k:k + cout]
    return out


def init_params(chans=CHANNELS, cin=IN_CHANNELS, cout=OUT_CHANNELS, seed=0) -> np.ndarray:
    """He-normal convolution weights, small head, zero biases."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co in layer_shapes(chans):
        fan = 27 * (cin if name == "e1a" else ci)
        w = rng.standard_normal((27, co, ci)) * math.sqrt(2.0 / fan)
        if name == "e1a":
            w[:, :, cin:] = 0.0
        p[name + ".w"] = w.astype(np.float32)
        p[name + ".b"] = np.zeros(co, np.float32)
    p["head.w"] = (rng.standard_normal((cout, chans[0])) * math.sqrt(1.0 / chans[0])
                   ).astype(np.float32)
    p["head.b"] = np.zeros(cout, np.float32)
    return pack_params(p, chans, cout)


class UNet:
    """The U-Net on one GPU for a fixed batch of cubic crops of side `side`."""

    def __init__(self, side: int, batch: int = 1, in_channels: int = IN_CHANNELS,
                 out_channels: int = OUT_CHANNELS, chans=CHANNELS, device: int = 0):
        self.side, self.batch = side, batch
        self.cin, self.cout, self.chans = in_channels, out_channels, tuple(chans)
        ch = (C.c_int32 * 4)(*chans)
        h = C.c_void_p()
        _check(_lib().hlbn_create(side, batch, in_channels, out_channels, ch, device,
                                  C.byref(h)))
        self._h = h
        self.nparam = int(_lib().hlbn_param_count(h))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib().hlbn_destroy(self._h)
            self._h = C.c_void_p()

    def set_params(self, v: np.ndarray) -> None:
        v = np.ascontiguousarray(v, np.float32)
        assert v.size == self.nparam
        _check(_lib().hlbn_set_params(self._h, _ptr(v)))

    def params(self) -> np.ndarray:
        v = np.empty(self.nparam, np.float32)
        _check(_lib().hlbn_get_params(self._h, _ptr(v)))
        return v

    def grads(self) -> np.ndarray:
        v = np.empty(self.nparam, np.float32)
        _check(_lib().hlbn_get_grads(self._h, _ptr(v)))
        return v

    def forward(self, x, out=None, device: bool = False):
        """x: [N,S,S,S,cin] binary32 (numpy, or a device tensor with
        device=True); returns the output array (numpy) or fills `out`."""
        if device:
            _check(_lib().hlbn_forward(self._h, _ptr(x), _ptr(out) if out is not None else None,
                                       1))
            return out
        x = np.ascontiguousarray(x, np.float32)
        assert x.shape == (self.batch, self.side, self.side, self.side, self.cin), x.shape
        y = np.empty((self.batch, self.side, self.side, self.side, self.cout), np.float32)
        _check(_lib().hlbn_forward(self._h, _ptr(x), _ptr(y), 0))
        return y

    def backward(self, target, mask, channel_weights=(1.0, 1.0, 1.0, 1.0), mask_count=None,
                 device: bool = False) -> float:
        w = np.ascontiguousarray(channel_weights, np.float32)
        loss = C.c_double()
        if device:
            cnt = float(mask_count)
            _check(_lib().hlbn_backward(self._h, _ptr(target), _ptr(mask), _ptr(w), cnt,
                                        C.byref(loss), 1))
            return loss.value
        t = np.ascontiguousarray(target, np.float32)
        m = np.ascontiguousarray(mask, np.uint8)
        cnt = float(m.sum()) if mask_count is None else float(mask_count)
        _check(_lib().hlbn_backward(self._h, _ptr(t), _ptr(m), _ptr(w), cnt, C.byref(loss), 0))
        return loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0) -> None:
        _check(_lib().hlbn_adam(self._h, lr, beta1, beta2, eps, weight_decay))

    def times(self):
        f, b = C.c_double(), C.c_double()
        _check(_lib().hlbn_last_times(self._h, C.byref(f), C.byref(b)))
        return f.value, b.value

    def device_bytes(self) -> int:
        return int(_lib().hlbn_device_bytes(self._h))
