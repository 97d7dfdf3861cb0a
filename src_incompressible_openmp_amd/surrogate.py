"""Python binding of the closure surrogate (include/mpcd_gpu.h, mpcd_gpu_net_*).

A 4-layer 3x3 convolutional network over the cell grid maps the red state's
cell features (occupancy deviation, cell velocity, temperature) to the liquid
closure's per-cell velocity correction; it runs on the GPU's bfloat16 matrix
cores with float32 sums, and trains there with the library's own forward,
backward and Adam kernels.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import gpu

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


def _lib():
    lib = gpu.load()
    if not getattr(lib, "_net_bound", False):
        lib.mpcd_gpu_net_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_void_p)]
        for name in ("mpcd_gpu_net_destroy", "mpcd_gpu_net_set_params", "mpcd_gpu_net_get_params",
                     "mpcd_gpu_net_forward", "mpcd_gpu_net_loss_grad", "mpcd_gpu_net_adam_step",
                     "mpcd_gpu_net_train"):
            getattr(lib, name).restype = ctypes.c_int
        lib._net_bound = True
    return lib


def num_params() -> int:
    return int(_lib().mpcd_gpu_net_num_params())


def _f32(a, shape_last):
    a = np.ascontiguousarray(a, np.float32)
    if a.shape[-1] != shape_last:
        raise ValueError(f"last axis must be {shape_last}")
    return a


class Net:
    """The surrogate on one grid shape; features [B, Nc, 4], outputs [B, Nc, 2]."""

    def __init__(self, Nx: int, Ny: int, perx: bool, pery: bool, max_batch: int = 1):
        self.Nx, self.Ny, self.Nc = Nx, Ny, Nx * Ny
        self.max_batch = max_batch
        self._h = ctypes.c_void_p()
        gpu._check(_lib().mpcd_gpu_net_create(Nx, Ny, int(perx), int(pery), max_batch, ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            _lib().mpcd_gpu_net_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, theta: np.ndarray) -> None:
        t = np.ascontiguousarray(theta, np.float64)
        if t.size != num_params():
            raise ValueError("parameter vector has the wrong size")
        gpu._check(_lib().mpcd_gpu_net_set_params(self._h, t.ctypes.data_as(_dp)))

    def get_params(self) -> np.ndarray:
        t = np.zeros(num_params())
        gpu._check(_lib().mpcd_gpu_net_get_params(self._h, t.ctypes.data_as(_dp)))
        return t

    def forward(self, feat: np.ndarray) -> np.ndarray:
        """outputs of any number of samples, max_batch at a time"""
        f = _f32(feat, 4).reshape(-1, self.Nc, 4)
        out = np.zeros((f.shape[0], self.Nc, 2), np.float32)
        for b0 in range(0, f.shape[0], self.max_batch):
            fb = np.ascontiguousarray(f[b0:b0 + self.max_batch])
            ob = np.zeros((fb.shape[0], self.Nc, 2), np.float32)
            gpu._check(_lib().mpcd_gpu_net_forward(self._h, fb.ctypes.data_as(_fp), fb.shape[0],
                                                   ob.ctypes.data_as(_fp)))
            out[b0:b0 + fb.shape[0]] = ob
        return out

    def loss_grad(self, feat: np.ndarray, target: np.ndarray):
        f = _f32(feat, 4).reshape(-1, self.Nc, 4)
        t = _f32(target, 2).reshape(-1, self.Nc, 2)
        loss = ctypes.c_double()
        g = np.zeros(num_params())
        gpu._check(_lib().mpcd_gpu_net_loss_grad(self._h, f.ctypes.data_as(_fp), t.ctypes.data_as(_fp), f.shape[0],
                                                 ctypes.byref(loss), g.ctypes.data_as(_dp)))
        return loss.value, g

    def adam_step(self, lr: float, grad: Optional[np.ndarray] = None) -> None:
        gp = None if grad is None else np.ascontiguousarray(grad, np.float64).ctypes.data_as(_dp)
        gpu._check(_lib().mpcd_gpu_net_adam_step(self._h, gp, ctypes.c_double(lr)))

    def train(self, feat: np.ndarray, target: np.ndarray, batch: int, iters: int, lr: float = 1e-3,
              seed: int = 1) -> np.ndarray:
        f = _f32(feat, 4).reshape(-1, self.Nc, 4)
        t = _f32(target, 2).reshape(-1, self.Nc, 2)
        losses = np.zeros(iters)
        gpu._check(_lib().mpcd_gpu_net_train(self._h, f.ctypes.data_as(_fp), t.ctypes.data_as(_fp), f.shape[0],
                                             batch, iters, ctypes.c_double(lr), ctypes.c_uint64(seed),
                                             losses.ctypes.data_as(_dp)))
        return losses


def mfma_probe(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Runs one 16x16x32 bf16 product: a, b are [64 lanes, 8] raw bf16 (uint16)."""
    a = np.ascontiguousarray(a, np.uint16)
    b = np.ascontiguousarray(b, np.uint16)
    d = np.zeros((64, 4), np.float32)
    u16p = ctypes.POINTER(ctypes.c_uint16)
    gpu._check(_lib().mpcd_gpu_net_mfma_probe(a.ctypes.data_as(u16p), b.ctypes.data_as(u16p), d.ctypes.data_as(_fp)))
    return d
