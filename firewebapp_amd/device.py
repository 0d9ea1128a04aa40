"""Device-side plumbing: grids, array marshalling, streams.

PyTorch-ROCm provides device memory and streams; every computation is a
call into libfirewave.so.  Arrays handed to the reference-style call surface
(``firewebapp_amd.physics_3d``) may be NumPy arrays (copied to the GPU and
back, so the reference's in-place semantics hold) or CUDA tensors (used in
place, zero copy).  float64 inputs select the FP64 kernels, float32 the FP32
kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

_DEV = None


def device():
    global _DEV
    if _DEV is None:
        if not torch.cuda.is_available():
            raise RuntimeError("firewebapp_amd needs a ROCm GPU (torch.cuda unavailable)")
        _DEV = torch.device("cuda", torch.cuda.current_device())
    return _DEV


def stream_ptr():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def fw_dtype(tdtype):
    if tdtype == torch.float64:
        return _lib.FW_F64
    if tdtype == torch.float32:
        return _lib.FW_F32
    raise TypeError(f"unsupported dtype {tdtype}")


class DeviceGrid:
    """fw_grid plus the device copies of the 1-D metric arrays it points to."""

    def __init__(self, nx, ny, nz, dx, dy, dz_arr, d_face_above, d_face_below,
                 z_face=None, n_z_bed=0, dtype=torch.float32):
        dev = device()
        f64 = dict(dtype=torch.float64, device=dev)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy = float(dx), float(dy)
        self.dtype = dtype
        self.dz = torch.as_tensor(np.asarray(dz_arr, np.float64), **f64).contiguous()
        self.d_above = torch.as_tensor(np.asarray(d_face_above, np.float64), **f64).contiguous()
        self.d_below = torch.as_tensor(np.asarray(d_face_below, np.float64), **f64).contiguous()
        if z_face is None:
            zf = np.concatenate([[0.0], np.cumsum(np.asarray(dz_arr, np.float64))])
        else:
            zf = np.asarray(z_face, np.float64)
        self.z_face = torch.as_tensor(zf, **f64).contiguous()
        self.n_z_bed = int(n_z_bed)
        # float64 reciprocals, IEEE quotients from the host (the kernels multiply
        # by them instead of dividing)
        inv = lambda a: torch.as_tensor(1.0 / np.asarray(a, np.float64), **f64).contiguous()
        self.inv_dz = inv(dz_arr)
        self.inv_d_above = inv(d_face_above)
        self.inv_d_below = inv(d_face_below)
        self.c = _lib.FwGrid(self.nx, self.ny, self.nz, fw_dtype(dtype),
                             self.dx, self.dy, self.dz.data_ptr(),
                             self.d_above.data_ptr(), self.d_below.data_ptr(),
                             self.z_face.data_ptr(), self.n_z_bed, 0,
                             self.inv_dz.data_ptr(), self.inv_d_above.data_ptr(),
                             self.inv_d_below.data_ptr())

    @property
    def ref(self):
        return C.byref(self.c)

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def ncell(self):
        return self.nx * self.ny * self.nz

    def empty(self, *lead, dtype=None):
        return torch.empty(*lead, *self.shape, dtype=dtype or self.dtype,
                           device=device())

    def zeros(self, *lead, dtype=None):
        return torch.zeros(*lead, *self.shape, dtype=dtype or self.dtype,
                           device=device())

    def with_dtype(self, dtype):
        if dtype == self.dtype:
            return self
        g = DeviceGrid.__new__(DeviceGrid)
        g.__dict__.update(self.__dict__)
        g.dtype = dtype
        g.c = _lib.FwGrid(self.nx, self.ny, self.nz, fw_dtype(dtype), self.dx,
                          self.dy, self.dz.data_ptr(), self.d_above.data_ptr(),
                          self.d_below.data_ptr(), self.z_face.data_ptr(),
                          self.n_z_bed, 0, self.inv_dz.data_ptr(),
                          self.inv_d_above.data_ptr(), self.inv_d_below.data_ptr())
        return g


class Marshal:
    """Moves reference-style arguments to the device and back.

    ``dev(a)`` returns a contiguous CUDA tensor of the working dtype for a
    NumPy array or tensor; ``writeback()`` copies results into the NumPy
    arrays that were registered with ``out(a)``.
    """

    def __init__(self, like):
        if isinstance(like, torch.Tensor):
            self.dtype = like.dtype
        else:
            self.dtype = torch.float64 if np.asarray(like).dtype == np.float64 else torch.float32
        self._back = []
        # Device copies must outlive the kernel launch that reads them: a
        # temporary freed after `ptr()` would let the caching allocator hand
        # its block to the next copy before the kernel has run.
        self._keep = []

    def dev(self, a, dtype=None, shape=None):
        """Device tensor for `a`; with `shape`, a size mismatch raises
        ValueError before any kernel can read past the buffer."""
        dt = dtype or self.dtype
        if shape is not None and tuple(a.shape) != tuple(shape):
            raise ValueError(f"array of shape {tuple(a.shape)} where {tuple(shape)} "
                             "is required")
        if isinstance(a, torch.Tensor):
            if a.is_cuda and a.dtype == dt and a.is_contiguous():
                return a
            t = a.to(device=device(), dtype=dt).contiguous()
        else:
            t = torch.as_tensor(np.ascontiguousarray(a), device=device()).to(dt).contiguous()
        self._keep.append(t)
        return t

    def out(self, a, dtype=None, shape=None):
        """Device buffer for an in/out argument, written back on writeback()."""
        t = self.dev(a, dtype, shape)
        if not (isinstance(a, torch.Tensor) and t.data_ptr() == a.data_ptr()):
            self._back.append((a, t))
        return t

    def writeback(self):
        torch.cuda.current_stream().synchronize()
        for a, t in self._back:
            if isinstance(a, torch.Tensor):
                a.copy_(t)
            else:
                a[...] = t.cpu().numpy().astype(a.dtype, copy=False)
        self._back.clear()
        self._keep.clear()


def grid_for(m: Marshal, shape, dx, dy, dz_arr, d_face_above, d_face_below,
             z_face=None, n_z_bed=0):
    nz, ny, nx = shape
    return DeviceGrid(nx, ny, nz, dx, dy, dz_arr, d_face_above, d_face_below,
                      z_face=z_face, n_z_bed=n_z_bed, dtype=m.dtype)


class _CudaArray:
    def __init__(self, addr, shape, dtype):
        typestr = {torch.float32: "<f4", torch.float64: "<f8",
                   torch.int32: "<i4", torch.uint8: "|u1"}[dtype]
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr,
                                         "data": (int(addr), False), "version": 2,
                                         "strides": None}


def device_view(addr, shape, dtype):
    """Zero-copy torch view of device memory owned by a libfirewave handle."""
    return torch.as_tensor(_CudaArray(addr, shape, dtype), device=device())
