"""Helpers shared by the GPU parity tests: device buffers (torch on the ROCm
device is only the allocator and copier), pointer arrays and checks."""
import ctypes

import numpy as np

EPS64 = np.finfo(np.float64).eps   # 2^-52
EPS32 = np.finfo(np.float32).eps


def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch


def to_dev(a, dtype=None):
    torch = torch_dev()
    a = np.ascontiguousarray(a if dtype is None else a.astype(dtype))
    return torch.from_numpy(a).to("cuda")


def to_host(t):
    torch = torch_dev()
    torch.cuda.synchronize()
    return t.cpu().numpy()


def ptr_array(t, count, stride_elems):
    """device array of device pointers base + i*stride (i < count)"""
    torch = torch_dev()
    base = t.data_ptr()
    item = t.element_size()
    addrs = np.array([base + i * stride_elems * item for i in range(count)], dtype=np.int64)
    return torch.from_numpy(addrs).to("cuda")


def handle():
    import pykblas
    torch = torch_dev()
    h = pykblas.Handle()
    h.set_stream(torch.cuda.current_stream().cuda_stream)
    return h


def P(x):
    return ctypes.c_void_p(x if isinstance(x, int) else x.data_ptr())
