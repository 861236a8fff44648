"""Device-array helpers for the GPU tests (torch provides device memory only)."""
from __future__ import annotations

import numpy as np


def to_device(torch, x: np.ndarray):
    """Copy a column-major numpy array to the device; returns (tensor, DeviceArray)."""
    from fftahip.plan import DeviceArray

    xf = np.asfortranarray(x)
    flat = np.ascontiguousarray(xf.ravel(order="F"))
    t = torch.from_numpy(flat).to("cuda")
    prec = 32 if xf.dtype == np.complex64 else 64
    return t, DeviceArray(t.data_ptr(), xf.shape, prec)


def empty_device(torch, shape, dtype):
    from fftahip.plan import DeviceArray

    n = int(np.prod(shape)) if len(shape) else 1
    tdt = torch.complex64 if np.dtype(dtype) == np.complex64 else torch.complex128
    t = torch.empty(n, dtype=tdt, device="cuda")
    prec = 32 if np.dtype(dtype) == np.complex64 else 64
    return t, DeviceArray(t.data_ptr(), shape, prec)


def to_host(torch, t, shape) -> np.ndarray:
    torch.cuda.synchronize()
    flat = t.cpu().numpy()
    return flat.reshape(shape, order="F")
