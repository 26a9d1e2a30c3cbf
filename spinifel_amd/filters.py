"""scipy.ndimage.gaussian_filter on device tensors (FP64, separable), for the
three boundary modes the merge and phasing steps use: 'reflect' (the
default, spinifel/sequential/autocorrelation.py:201), 'constant' (cval 0,
spinifel/sequential/phasing.py:258-259) and 'wrap' (:215-216).  The kernel
is scipy's: weights exp(-x^2 / 2 sigma^2) on |x| <= int(truncate sigma + 0.5),
normalised to sum 1, applied along each axis in turn."""
import numpy as np
import torch


def _index(n, radius, mode, device):
    i = np.arange(-radius, n + radius)
    valid = None
    if mode == "reflect":
        i = np.mod(i, 2 * n)
        i = np.where(i >= n, 2 * n - 1 - i, i)
    elif mode == "wrap":
        i = np.mod(i, n)
    elif mode == "constant":
        valid = (i >= 0) & (i < n)
        i = np.clip(i, 0, n - 1)
    else:
        raise ValueError(f"mode {mode!r}: 'reflect', 'wrap' or 'constant'")
    idx = torch.as_tensor(i, dtype=torch.long, device=device)
    if valid is not None:
        valid = torch.as_tensor(valid, dtype=torch.float64, device=device)
    return idx, valid


def gaussian_filter(a, sigma, mode="reflect", truncate=4.0):
    """gaussian_filter(a, sigma, mode=mode, truncate=truncate) of a device tensor."""
    radius = int(truncate * float(sigma) + 0.5)
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    w = np.exp(-0.5 * (x / sigma) ** 2)
    w /= w.sum()
    out = a.to(torch.float64)
    for ax in range(out.dim()):
        n = out.shape[ax]
        idx, valid = _index(n, radius, mode, out.device)
        pad = out.index_select(ax, idx)
        if valid is not None:
            shape = [1] * out.dim()
            shape[ax] = valid.numel()
            pad = pad * valid.view(shape)
        acc = torch.zeros_like(out)
        for t in range(2 * radius + 1):
            acc += float(w[t]) * pad.narrow(ax, t, n)
        out = acc
    return out
