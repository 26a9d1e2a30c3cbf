"""Data preparation, drop-in for spinifel/prep.py (and the binning part of
spinifel/sequential/prep.py:get_data).

``clipping`` followed by 2^n x 2^n binning runs on the GPU in one pass
(``mtip_clip_bin``): ``binning_sum`` for the images, ``binning_mean`` for the
pixel positions, ``binning_index`` for the pixel index map, with the
reference's reduction order and types (float results equal numpy's bit for
bit).  Functions take numpy arrays (returned as numpy, like the reference)
or device tensors (returned on the device, for a pipeline that uploads raw
frames once).

The HDF5 loaders (prep.py:142-187: ``pixel_position_reciprocal`` 3-first or
3-last, ``pixel_index_map``, ``intensities``, ``orientations``, ``volume``)
keep the reference's logic but need h5py, which is not installed in this
image: they raise ImportError naming it.  The SAXS/PNG/MRC exports are not
part of the data path and are left out.
"""
import sys

import numpy as np
import torch

from . import _lib, engine
from .settings import settings as _settings

_DTYPES = {torch.float32: 0, torch.float64: 1, torch.int32: 2, torch.int64: 3}
_MODES = {"sum": 0, "mean": 1, "index": 2}


def _device():
    return torch.device("cuda", int(getattr(_settings, "device", 0)))


def clip_bin(arr, n_clip, n_bin, mode):
    """clipping(arr, n_clip) then bin2nx2n_<mode>(., n_bin) on the GPU.
    arr: (..., panels, sa, sb); numpy in -> numpy out, tensor in -> tensor out."""
    as_numpy = not torch.is_tensor(arr)
    t = torch.as_tensor(np.ascontiguousarray(arr)) if as_numpy else arr
    engine.require_gpu()
    t = t.to(_device())
    if t.dtype not in _DTYPES:
        t = t.to(torch.float64 if t.is_floating_point() else torch.int64)
    t = t.contiguous()
    if t.dim() < 3:
        raise ValueError("clip_bin needs (..., panels, sa, sb)")
    *lead, panels, sa, sb = t.shape
    if panels not in (1, 4):
        print("Clipping function doesn't currently accept a detector with %i panels" % panels)
    n_lead = int(np.prod(lead)) if lead else 1
    k = n_clip + n_bin
    out = torch.empty(tuple(lead) + (panels, sa >> k, sb >> k), dtype=t.dtype, device=t.device)
    _lib.check(_lib.load().mtip_clip_bin(engine._ptr(t), _DTYPES[t.dtype], n_lead, panels, sa,
                                         sb, int(n_clip), int(n_bin), _MODES[mode],
                                         engine._ptr(out), engine._stream()), "mtip_clip_bin")
    return out.cpu().numpy() if as_numpy else out


def bin2nx2n_sum(arr, n):
    return clip_bin(arr, 0, n, "sum")


def bin2nx2n_mean(arr, n):
    return clip_bin(arr, 0, n, "mean")


def bin2nx2n_index(arr, n):
    return clip_bin(arr, 0, n, "index")


def bin2x2_sum(arr):
    return clip_bin(arr, 0, 1, "sum")


def bin2x2_mean(arr):
    return clip_bin(arr, 0, 1, "mean")


def bin2x2_index(arr):
    return clip_bin(arr, 0, 1, "index")


def clipping(arr, n):
    return clip_bin(arr, n, 0, "sum")


def clipping_index(arr, n):
    """prep.py:108-113: clip, then subtract the minimum."""
    out = clipping(arr, n)
    out -= out.min()
    return out


def binning_sum(arr, settings=None):
    s = _settings if settings is None else settings
    return clip_bin(arr, s.N_clipping, s.N_binning, "sum")


def binning_mean(arr, settings=None):
    s = _settings if settings is None else settings
    return clip_bin(arr, s.N_clipping, s.N_binning, "mean")


def binning_index(arr, settings=None):
    s = _settings if settings is None else settings
    # the minimum is subtracted after clipping, before binning (prep.py:131-134)
    c = clipping_index(arr, s.N_clipping)
    return clip_bin(c, 0, s.N_binning, "index")


def compute_pixel_distance(pixel_position_reciprocal):
    """prep.py:170-172 -> |q| per pixel."""
    p = pixel_position_reciprocal
    if torch.is_tensor(p):
        return torch.sqrt((p ** 2).sum(dim=0))
    return np.sqrt((np.asarray(p) ** 2).sum(axis=0))


# ------------------------------------------------------------------ HDF5

def _h5py():
    try:
        import h5py  # noqa: F401
    except ImportError as e:      # not installed in this image
        raise ImportError("spinifel_amd.prep loaders need h5py (the reference reads its "
                          "HDF5 input with it); it is not installed here") from e
    return sys.modules["h5py"]


def load_pixel_position_reciprocal(pixel_position_reciprocal, data_path):
    """prep.py:142-150: 3-first or 3-last on disk, 3-first in memory."""
    with _h5py().File(data_path, "r") as h5f:
        p = h5f["pixel_position_reciprocal"][:]
        pixel_position_reciprocal[:] = p if p.shape[0] == 3 else np.moveaxis(p, -1, 0)


def load_pixel_index_map(pixel_index_map, data_path):
    """prep.py:153-157."""
    with _h5py().File(data_path, "r") as h5f:
        pixel_index_map[:] = np.moveaxis(h5f["pixel_index_map"][:], -1, 0)


def load_slices(slices, i_start, i_end, data_path):
    """prep.py:160-168 (exits like the reference when the file is short)."""
    with _h5py().File(data_path, "r") as h5f:
        if h5f["intensities"].shape[0] >= i_end:
            slices[:] = h5f["intensities"][i_start:i_end]
        else:
            sys.exit(f"Error: Not enough intensity slices (max:{h5f['intensities'].shape[0]:d})")


def load_orientations(orientations, i_start, i_end, data_path):
    """prep.py:171-174."""
    with _h5py().File(data_path, "r") as h5f:
        orientations[:] = h5f["orientations"][i_start:i_end]
