"""Forward (type-2) and adjoint (type-1) operators and q-point generation,
drop-in for the parts of spinifel/autocorrelation.py on the orientation-matching
and merge paths.

``forward`` keeps the reference signature ``forward(ugrid, H_, K_, L_,
support, use_recip_sym, N)`` (spinifel/autocorrelation.py:32-61) with its
checks: the domain assert ``max |H_, K_, L_| < 3 pi`` (:46), the realness
assert when ``use_recip_sym`` (:49-50) and ``ugrid *= support`` (:53-54).
The transform runs on the GPU (libmtip); only its real part is computed
because that is all the reference uses (orientation_matching.py:241,
slicing.py:43), so the result is returned as a real float64 array
(``.real`` on it still works).
"""
import ctypes

import numpy as np
import torch

from . import _lib, engine
from .settings import settings as _settings

_PLANS = {}


def get_plan(M, eps=None, device=None):
    """Cached NUFFT plan per (M, eps, device) -- the persistent-plan reuse of
    spinifel/extern/nufft_ext.py:271-280."""
    eps = _settings.nufft_eps if eps is None else eps
    device = _settings.device if device is None else device
    key = (int(M), float(eps), int(device))
    if key not in _PLANS:
        _PLANS[key] = engine.NufftPlan(M, eps, device)
    return _PLANS[key]


def forward(ugrid, H_, K_, L_, support, use_recip_sym, N, eps=None):
    """Re of the type-2 NUFFT of ``ugrid`` at (H_, K_, L_) -> float64 (N,)."""
    H_ = np.asarray(H_, dtype=np.float64).ravel()
    K_ = np.asarray(K_, dtype=np.float64).ravel()
    L_ = np.asarray(L_, dtype=np.float64).ravel()
    assert H_.shape == K_.shape == L_.shape
    assert np.max(np.abs(np.array([H_, K_, L_]))) < 3 * np.pi
    ugrid = np.asarray(ugrid)
    if use_recip_sym:
        assert np.all(np.isreal(ugrid))
    if support is not None:
        ugrid = ugrid * support
    ugrid = np.real(ugrid)
    if ugrid.ndim != 3 or len(set(ugrid.shape)) != 1:
        raise ValueError(f"model must be cubic, got {ugrid.shape}")
    plan = get_plan(ugrid.shape[0], eps)
    plan.set_model(ugrid)
    out = plan.execute(H_, K_, L_).double().cpu().numpy()
    if out.shape[0] != N:
        raise ValueError(f"N = {N} but {out.shape[0]} points were given")
    return out


def adjoint(nuvect, H_, K_, L_, M, use_recip_sym=True, support=None, eps=None):
    """Re of the type-1 NUFFT (isign +1) of ``nuvect`` at (H_, K_, L_) onto the
    M^3 grid, times ``support``, divided by M^3 -> float64 (M, M, M).

    Reference signature and checks: spinifel/autocorrelation.py:97-128 (domain
    assert :114, support :121-122, real part :124-125, / M**3 :127).  The GPU
    path computes the real part of the transform of real data, which is what
    every reference caller asks for (real slices, use_recip_sym=True,
    sequential/autocorrelation.py:158-166, :195-203); complex data or
    ``use_recip_sym=False`` raise NotImplementedError rather than return a
    different result.
    """
    H_ = np.asarray(H_, dtype=np.float64).ravel()
    K_ = np.asarray(K_, dtype=np.float64).ravel()
    L_ = np.asarray(L_, dtype=np.float64).ravel()
    assert H_.shape == K_.shape == L_.shape
    assert np.max(np.abs(np.array([H_, K_, L_]))) < 3 * np.pi
    nuvect = np.asarray(nuvect).ravel()
    if np.iscomplexobj(nuvect):
        if np.any(nuvect.imag != 0):
            raise NotImplementedError("adjoint of complex data (GPU path takes real data)")
        nuvect = nuvect.real
    if not use_recip_sym:
        raise NotImplementedError("use_recip_sym=False (GPU path returns the real part)")
    if nuvect.shape[0] != H_.shape[0]:
        raise ValueError(f"{nuvect.shape[0]} values for {H_.shape[0]} points")
    plan = get_plan(int(M), eps)
    out = plan.adjoint_points(H_, K_, L_, values=nuvect.astype(np.float32),
                              support=None if support is None else
                              np.ascontiguousarray(np.broadcast_to(
                                  np.asarray(support, np.float64), (int(M),) * 3)))
    return out.cpu().numpy()


def _rotate(orientations, pixel_position_reciprocal, scale):
    lib = _lib.load()
    engine.require_gpu()
    dev = torch.device("cuda", _settings.device)
    quats = torch.as_tensor(np.ascontiguousarray(orientations, np.float64).reshape(-1, 4)).to(dev)
    ppr = np.asarray(pixel_position_reciprocal, np.float64)
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1))).to(dev)
    N_o, N_p = quats.shape[0], q.shape[1]
    outs = [torch.empty(N_o * N_p, dtype=torch.float64, device=dev) for _ in range(3)]
    st = engine._stream()
    for s in range(0, N_o, 65535):
        e = min(N_o, s + 65535)
        _lib.check(lib.mtip_rotate_positions(
            ctypes.c_void_p(quats[s:].data_ptr()), e - s, engine._ptr(q), N_p, float(scale),
            ctypes.c_void_p(outs[0][s * N_p:].data_ptr()),
            ctypes.c_void_p(outs[1][s * N_p:].data_ptr()),
            ctypes.c_void_p(outs[2][s * N_p:].data_ptr()), st), "mtip_rotate_positions")
    shape = (N_o,) + ppr.shape[1:]
    return tuple(o.cpu().numpy().reshape(shape) for o in outs)


def gen_nonuniform_positions(orientations, pixel_position_reciprocal):
    """H, K, L of shape (N_orientations,) + det_shape (autocorrelation.py:290-305);
    rotation in FP64 on the GPU."""
    orientations = np.asarray(orientations)
    if orientations.shape[0] == 0:
        print("WARNING: gen_nonuniform_positions got empty orientation - returning h,k,l for Null rotation")
        ppr = np.asarray(pixel_position_reciprocal)
        z = np.zeros((0,) + ppr.shape[1:])
        return z, z.copy(), z.copy()
    return _rotate(orientations, pixel_position_reciprocal, 1.0)


def gen_nonuniform_normalized_positions(orientations, pixel_position_reciprocal,
                                        reciprocal_extent, oversampling):
    """Flattened H_, K_, L_ scaled by pi / (E oversampling) (autocorrelation.py:310-323)."""
    H, K, L = gen_nonuniform_positions(orientations, pixel_position_reciprocal)
    H_ = H.flatten() / reciprocal_extent * np.pi / oversampling
    K_ = K.flatten() / reciprocal_extent * np.pi / oversampling
    L_ = L.flatten() / reciprocal_extent * np.pi / oversampling
    return H_, K_, L_
