"""Model-slice generation, drop-in for spinifel/slicing.py.

``gen_model_slices`` and ``gen_model_slices_batch`` keep the reference
signatures (spinifel/slicing.py:10-13, :50-59).  The reference body of
``gen_model_slices`` cannot run (it passes nine arguments to the
seven-parameter ``autocorrelation.forward``, and ``forward_cpu`` /
``forward_gpu`` are defined nowhere); this implements its evident intent:
Re of the type-2 NUFFT of ``ac`` at the rotated, normalised pixel q-points,
flattened orientation-major to ``(N_orientations * N_pixels,)``.

``override_forward_with``: ``None`` and ``'gpu'`` run the HIP path;
``'cpu'`` raises -- this build has no CPU slicing path.
``ac_support_size`` is unused, as in the reference (:24-26).
"""
import numpy as np
import torch

from . import engine
from .autocorrelation import get_plan


def _slices_device(ac, ref_orientations, pixel_position_reciprocal, reciprocal_extent,
                   oversampling, N_pixels, eps=None):
    ac = np.asarray(ac)
    assert np.all(np.isreal(ac))
    ac = np.real(ac)
    plan = get_plan(ac.shape[0], eps)
    plan.set_model(ac)
    quats = torch.as_tensor(np.ascontiguousarray(ref_orientations, np.float64).reshape(-1, 4))
    ppr = np.asarray(pixel_position_reciprocal, np.float64)
    q = torch.as_tensor(np.ascontiguousarray(ppr.reshape(3, -1)))
    if q.shape[1] != N_pixels:
        raise ValueError(f"N_pixels = {N_pixels} but the detector has {q.shape[1]} pixels")
    scale = np.pi / (float(reciprocal_extent) * float(oversampling))
    r = plan.slices(quats, q, scale, ld=N_pixels, order=engine.auto_tiling(plan, ppr.shape[1:]))
    if int(r["bad"].item()) > 0:
        raise AssertionError("NU points outside [-3pi, 3pi)")
    return r["s32"]


def gen_model_slices(ac, ref_orientations, pixel_position_reciprocal, reciprocal_extent,
                     oversampling, ac_support_size, N_pixels, override_forward_with=None):
    """Slices of ``ac`` at the given quaternions -> float64 (N_o * N_pixels,)."""
    if override_forward_with not in (None, "gpu"):
        if override_forward_with == "cpu":
            raise NotImplementedError("no CPU slicing path in this build (GPU only)")
        raise ValueError(f"override_forward_with={override_forward_with!r}")
    s = _slices_device(ac, ref_orientations, pixel_position_reciprocal, reciprocal_extent,
                       oversampling, N_pixels)
    return s.double().cpu().numpy().reshape(-1)


def gen_model_slices_batch(ac, ref_orientations, pixel_position_reciprocal, reciprocal_extent,
                           oversampling, ac_support_size, N_pixels, batch_size=None,
                           override_forward_with=None):
    """Batched variant (slicing.py:50-95): same values, float64, asserts divisibility."""
    ref_orientations = np.asarray(ref_orientations)
    N_orientations = ref_orientations.shape[0]
    N_batch_size = N_orientations if batch_size is None else batch_size
    assert N_orientations % N_batch_size == 0, \
        f"N_orientations ({N_orientations}) must be divisible by N_batch_size ({N_batch_size})"
    out = np.zeros((N_pixels * N_orientations,))
    for i in range(N_orientations // N_batch_size):
        st, en = i * N_batch_size, (i + 1) * N_batch_size
        out[st * N_pixels:en * N_pixels] = gen_model_slices(
            ac, ref_orientations[st:en], pixel_position_reciprocal, reciprocal_extent,
            oversampling, ac_support_size, N_pixels, override_forward_with)
    return out
