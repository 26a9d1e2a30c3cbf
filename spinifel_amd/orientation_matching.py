"""Orientation matching, drop-in for spinifel's ``match`` /
``slicing_and_match`` (spinifel/mpi/orientation_matching.py:8-16 ->
spinifel/sequential/orientation_matching.py:189-256).

Same signature, same settings fields read from the settings singleton
(``N_orientations``, ``N_batch_size``, ``reduced_det_shape``,
``oversampling``), same checks and return value:

* ``assert slices_.shape == (N_slices,) + reduced_det_shape`` (:210);
* ``np.zeros((0, 4))`` when there are no images (:213-214);
* ``assert N_orientations % N_batch_size == 0`` (:224);
* ``E = pixel_distance_reciprocal.max()`` (:221), q scaled by pi / (E os);
* model slices scaled by the global ratio ``slices_.std() / model.std()``
  (:245-248), computed in FP64 on the device;
* ``ref_orientations[argmin]`` per image (:252-256), first index on ties.

Everything between the inputs and the index runs on the GPU through
libmtip (FP64 rotation, ES-kernel NUFFT, MFMA cross term with fused top-K,
optional FP32/FP64 rescoring).  ``ref_orientations=None`` uses this
package's super-Fibonacci grid, not skopi's ``get_uniform_quat`` (skopi is
absent, so that grid cannot be reproduced).
"""
import time

import numpy as np
import torch

from . import engine, quaternions
from .autocorrelation import get_plan
from .settings import settings as _settings


def resolve_form(s, precision=None, topk=None, rescore=None):
    """(precision, K, rescore) of a match from the arguments and the settings.

    BF16 / FP8: top-8 rescored in FP32 then FP64.  FP32: with
    ``use_single_prec`` the plain FP32 argmin (K = 1, no rescoring); without
    it (the reference's default, float64 distances,
    spinifel/sequential/orientation_matching.py:32-37) the top-4 of the FP32
    cross term are rescored in FP64, so the index is the FP64 argmin.
    Explicit ``topk`` / ``rescore`` (arguments, then settings) override."""
    precision = precision or getattr(s, "precision", "fp32")
    if precision not in ("fp32", "bf16", "fp8"):
        raise ValueError(f"precision must be fp32, bf16 or fp8, got {precision!r}")
    single = bool(getattr(s, "use_single_prec", False))
    if rescore is None:
        rescore = getattr(s, "rescore", None)
    if rescore is None:
        rescore = not (precision == "fp32" and single)
    K = topk or getattr(s, "topk", 0) or (8 if precision != "fp32" else (4 if rescore else 1))
    if K not in (1, 2, 4, 8):
        raise ValueError("topk must be 1, 2, 4 or 8")
    return precision, K, bool(rescore)


def slicing_and_match(ac, slices_, pixel_position_reciprocal, pixel_distance_reciprocal,
                      ref_orientations=None, *, settings=None, precision=None, topk=None,
                      rescore=None, return_details=False):
    """Quaternions (N_slices, 4) of the best-matching reference orientation per image."""
    s = settings or _settings
    st_init = time.monotonic()
    N_orientations = s.N_orientations
    N_batch_size = s.N_batch_size
    det = tuple(s.reduced_det_shape)
    N_pixels = int(np.prod(det))
    slices_ = np.asarray(slices_)
    N_slices = slices_.shape[0]
    assert slices_.shape == (N_slices,) + det
    if not N_slices:
        return np.zeros((0, 4))
    if ref_orientations is None:
        ref_orientations = quaternions.uniform_quaternions(N_orientations)
    else:
        if s.verbose:
            print(f"Warning: {ref_orientations.shape[0]} referenced orientations were given (unit test).")
        if ref_orientations.shape[0] < N_orientations:
            raise ValueError(f"{ref_orientations.shape[0]} reference orientations given, "
                             f"N_orientations = {N_orientations}")
    ref_orientations = np.asarray(ref_orientations)
    assert N_orientations % N_batch_size == 0, "N_orientations must be divisible by N_batch_size"
    precision, K, do_rescore = resolve_form(s, precision, topk, rescore)

    ac = np.asarray(ac)
    assert np.all(np.isreal(ac))
    ac = np.real(ac)
    reciprocal_extent = float(np.asarray(pixel_distance_reciprocal).max())
    ppr = np.asarray(pixel_position_reciprocal, np.float64)
    if int(np.prod(ppr.shape[1:])) != N_pixels or ppr.shape[0] != 3:
        raise ValueError(f"pixel_position_reciprocal {ppr.shape} does not match {det}")

    plan = get_plan(ac.shape[0])
    plan.set_model(ac)
    scale = np.pi / (reciprocal_extent * s.oversampling)
    st_slice = time.monotonic()
    sl = engine.SliceSet.compute(plan, ref_orientations[:N_orientations], ppr.reshape(3, -1),
                                 scale, precision, order=engine.auto_tiling(plan, ppr.shape[1:]))
    # the images go up on a copy stream while the slicing runs
    imgs = engine.ImageSet(engine.upload_rows(slices_.reshape(N_slices, N_pixels), plan.device),
                           plan.device)
    if int(sl.bad.item()) > 0:
        raise AssertionError(f"{int(sl.bad.item())} NU points outside [-3pi, 3pi)")
    en_slice = time.monotonic()
    r = engine.match_resident(sl, imgs, precision=precision, K=K, rescore_fp64=do_rescore)
    index = r["index"].long().cpu().numpy()
    en_match = time.monotonic()
    if s.verbose:
        print(f"Data/Model std ratio: {float(r['alpha'].item())}.", flush=True)
        print(f"Match tot:{en_match - st_init:.2f}s. slice={en_slice - st_slice:.2f}s. "
              f"match={en_match - en_slice:.2f}s. slice_oh={st_slice - st_init:.2f}s.")
    if return_details:
        out = {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in r.items()}
        out["index"] = index
        out["orientations"] = ref_orientations[index]
        return out
    return ref_orientations[index]


def match(ac, slices_, pixel_position_reciprocal, pixel_distance_reciprocal,
          ref_orientations=None):
    """spinifel/mpi/orientation_matching.py:8-16: each rank (here: the one GPU)
    matches its own images."""
    return slicing_and_match(ac, slices_, pixel_position_reciprocal, pixel_distance_reciprocal,
                             ref_orientations=ref_orientations)
