"""The stateful class path the reconstruction drivers call:
``NUFFT`` (extern/nufft_ext.py:36-284) and ``SNM``
(sequential/orientation_matching.py:40-186), as in spinifel/mpi/main.py:151-160,
:250 and spinifel/sequential/main.py:36-61:

    nufft = NUFFT(settings, pixel_position_reciprocal, pixel_distance_reciprocal)
    snm = SNM(settings, slices_, pixel_position_reciprocal, pixel_distance_reciprocal, nufft)
    orientations = snm.slicing_and_match(ac_phased)

The data images stay resident in HBM between calls (the class path's point),
the NUFFT plan persists, and every step of ``slicing_and_match`` runs on the
GPU: slicing of all orientations, the class path's per-batch scaling of the
model slices (each batch of N_batch_size orientations scaled by the images'
std over the batch's own std, orientation_matching.py:165), distances and
argmin.  Differences a caller can observe:
  * ``ref_orientations`` defaults to this package's super-Fibonacci grid, not
    skopi's ``get_uniform_quat`` (skopi is absent); pass ``ref_orientations``
    for a given grid;
  * ``NUFFT.forward`` returns the real part of the transform as a float32
    device tensor (the class path only uses ``.real``,
    orientation_matching.py:163);
  * the transform tolerance is ``settings.nufft_eps`` (1e-6 by default) where
    the reference fixes 1e-12 (nufft_ext.py:62).
"""
import numpy as np
import torch

from . import engine
from .autocorrelation import get_plan
from .quaternions import uniform_quaternions


class NUFFT:
    """Plan, rotated pixel geometry and reference grid of the class path
    (extern/nufft_ext.py:36-144)."""

    def __init__(self, settings, pixel_position_reciprocal, pixel_distance_reciprocal,
                 images_per_rank=None, ref_orientations=None):
        engine.require_gpu()
        self.N_orientations = int(settings.N_orientations)
        self.N_batch_size = int(settings.N_batch_size)
        self.reduced_det_shape = tuple(settings.reduced_det_shape)
        self.oversampling = settings.oversampling
        self.N_pixels = int(np.prod(self.reduced_det_shape))
        self.N_images = getattr(settings, "N_images_per_rank", 0) if images_per_rank is None \
            else int(images_per_rank)
        self.eps = float(getattr(settings, "nufft_eps", 1e-6))
        self.isign = -1
        self.device = torch.device("cuda", int(getattr(settings, "device", 0)))
        q = np.asarray(ref_orientations, np.float64) if ref_orientations is not None else \
            uniform_quaternions(self.N_orientations)
        self.ref_orientations = q.reshape(-1, 4)
        self.reciprocal_extent = float(np.asarray(pixel_distance_reciprocal).max())
        self.pixel_position_reciprocal = np.asarray(pixel_position_reciprocal, np.float64)
        if int(np.prod(self.pixel_position_reciprocal.shape[1:])) != self.N_pixels:
            raise ValueError("pixel_position_reciprocal does not match reduced_det_shape")
        self.mult = np.pi / (self.oversampling * self.reciprocal_extent)
        self._q = torch.as_tensor(np.ascontiguousarray(
            self.pixel_position_reciprocal.reshape(3, -1))).to(self.device)
        self._quats = torch.as_tensor(np.ascontiguousarray(self.ref_orientations)).to(self.device)
        self._model_key = None

    def update_fields(self, n_images_per_rank):
        """extern/nufft_ext.py:147-182: the image count changed (buffers are
        per call here, so only the count is kept)."""
        self.N_images = int(n_images_per_rank)

    def plan(self, M):
        return get_plan(M, self.eps, self.device.index or 0)

    def set_model(self, ac):
        """Load ac (real, M^3) into the plan's fine grid unless it is already there."""
        ac = np.asarray(ac) if not torch.is_tensor(ac) else ac
        if torch.is_tensor(ac):
            key = (ac.data_ptr(), tuple(ac.shape), ac._version)
        else:
            key = None     # host arrays: always reload
        plan = self.plan(ac.shape[0])
        if key is None or key != self._model_key:
            plan.set_model(ac if torch.is_tensor(ac) else np.real(ac).astype(np.float32))
            self._model_key = key
        return plan

    def slices(self, ac, st=0, en=None, precision="fp32"):
        """engine.SliceSet of orientations [st, en) of the reference grid."""
        en = self.N_orientations if en is None else en
        plan = self.set_model(ac)
        sl = engine.SliceSet.compute(plan, self._quats[st:en], self._q, self.mult, precision,
                                     order=engine.auto_tiling(plan, self.pixel_position_reciprocal.shape[1:]))
        if int(sl.bad.item()) > 0:
            # the assert of nufft_ext.py:144, checked where the points are formed
            raise AssertionError("rotated pixel positions outside [-3pi, 3pi)")
        return sl

    def forward(self, ugrid, st, en, support, use_recip_sym, N):
        """extern/nufft_ext.py:248-284: the transform at orientations [st, en),
        flat (N,) orientation-major; real part, float32 on the device."""
        if use_recip_sym:
            g = ugrid if not torch.is_tensor(ugrid) else ugrid.cpu().numpy()
            assert np.all(np.real(g) == g), "ugrid must be real"
        if support is not None and not (np.isscalar(support) and support == 1):
            ugrid = ugrid * support
        sl = self.slices(ugrid, st, en)
        out = sl.f32[:, :sl.N_p].reshape(-1)
        if out.numel() != N:
            raise ValueError(f"N = {N} but {en - st} orientations x {self.N_pixels} pixels")
        return out

    def adjoint(self, nuvect, H_, K_, L_, support, use_reciprocal_symmetry, M):
        """extern/nufft_ext.py:287-327: type 1 (isign +1) of real ``nuvect`` at
        (H_, K_, L_) onto M^3, times support, real part, / M^3 -> float64
        (M, M, M) device tensor.  Complex data or use_reciprocal_symmetry=False
        raise NotImplementedError (the GPU path returns the real part)."""
        if not use_reciprocal_symmetry:
            raise NotImplementedError("use_reciprocal_symmetry=False (GPU path returns the real part)")
        if torch.is_tensor(nuvect):
            if nuvect.is_complex():
                if bool((nuvect.imag != 0).any()):
                    raise NotImplementedError("adjoint of complex data")
                nuvect = nuvect.real
        else:
            nuvect = np.asarray(nuvect)
            if np.iscomplexobj(nuvect):
                if np.any(nuvect.imag != 0):
                    raise NotImplementedError("adjoint of complex data")
                nuvect = nuvect.real
        sup = None
        if support is not None and not (np.isscalar(support) and support == 1):
            sup = support if torch.is_tensor(support) else np.ascontiguousarray(
                np.broadcast_to(np.asarray(support, np.float64), (int(M),) * 3))
        return self.plan(int(M)).adjoint_points(H_, K_, L_, values=torch.as_tensor(nuvect).reshape(-1),
                                                support=sup)


class SNM:
    """Slicing and matching with the data images resident
    (sequential/orientation_matching.py:40-186)."""

    def __init__(self, settings, slices_, pixel_position_reciprocal, pixel_distance_reciprocal,
                 nufft, precision=None):
        self.N_orientations = int(settings.N_orientations)
        self.N_batch_size = int(settings.N_batch_size)
        self.reduced_det_shape = tuple(settings.reduced_det_shape)
        self.oversampling = settings.oversampling
        self.N_pixels = int(np.prod(self.reduced_det_shape))
        self.N = self.N_pixels * self.N_orientations
        assert self.N_orientations % self.N_batch_size == 0, \
            "N_orientations must be divisible by N_batch_size"
        slices_ = np.asarray(slices_) if not torch.is_tensor(slices_) else slices_
        self.N_slices = slices_.shape[0]
        assert tuple(slices_.shape) == (self.N_slices,) + self.reduced_det_shape
        self.reciprocal_extent = float(np.asarray(pixel_distance_reciprocal).max())
        self.pixel_position_reciprocal = pixel_position_reciprocal
        self.nufft = nufft
        self.precision = precision or getattr(settings, "precision", "fp32")
        # FP64 rescoring unless use_single_prec (the reference's f_type rule)
        from .orientation_matching import resolve_form
        _, self.K, self.rescore = resolve_form(settings, self.precision)
        self.images = None
        if self.N_slices:
            self.images = engine.ImageSet(slices_.reshape(self.N_slices, self.N_pixels),
                                          nufft.device)
            n = self.N_slices * self.N_pixels
            mean = self.images.row_sum.sum() / n
            # slices_.std() (:74), population, FP64 on the device
            self.slices_std = torch.sqrt(self.images.row_sumsq.sum() / n - mean * mean)
        self.last = None

    def batch_multipliers(self, sl):
        """Per orientation the class path's scale slices_std / std(its batch) (:165)."""
        nb = self.N_orientations // self.N_batch_size
        offs = np.arange(nb + 1, dtype=np.int64) * self.N_batch_size
        # segments = orientation batches against the images: std(batch) / std(images)
        return engine.segment_scaling_ratio(sl.row_sum, sl.row_sumsq, offs, self.images.row_sum,
                                            self.images.row_sumsq, self.N_pixels)

    def slicing_and_match(self, ac):
        """Orientations (N_slices, 4) of the data images (:122-186)."""
        if not self.N_slices:
            return np.zeros((0, 4))
        sl = self.nufft.slices(ac, 0, self.N_orientations)
        ratio = self.batch_multipliers(sl)
        # model slices scaled in place by the class path's per-batch factor
        scaled = engine.residual(sl.f32, "fp32", sl.N_p, mult=ratio, invert=True)
        model = engine.SliceSet.from_rows(scaled.data, sl.N_p)
        one = engine._one(self.nufft.device)
        r = engine.match_resident(model, self.images, self.precision, K=self.K,
                                  rescore_fp64=self.rescore, alpha=one)
        index = r["index"].long().cpu().numpy()
        self.last = dict(index=index, mult=(1.0 / ratio).cpu().numpy(), result=r)
        return self.nufft.ref_orientations[index]
