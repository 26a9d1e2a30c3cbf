"""Phasing (M-TIP's last step) on the GPU, drop-in for ``phase`` of
spinifel/sequential/phasing.py:228-314:

    ac_phased, support_, rho_ = phase(generation, ac, support_, rho_)

Cycles of error reduction (ER), hybrid input-output (HIO) and shrink-wrap
recover a density whose Fourier amplitudes match the square root of the
autocorrelation's transform.  Every iteration is one FP64 real-to-complex
FFT pair on the M^3 grid (rocFFT through torch.fft) around two libmtip
kernels: the amplitude replacement on the half spectrum
(``mtip_phase_amplitudes``) and the fused ER / HIO update
(``mtip_phase_update``).  The density is real and its spectrum Hermitian, so
the half spectrum carries all of it and the inverse is the real part the
reference keeps.  The array conventions are the reference's: trailing
underscores are ifftshifted arrays, FFTs of them unshifted (:27-41).

Differences a caller can observe:
  * ``rho_=None`` draws ``support_ * numpy.random.rand`` like the CPU path
    (the GPU path of the reference draws from cupy's generator);
  * the PNG dumps (``image.show_volume``, :298-305) are not written;
  * results return as numpy arrays, as the reference's path returns them
    after its conversion (:290-296).
"""
import numpy as np
import torch

from . import _lib, engine, filters
from .settings import settings as _settings


def _amplitudes_from_ac(ac, dev):
    """:257-270 -- filtered ac (shifted), Fourier amplitudes (half spectrum)."""
    M = ac.shape[0]
    a = torch.as_tensor(np.real(np.asarray(ac)) if not torch.is_tensor(ac) else ac.real)
    a = a.to(dev, torch.float64)
    ac_filt = filters.gaussian_filter(torch.clamp(a, min=0.0), 1, "constant", truncate=2)
    ac_filt_ = torch.fft.ifftshift(ac_filt)
    amplitudes_ = torch.sqrt(torch.abs(torch.fft.rfftn(ac_filt_))).contiguous()
    return ac_filt_, amplitudes_


def create_support_(ac_, M, Mquat):
    """:85-107 -- central cube of edge M - 2 Mquat and ac_ > 1e-2 max (device)."""
    sq = torch.zeros((M, M, M), dtype=torch.bool, device=ac_.device)
    sq[Mquat:M - Mquat, Mquat:M - Mquat, Mquat:M - Mquat] = True
    return torch.fft.ifftshift(sq) & (ac_ > 1e-2 * ac_.max())


def _dense(t):
    """True when t's elements fill its storage span exactly (a permuted
    C-order layout), so memory order is a bijection onto its elements."""
    if t.storage_offset() != 0:
        return False
    expect = 1
    for st, sz in sorted(zip(t.stride(), t.size())):
        if sz != 1 and st != expect:
            return False
        expect *= sz
    return True


class _Phaser:
    """The iteration state: density, support, amplitudes, on one stream."""

    def __init__(self, amplitudes_, support_, rho_, M):
        self.lib = _lib.load()
        self.M = M
        self.amp = amplitudes_
        self.sup = support_.to(torch.uint8).contiguous()
        self.rho = rho_.contiguous()
        self.n = M ** 3

    def step(self, mode, beta=0.0, rho_max=float("inf")):
        """One ER (mode 0) or HIO (mode 1) iteration (:129-194)."""
        X = torch.fft.rfftn(self.rho)
        # rocFFT's half spectrum comes back in a permuted dense layout (strides
        # (M, 1, M^2) for (M, M, M/2+1)); the kernel walks memory linearly, so
        # the amplitudes are kept in the spectrum's own layout
        if X.stride() != self.amp.stride():
            if not _dense(X):
                X = X.contiguous()
            amp = torch.empty_strided(X.size(), X.stride(), dtype=torch.float64, device=X.device)
            amp.copy_(self.amp)
            self.amp = amp
        _lib.check(self.lib.mtip_phase_amplitudes(engine._ptr(X), engine._ptr(self.amp), X.numel(),
                                                  engine._stream()), "mtip_phase_amplitudes")
        rho_mod = torch.fft.irfftn(X, s=(self.M,) * 3).contiguous()
        _lib.check(self.lib.mtip_phase_update(engine._ptr(self.rho), engine._ptr(rho_mod),
                                              engine._ptr(self.sup), self.n, int(mode),
                                              float(beta), float(rho_max), engine._stream()),
                   "mtip_phase_update")

    def shrink_wrap(self, sigma, method=None, weight=1., cutoff=0.05):
        """:198-224 -- support = gaussian(|rho|, sigma, wrap, truncate 2) > threshold."""
        rho_abs_ = torch.abs(self.rho)
        g = filters.gaussian_filter(rho_abs_, sigma, "wrap", truncate=2)
        if method is None:
            method = "std"
        if method == "std":
            threshold = torch.std(g, unbiased=False) * weight
        elif method == "max":
            threshold = rho_abs_.max() * cutoff * weight
        else:
            raise ValueError(f"Invalid method: {method}. Options are 'std' or 'max'.")
        self.sup = (g > threshold).to(torch.uint8)


def recenter(rho_, support_, M):
    """:63-81 -- roll density and support so the centre of mass of |rho| is at
    the origin (device tensors, returned rolled)."""
    ls = torch.linspace(-1, 1, M + 1, dtype=torch.float64, device=rho_.device)
    ls = (ls[:-1] + ls[1:]) / 2
    grids = torch.meshgrid(ls, ls, ls, indexing="ij")
    a = torch.abs(rho_)
    den = a.sum()
    for i in range(3):
        num = (a * torch.fft.ifftshift(grids[i])).sum()
        # numpy's round (half to even), as torch.round
        shift = int(torch.round(num / den * M / 2))
        rho_ = torch.roll(rho_, -shift, i)
        support_ = torch.roll(support_, -shift, i)
    return rho_, support_


def phase(generation, ac, support_=None, rho_=None, method=None, weight=1., settings=None):
    """:228-314 -> (ac_phased float64 (M, M, M), support_ bool, rho_ float64),
    numpy arrays; support_ and rho_ in the shifted (_) convention."""
    engine.require_gpu()
    s = _settings if settings is None else settings
    Mquat = int(s.Mquat)
    M = 4 * Mquat + 1
    dev = torch.device("cuda", int(getattr(s, "device", 0)))
    ac_filt_, amplitudes_ = _amplitudes_from_ac(ac, dev)
    if support_ is None:
        support_ = create_support_(ac_filt_, M, Mquat)
    support_ = torch.as_tensor(np.asarray(support_) if not torch.is_tensor(support_) else support_)
    support_ = support_.to(dev).bool()
    if rho_ is None:
        rho_ = support_.cpu().numpy() * np.random.rand(*support_.shape)
    rho_ = torch.as_tensor(np.asarray(rho_) if not torch.is_tensor(rho_) else rho_)
    rho_ = rho_.to(dev, torch.float64)
    ph = _Phaser(amplitudes_, support_, rho_, M)
    nER, nHIO, beta = int(s.nER), int(s.nHIO), float(s.beta)
    for _ in range(int(s.N_phase_loops)):
        for _ in range(nER):
            ph.step(0)
        for _ in range(nHIO):
            ph.step(1, beta)
        for _ in range(nER):
            ph.step(0)
        ph.shrink_wrap(1, method=method, weight=weight)
    for _ in range(nER):
        ph.step(0)
    rho_, sup = recenter(ph.rho, ph.sup.bool(), M)
    intensities_phased_ = torch.abs(torch.fft.fftn(rho_)) ** 2
    ac_phased_ = torch.abs(torch.fft.ifftn(intensities_phased_))
    ac_phased = torch.fft.fftshift(ac_phased_)
    return (ac_phased.cpu().numpy().astype(np.float64), sup.cpu().numpy(),
            rho_.cpu().numpy())
