"""ORACLE -- test infrastructure, not product code.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module.

FP64 CPU restatement of spinifel's phasing step (M-TIP's last step,
spinifel/sequential/phasing.py:46-314) with numpy FFTs:

* ``center_of_mass`` / ``recenter``: :46-81;
* ``create_support_``: :85-107 (central cube of edge M - 2 Mquat and
  ac_ > 1e-2 max);
* ``step_phase``: :170-194 (replace the Fourier amplitudes except the
  masked peak, inverse transform, real part, support star);
* ``ER`` / ``HIO``: :129-167 (with the rho_max clamp);
* ``shrink_wrap``: :198-224 (gaussian_filter(|rho|, sigma, mode 'wrap',
  truncate 2), threshold std * weight or max * cutoff * weight);
* ``gaussian_filter``: scipy.ndimage.gaussian_filter for the two modes the
  step uses ('constant' and 'wrap'), separable FP64;
* ``phase``: :228-314 with the initial density given (the reference draws
  ``support_ * rand`` when ``rho_`` is None); the PNG dumps of
  ``image.show_volume`` are left out.

Parity status: ``gaussian_filter`` against scipy.ndimage (importable here);
``step_phase`` by what it must do (its output's Fourier amplitudes equal the
given ones off the masked peak; a consistent density is a fixed point of
ER); ``HIO`` by its closed form on a hand-made case; ``phase`` on a compact
positive object by the agreement of the phased autocorrelation with the
filtered input (tests/test_oracle_phasing.py).
"""
import numpy as np


def _axis_index(n, radius, mode):
    i = np.arange(-radius, n + radius)
    if mode == "wrap":
        return np.mod(i, n), None
    if mode == "constant":
        valid = (i >= 0) & (i < n)
        return np.clip(i, 0, n - 1), valid
    raise ValueError(mode)


def gaussian_filter(a, sigma, mode, truncate=4.0):
    """scipy.ndimage.gaussian_filter(a, sigma, mode=mode, truncate=truncate), FP64,
    mode 'wrap' or 'constant' (cval 0)."""
    radius = int(truncate * float(sigma) + 0.5)
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    w = np.exp(-0.5 * (x / sigma) ** 2)
    w /= w.sum()
    out = np.asarray(a, dtype=np.float64)
    for ax in range(out.ndim):
        n = out.shape[ax]
        idx, valid = _axis_index(n, radius, mode)
        pad = np.take(out, idx, axis=ax)
        if valid is not None:
            shape = [1] * out.ndim
            shape[ax] = valid.size
            pad = pad * valid.reshape(shape)
        acc = np.zeros_like(out)
        for t in range(2 * radius + 1):
            acc += w[t] * np.take(pad, np.arange(t, t + n), axis=ax)
        out = acc
    return out


def center_of_mass(rho_, hkl_, M):
    """phasing.py:46-59."""
    rho_ = np.abs(rho_)
    num = (rho_ * hkl_).sum(axis=(1, 2, 3))
    den = rho_.sum()
    return np.round(num / den * M / 2)


def recenter(rho_, support_, M):
    """phasing.py:63-81 (in place)."""
    ls = np.linspace(-1, 1, M + 1)
    ls = (ls[:-1] + ls[1:]) / 2
    hkl_list = np.meshgrid(ls, ls, ls, indexing="ij")
    hkl_ = np.stack([np.fft.ifftshift(coord) for coord in hkl_list])
    vect = center_of_mass(rho_, hkl_, M)
    for i in range(3):
        shift = int(vect[i])
        rho_[:] = np.roll(rho_, -shift, i)
        support_[:] = np.roll(support_, -shift, i)


def create_support_(ac_, M, Mquat):
    """phasing.py:85-107."""
    sl = slice(Mquat, -Mquat)
    square_support = np.zeros((M, M, M), dtype=np.bool_)
    square_support[sl, sl, sl] = 1
    square_support_ = np.fft.ifftshift(square_support)
    thresh_support_ = ac_ > 1e-2 * ac_.max()
    return np.logical_and(square_support_, thresh_support_)


def step_phase(rho_, amplitudes_, amp_mask_, support_):
    """phasing.py:170-194."""
    rho_hat_ = np.fft.fftn(rho_)
    phases_ = np.angle(rho_hat_)
    rho_hat_mod_ = np.where(amp_mask_, amplitudes_ * np.exp(1j * phases_), rho_hat_)
    rho_mod_ = np.fft.ifftn(rho_hat_mod_).real
    support_star_ = np.logical_and(support_, rho_mod_ > 0)
    return rho_mod_, support_star_


def ER(rho_, amplitudes_, amp_mask_, support_, rho_max):
    """phasing.py:129-146 (in place)."""
    rho_mod_, support_star_ = step_phase(rho_, amplitudes_, amp_mask_, support_)
    rho_[:] = np.where(support_star_, rho_mod_, 0)
    i_overmax = rho_mod_ > rho_max
    rho_[i_overmax] = rho_max


def HIO(beta, rho_, amplitudes_, amp_mask_, support_, rho_max):
    """phasing.py:150-167 (in place)."""
    rho_mod_, support_star_ = step_phase(rho_, amplitudes_, amp_mask_, support_)
    rho_[:] = np.where(support_star_, rho_mod_, rho_ - beta * rho_mod_)
    i_overmax = rho_mod_ > rho_max
    rho_[i_overmax] += 2 * beta * rho_mod_[i_overmax] - rho_max


def shrink_wrap(sigma, rho_, support_, method=None, weight=1., cutoff=0.05):
    """phasing.py:198-224 (in place on support_)."""
    rho_abs_ = np.absolute(rho_)
    rho_gauss_ = gaussian_filter(rho_abs_, sigma, "wrap", truncate=2)
    if method is None:
        method = "std"
    if method == "std":
        threshold = np.std(rho_gauss_) * weight
    elif method == "max":
        threshold = rho_abs_.max() * cutoff * weight
    else:
        raise ValueError(f"Invalid method: {method}. Options are 'std' or 'max'.")
    support_[:] = rho_gauss_ > threshold


def amplitudes_from_ac(ac):
    """phasing.py:257-267: filtered ac, its shifted copy, Fourier amplitudes, mask."""
    M = ac.shape[0]
    ac_filt = gaussian_filter(np.maximum(np.real(ac), 0), 1, "constant", truncate=2)
    ac_filt_ = np.fft.ifftshift(ac_filt)
    amplitudes_ = np.sqrt(np.abs(np.fft.fftn(ac_filt_)))
    amp_mask_ = np.ones((M, M, M), dtype=np.bool_)
    amp_mask_[0, 0, 0] = 0
    return ac_filt_, amplitudes_, amp_mask_


def phase(generation, ac, Mquat, rho_, support_=None, method=None, weight=1.,
          nER=10, nHIO=5, N_phase_loops=5, beta=0.9):
    """phasing.py:228-314 -> (ac_phased, support_, rho_); rho_ is required."""
    M = 4 * Mquat + 1
    ac = np.asarray(ac)
    ac_filt_, amplitudes_, amp_mask_ = amplitudes_from_ac(ac)
    if support_ is None:
        support_ = create_support_(ac_filt_, M, Mquat)
    support_ = np.array(support_)
    rho_ = np.array(rho_, dtype=np.float64)
    rho_max = np.inf
    for _ in range(N_phase_loops):
        for _ in range(nER):
            ER(rho_, amplitudes_, amp_mask_, support_, rho_max)
        for _ in range(nHIO):
            HIO(beta, rho_, amplitudes_, amp_mask_, support_, rho_max)
        for _ in range(nER):
            ER(rho_, amplitudes_, amp_mask_, support_, rho_max)
        shrink_wrap(1, rho_, support_, method=method, weight=weight)
    for _ in range(nER):
        ER(rho_, amplitudes_, amp_mask_, support_, rho_max)
    recenter(rho_, support_, M)
    intensities_phased_ = np.abs(np.fft.fftn(rho_)) ** 2
    ac_phased_ = np.abs(np.fft.ifftn(intensities_phased_))
    ac_phased = np.fft.fftshift(ac_phased_)
    return ac_phased.astype(np.float64), support_, rho_
