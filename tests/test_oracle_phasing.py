"""Pins the oracle's phasing (ER / HIO / shrink-wrap) by what each step must
do, and its Gaussian filter against scipy (importable here)."""
import numpy as np
import pytest

from helpers import rel_l2
from oracle import phasing


@pytest.mark.parametrize("mode,sigma,truncate", [("wrap", 1, 2), ("constant", 1, 2),
                                                 ("wrap", 0.7, 4.0)])
def test_gaussian_filter_matches_scipy(mode, sigma, truncate):
    from scipy.ndimage import gaussian_filter
    a = np.random.default_rng(1).standard_normal((9, 10, 11))
    np.testing.assert_allclose(phasing.gaussian_filter(a, sigma, mode, truncate),
                               gaussian_filter(a, sigma=sigma, mode=mode, truncate=truncate),
                               rtol=1e-12, atol=1e-14)


def _object(M, seed):
    """A compact positive density in the central region, in shifted (_) order."""
    rng = np.random.default_rng(seed)
    rho = np.zeros((M,) * 3)
    c, r = M // 2, M // 6
    rho[c - r:c + r + 1, c - r:c + r + 1, c - r:c + r + 1] = rng.uniform(0.5, 1.5, (2 * r + 1,) * 3)
    return np.fft.ifftshift(rho)


def test_step_phase_imposes_amplitudes_and_er_fixed_point():
    M = 17
    rho0_ = _object(M, 2)
    amp_ = np.abs(np.fft.fftn(rho0_))
    mask = np.ones((M,) * 3, dtype=bool)
    mask[0, 0, 0] = False
    sup = rho0_ > 0
    # a different density gets the given amplitudes (off the masked peak)
    other = np.random.default_rng(3).uniform(0, 1, (M,) * 3)
    rho_mod, star = phasing.step_phase(other, amp_, mask, sup)
    got = np.abs(np.fft.fftn(rho_mod))
    np.testing.assert_allclose(got[mask], amp_[mask], rtol=1e-10, atol=1e-10)
    assert np.all(star <= sup) and np.all(star == (sup & (rho_mod > 0)))
    # a consistent density is a fixed point of ER
    rho = rho0_.copy()
    phasing.ER(rho, amp_, mask, sup, np.inf)
    assert rel_l2(rho, rho0_) < 1e-12


def test_hio_closed_form_and_clamps():
    M = 9
    rng = np.random.default_rng(4)
    rho = rng.standard_normal((M,) * 3)
    amp = np.abs(np.fft.fftn(rng.uniform(0, 1, (M,) * 3)))
    mask = np.ones((M,) * 3, dtype=bool)
    mask[0, 0, 0] = False
    sup = rng.uniform(size=(M,) * 3) > 0.5
    rho_mod, star = phasing.step_phase(rho, amp, mask, sup)
    h = rho.copy()
    phasing.HIO(0.9, h, amp, mask, sup, np.inf)
    np.testing.assert_allclose(h, np.where(star, rho_mod, rho - 0.9 * rho_mod), atol=1e-14)
    cap = float(np.quantile(rho_mod, 0.8))
    e = rho.copy()
    phasing.ER(e, amp, mask, sup, cap)
    ref = np.where(star, rho_mod, 0.0)
    ref[rho_mod > cap] = cap
    np.testing.assert_allclose(e, ref, atol=1e-14)


def test_shrink_wrap_threshold_rules():
    M = 11
    rho = _object(M, 5)
    sup = np.zeros((M,) * 3, dtype=bool)
    phasing.shrink_wrap(1, rho, sup)
    g = phasing.gaussian_filter(np.abs(rho), 1, "wrap", 2)
    assert np.array_equal(sup, g > g.std())
    phasing.shrink_wrap(1, rho, sup, method="max", cutoff=0.1)
    assert np.array_equal(sup, g > 0.1 * np.abs(rho).max())
    with pytest.raises(ValueError):
        phasing.shrink_wrap(1, rho, sup, method="median")


def test_phase_recovers_the_autocorrelation_of_a_compact_object():
    """The density phased from the autocorrelation of a compact positive
    object reproduces that autocorrelation (after the reference's own
    filtering of the input) closely."""
    Mquat = 5
    M = 4 * Mquat + 1
    rho0_ = _object(M, 6)
    ac = np.fft.fftshift(np.real(np.fft.ifftn(np.abs(np.fft.fftn(rho0_)) ** 2)))
    ac_filt_, _, _ = phasing.amplitudes_from_ac(ac)
    support_ = phasing.create_support_(ac_filt_, M, Mquat)
    rho_init = support_ * np.random.default_rng(7).uniform(size=(M,) * 3)
    ac_phased, sup, rho = phasing.phase(0, ac, Mquat, rho_init, None)
    target = np.fft.fftshift(ac_filt_)
    c = np.corrcoef(ac_phased.ravel(), target.ravel())[0, 1]
    assert c > 0.95, c
    assert rho.shape == (M,) * 3 and sup.dtype == bool


def test_recenter_moves_a_known_offset_to_the_origin():
    """Known answer for recenter / center_of_mass (sequential/phasing.py:46-81):
    an object whose centre of mass sits at the signed voxel offset (5, -3, 7)
    of the fftshifted grid is rolled by (-5, 3, -7), its support with it, so
    the object ends up centred on voxel 0.  A flipped roll sign would move it
    to (10, -6, 14); a transposed axis would leave it off centre."""
    M = 21
    rho = np.zeros((M, M, M))
    sup = np.zeros((M, M, M), dtype=bool)
    ax = [np.arange(4, 7), np.arange(-4, -1) % M, np.arange(6, 9)]     # centres 5, -3, 7
    rho[np.ix_(*ax)] = np.random.default_rng(3).uniform(1.0, 1.0 + 1e-3, (3, 3, 3))
    sup[np.ix_(*ax)] = True
    rho0, sup0 = rho.copy(), sup.copy()
    ls = np.linspace(-1, 1, M + 1)
    ls = (ls[:-1] + ls[1:]) / 2
    hkl = np.stack([np.fft.ifftshift(c) for c in np.meshgrid(ls, ls, ls, indexing="ij")])
    np.testing.assert_array_equal(phasing.center_of_mass(rho, hkl, M), [5, -3, 7])
    phasing.recenter(rho, sup, M)
    want = np.roll(rho0, (-5, 3, -7), axis=(0, 1, 2))
    np.testing.assert_array_equal(rho, want)
    np.testing.assert_array_equal(sup, np.roll(sup0, (-5, 3, -7), axis=(0, 1, 2)))
    centred = [np.arange(-1, 2) % M] * 3
    assert rho[np.ix_(*centred)].min() >= 1.0
    assert rho.sum() == pytest.approx(rho[np.ix_(*centred)].sum(), rel=1e-15)
    np.testing.assert_array_equal(phasing.center_of_mass(rho, hkl, M), [0, 0, 0])


def test_center_of_mass_uses_absolute_density_and_rounds_half_to_even():
    """|rho| enters the centre of mass (phasing.py:55); a two-voxel object at
    offsets 5 and 6 has its centre at 5.5, which np.round takes to 6."""
    M = 15
    ls = np.linspace(-1, 1, M + 1)
    ls = (ls[:-1] + ls[1:]) / 2
    hkl = np.stack([np.fft.ifftshift(c) for c in np.meshgrid(ls, ls, ls, indexing="ij")])
    rho = np.zeros((M, M, M))
    rho[5, 0, 0] = -2.0
    rho[6, 0, 0] = 2.0
    np.testing.assert_array_equal(phasing.center_of_mass(rho, hkl, M), [6, 0, 0])
