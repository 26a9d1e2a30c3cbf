"""GPU parity of the phasing step (ER / HIO / shrink-wrap, libmtip kernels +
rocFFT) against the oracle's FP64 restatement of
spinifel/sequential/phasing.py: single iterations to 1e-12, a short run with
the same initial density to 1e-8 (same supports), and the full default run
on a compact object by the quality of the phased autocorrelation."""
from types import SimpleNamespace

import numpy as np
import pytest

from helpers import rel_l2
from oracle import phasing as oph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    from spinifel_amd import engine
    engine.require_gpu()
    return torch.device("cuda", 0)


def _object_ac(Mquat, seed):
    M = 4 * Mquat + 1
    rng = np.random.default_rng(seed)
    rho = np.zeros((M,) * 3)
    c, r = M // 2, M // 6
    rho[c - r:c + r + 1, c - r:c + r + 1, c - r:c + r + 1] = rng.uniform(0.5, 1.5, (2 * r + 1,) * 3)
    rho_ = np.fft.ifftshift(rho)
    ac = np.fft.fftshift(np.real(np.fft.ifftn(np.abs(np.fft.fftn(rho_)) ** 2)))
    return M, ac


def test_single_er_and_hio_vs_oracle(torch_dev):
    import torch
    from spinifel_amd import phasing
    M = 17
    rng = np.random.default_rng(1)
    ac = rng.uniform(0, 1, (M,) * 3)
    ac = ac + ac[::-1, ::-1, ::-1]
    ac_filt_, amp_, mask = oph.amplitudes_from_ac(ac)
    sup = rng.uniform(size=(M,) * 3) > 0.4
    rho0 = rng.standard_normal((M,) * 3)
    _, amp_half = phasing._amplitudes_from_ac(ac, torch_dev)
    np.testing.assert_allclose(amp_half.cpu().numpy(), amp_[:, :, :M // 2 + 1], rtol=1e-12,
                               atol=1e-12)
    for mode in (0, 1):
        ph = phasing._Phaser(amp_half, torch.as_tensor(sup, device=torch_dev),
                             torch.as_tensor(rho0, device=torch_dev), M)
        ph.step(mode, 0.9)
        ref = rho0.copy()
        if mode == 0:
            oph.ER(ref, amp_, mask, sup, np.inf)
        else:
            oph.HIO(0.9, ref, amp_, mask, sup, np.inf)
        assert rel_l2(ph.rho.cpu().numpy(), ref) < 1e-12
    # the rho_max clamp of ER
    ph = phasing._Phaser(amp_half, torch.as_tensor(sup, device=torch_dev),
                         torch.as_tensor(rho0, device=torch_dev), M)
    ph.step(0, 0.0, 0.05)
    ref = rho0.copy()
    oph.ER(ref, amp_, mask, sup, 0.05)
    np.testing.assert_allclose(ph.rho.cpu().numpy(), ref, rtol=0, atol=1e-12)


def test_short_phase_vs_oracle(torch_dev):
    from spinifel_amd import phasing
    print("case: phase Mquat=4, 1 loop of 3 ER / 2 HIO / 3 ER", flush=True)
    Mquat = 4
    M, ac = _object_ac(Mquat, 2)
    s = SimpleNamespace(Mquat=Mquat, nER=3, nHIO=2, N_phase_loops=1, beta=0.9, device=0)
    ac_filt_, _, _ = oph.amplitudes_from_ac(ac)
    sup0 = oph.create_support_(ac_filt_, M, Mquat)
    rho0 = sup0 * np.random.default_rng(3).uniform(size=(M,) * 3)
    got = phasing.phase(0, ac, None, rho0, settings=s)
    ref = oph.phase(0, ac, Mquat, rho0, None, nER=3, nHIO=2, N_phase_loops=1, beta=0.9)
    assert np.array_equal(got[1], ref[1])
    assert rel_l2(got[2], ref[2]) < 1e-8
    assert rel_l2(got[0], ref[0]) < 1e-8


def test_default_phase_recovers_autocorrelation(torch_dev):
    """Default settings (5 loops of 10 ER / 5 HIO / 10 ER + shrink-wrap, then
    10 ER) at M = 81: the phased autocorrelation of a compact object matches
    the filtered input, as the oracle's does."""
    from spinifel_amd import phasing
    from spinifel_amd.settings import SpinifelSettings
    print("case: phase Mquat=20 (M=81) default settings", flush=True)
    Mquat = 20
    M, ac = _object_ac(Mquat, 4)
    ac_filt_, _, _ = oph.amplitudes_from_ac(ac)
    target = np.fft.fftshift(ac_filt_)
    s = SpinifelSettings()
    assert (s.Mquat, s.nER, s.nHIO, s.N_phase_loops, s.beta) == (20, 10, 5, 5, 0.9)
    np.random.seed(5)
    ac_phased, sup, rho = phasing.phase(0, ac, settings=s)
    assert ac_phased.shape == (M,) * 3 and sup.dtype == bool and rho.dtype == np.float64
    c = np.corrcoef(ac_phased.ravel(), target.ravel())[0, 1]
    assert c > 0.95, c


def test_recenter_known_offset(torch_dev):
    """The GPU recenter rolls an object centred at the signed voxel offset
    (5, -3, 7) by (-5, 3, -7), density and support alike, exactly as the
    oracle (pinned by its own known-answer test) does."""
    import torch
    from spinifel_amd import phasing
    M = 21
    rho = np.zeros((M,) * 3)
    sup = np.zeros((M,) * 3, dtype=bool)
    ax = [np.arange(4, 7), np.arange(-4, -1) % M, np.arange(6, 9)]
    rho[np.ix_(*ax)] = np.random.default_rng(3).uniform(1.0, 1.5, (3, 3, 3))
    sup[np.ix_(*ax)] = True
    r_gpu, s_gpu = phasing.recenter(torch.as_tensor(rho, device=torch_dev),
                                    torch.as_tensor(sup, device=torch_dev), M)
    r_ora, s_ora = rho.copy(), sup.copy()
    oph.recenter(r_ora, s_ora, M)
    np.testing.assert_array_equal(r_gpu.cpu().numpy(), r_ora)
    np.testing.assert_array_equal(s_gpu.cpu().numpy(), s_ora)
    np.testing.assert_array_equal(r_ora, np.roll(rho, (-5, 3, -7), axis=(0, 1, 2)))
