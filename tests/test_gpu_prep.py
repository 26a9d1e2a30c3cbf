"""GPU parity of the data preparation (clip + bin in one pass, libmtip)
against the oracle's restatement of spinifel/prep.py: bit-exact for every
mode, dtype and detector layout the reference handles."""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import prep as oprep

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    from spinifel_amd import engine
    engine.require_gpu()
    return torch.device("cuda", 0)


@pytest.mark.parametrize("shape", [(3, 1, 64, 96), (2, 4, 32, 32), (1, 2, 16, 16)])
@pytest.mark.parametrize("n_clip,n_bin", [(0, 0), (0, 2), (1, 1), (1, 3), (2, 0)])
def test_sum_and_mean_bitwise(torch_dev, shape, n_clip, n_bin):
    from spinifel_amd import prep
    rng = np.random.default_rng(n_clip * 10 + n_bin)
    imgs = rng.poisson(2.0, shape).astype(np.float32)
    ppr = rng.standard_normal(shape)
    np.testing.assert_array_equal(prep.clip_bin(imgs, n_clip, n_bin, "sum"),
                                  oprep.binning(imgs, n_clip, n_bin, "sum"))
    np.testing.assert_array_equal(prep.clip_bin(ppr, n_clip, n_bin, "mean"),
                                  oprep.binning(ppr, n_clip, n_bin, "mean"))


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_index_bitwise(torch_dev, dtype):
    from spinifel_amd import prep
    r = np.arange(64)
    idx = np.stack(np.meshgrid(r, r, indexing="ij"))[:, None].astype(dtype) - 7   # (2,1,64,64)
    s = SimpleNamespace(N_clipping=1, N_binning=2)
    got = prep.binning_index(idx, settings=s)
    ref = oprep.binning(idx, 1, 2, "index")
    assert got.dtype == ref.dtype
    np.testing.assert_array_equal(got, ref)


def test_device_tensors_and_settings(torch_dev):
    import torch
    from spinifel_amd import prep
    s = SimpleNamespace(N_clipping=1, N_binning=1)
    imgs = torch.rand((5, 1, 128, 128), device=torch_dev)
    out = prep.binning_sum(imgs, settings=s)
    assert out.is_cuda and tuple(out.shape) == (5, 1, 32, 32)
    np.testing.assert_array_equal(out.cpu().numpy(), oprep.binning(imgs.cpu().numpy(), 1, 1, "sum"))
    with pytest.raises(ValueError):
        prep.clip_bin(np.zeros((1, 1, 12, 12), np.float32), 0, 3, "sum")     # 12 not a multiple of 8
    with pytest.raises(ValueError):
        prep.clip_bin(np.zeros((1, 1, 8, 8), np.int32), 0, 1, "mean")
