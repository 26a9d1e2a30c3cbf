"""The C-ABI library loads, exports every declared symbol and rejects bad
arguments before touching a GPU; the product path refuses to run without
one (no CPU fallback)."""
import ctypes

import numpy as np
import pytest

from spinifel_amd import _lib


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    names = _lib.header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared entry point has a ctypes signature in the binding
    assert set(names) == set(_lib._SIGS), set(names) ^ set(_lib._SIGS)


def test_abi_version_and_workspace_size():
    lib = _lib.load()
    assert lib.mtip_abi_version() == 5
    assert lib.mtip_match_workspace_size(10000, 10000, 8, 1) == 10000 * 79 * 8 * 8
    assert lib.mtip_match_workspace_size(1, 1, 1, 0) == 8
    assert lib.mtip_match_workspace_size(0, 5, 1, 0) == 0


def test_argument_errors_are_reported_without_a_gpu():
    lib = _lib.load()
    nul = ctypes.c_void_p()
    fake = ctypes.c_void_p(0x1000)    # never dereferenced: argument checks fail first
    rc = lib.mtip_match_topk(7, fake, 16, nul, fake, 4, fake, 16, nul, 4, 16, fake, 1, fake, 0,
                             fake, fake, nul)
    assert rc == _lib.MTIP_ERR_ARG
    assert "precision" in _lib.last_error()
    rc = lib.mtip_match_topk(0, fake, 16, nul, fake, 4, fake, 16, nul, 4, 16, fake, 3, fake, 0,
                             fake, fake, nul)
    assert rc == _lib.MTIP_ERR_ARG and "K must be" in _lib.last_error()
    rc = lib.mtip_match_topk(0, fake, 32, nul, fake, 4, fake, 32, nul, 4, 20, fake, 1, fake, 0,
                             fake, fake, nul)
    assert rc == _lib.MTIP_ERR_ARG and "multiple of 16" in _lib.last_error()
    rc = lib.mtip_match_topk(0, fake, 16, nul, fake, 4, fake, 16, nul, 4, 16, fake, 1, fake, 0,
                             fake, fake, nul)
    assert rc == _lib.MTIP_ERR_NOMEM and "workspace" in _lib.last_error()
    rc = lib.mtip_match_topk(2, fake, 16, nul, fake, 4, fake, 16, nul, 4, 16, fake, 1, fake, 4096,
                             fake, fake, nul)
    assert rc == _lib.MTIP_ERR_ARG and "FP8" in _lib.last_error()
    # split edge tiles: unknown flags, and the partial sums' bytes beside the
    # keys (10,000 x 10,000, K = 8, 256-row tile: 79 thin tiles in 3 parts of
    # 256 x 16 floats each) are part of the workspace the call needs
    keys = 10000 * 40 * 8 * 8
    rc = lib.mtip_match_partial_ex(1, fake, 16384, nul, fake, 10000, fake, 16384, nul, 10000,
                                   16384, fake, 8, 256, 2, fake, keys, nul)
    assert rc == _lib.MTIP_ERR_ARG and "flags" in _lib.last_error()
    rc = lib.mtip_match_partial_ex(1, fake, 16384, nul, fake, 10000, fake, 16384, nul, 10000,
                                   16384, fake, 8, 256, 1, fake, keys + 79 * 3 * 4096 * 4 - 1, nul)
    assert rc == _lib.MTIP_ERR_NOMEM and "workspace" in _lib.last_error()
    rc = lib.mtip_heap_sort(nul, 4, 4, 9, nul, nul)
    assert rc == _lib.MTIP_ERR_ARG
    rc = lib.mtip_euclidean_distance(fake, fake, 4, 4, 18, fake, nul)
    assert rc == _lib.MTIP_ERR_ARG and "multiple of 4" in _lib.last_error()
    rc = lib.mtip_plan_create(1, 1e-6, 0, ctypes.byref(ctypes.c_void_p()))
    assert rc == _lib.MTIP_ERR_ARG and "M must be" in _lib.last_error()
    rc = lib.mtip_residual_rows(nul, 4, 8, 4, nul, 0, nul, 0, nul, 8, nul, nul, nul)
    assert rc == _lib.MTIP_ERR_ARG
    with pytest.raises(ValueError):
        _lib.check(_lib.MTIP_ERR_ARG, "x")
    with pytest.raises(AssertionError):
        _lib.check(_lib.MTIP_ERR_DOMAIN, "x")
    with pytest.raises(MemoryError):
        _lib.check(_lib.MTIP_ERR_NOMEM, "x")


def test_long_rows_need_the_128_row_tile():
    # the 256-row GEMM tile addresses its rows with 32-bit offsets: 256 rows of
    # 2^22 bf16 values are 2^31 bytes, reported before any GPU work
    lib = _lib.load()
    fake = ctypes.c_void_p(0x1000)    # never dereferenced
    nul = ctypes.c_void_p()
    n, ld = 4096, 1 << 22
    ws = lib.mtip_match_workspace_size(n, n, 8, 1)
    # the built-in rule: the 256-row tile from 192 tiles of 256 x 256 (BF16 / FP8)
    assert lib.mtip_gemm_tile(1, n, n) == 256
    assert lib.mtip_gemm_tile(1, 500, 58642) == 256
    assert lib.mtip_gemm_tile(1, 1024, 1024) == 128
    assert lib.mtip_gemm_tile(0, n, n) == 128
    rc = lib.mtip_match_partial(1, fake, ld, nul, fake, n, fake, ld, nul, n, 16, fake, 8, 0, fake,
                                ws, nul)
    assert rc == _lib.MTIP_ERR_UNSUPPORTED and "128-row tile" in _lib.last_error()


def test_product_path_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from spinifel_amd import engine, orientation_matching, settings as st
    with pytest.raises(RuntimeError, match="GPU"):
        engine.NufftPlan(21, 1e-6, 0)
    s = st.SpinifelSettings(N_orientations=4, N_batch_size=4, reduced_det_shape=(1, 4, 4))
    with pytest.raises(RuntimeError, match="GPU"):
        orientation_matching.slicing_and_match(
            np.ones((5, 5, 5), np.float32), np.ones((2, 1, 4, 4), np.float32),
            np.ones((3, 1, 4, 4)), np.ones((1, 4, 4)), np.eye(4), settings=s)


def test_argument_errors_of_adjoint_phasing_and_prep_entry_points():
    """The type-1, phasing and prep entry points reject bad arguments before
    any device work."""
    lib = _lib.load()
    nul = ctypes.c_void_p()
    fake = ctypes.c_void_p(0x1000)
    assert lib.mtip_plan_spread_begin(nul, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_plan_spread_slices(nul, fake, 1, fake, 1, nul, 0, 0, 1.0, nul, 0, nul,
                                       nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_plan_spread_points(nul, fake, fake, fake, 4, nul, nul, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_plan_spread_finish(nul, nul, 1.0, fake, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_phase_update(fake, fake, fake, 8, 2, 0.9, 1.0, nul) == _lib.MTIP_ERR_ARG
    assert "mode" in _lib.last_error()
    assert lib.mtip_phase_update(fake, fake, fake, -1, 0, 0.9, 1.0, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_phase_amplitudes(nul, fake, 8, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_phase_amplitudes(nul, nul, 0, nul) == _lib.MTIP_OK     # nothing to do
    # clip + bin: dtype, mode, mean of integers, edge multiples, null pointers
    assert lib.mtip_clip_bin(fake, 7, 1, 1, 8, 8, 0, 1, 0, fake, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_clip_bin(fake, 0, 1, 1, 8, 8, 0, 1, 3, fake, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_clip_bin(fake, 2, 1, 1, 8, 8, 0, 1, 1, fake, nul) == _lib.MTIP_ERR_ARG
    assert "float" in _lib.last_error()
    assert lib.mtip_clip_bin(fake, 0, 1, 1, 12, 12, 0, 3, 0, fake, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_clip_bin(fake, 0, 1, 1, 8, 8, 0, 5, 0, fake, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_clip_bin(nul, 0, 1, 1, 8, 8, 0, 1, 0, fake, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_clip_bin(nul, 0, 0, 1, 8, 8, 0, 1, 0, nul, nul) == _lib.MTIP_OK


def test_prep_loaders_name_h5py():
    """The HDF5 loaders keep the reference's logic and say what is missing."""
    pytest.importorskip("torch")
    try:
        import h5py  # noqa: F401
        pytest.skip("h5py is installed")
    except ImportError:
        pass
    from spinifel_amd import prep
    with pytest.raises(ImportError, match="h5py"):
        prep.load_slices(np.zeros((1, 1, 4, 4)), 0, 1, "/nonexistent.h5")


def test_argument_errors_of_merge_entry_points():
    lib = _lib.load()
    nul = ctypes.c_void_p()
    fake = ctypes.c_void_p(0x1000)
    assert lib.mtip_merge_pad(fake, nul, 9, 8, fake, fake, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_merge_pad(nul, nul, 9, 18, fake, fake, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_merge_combine(fake, fake, fake, nul, 0, 18, 0.0, 0.0, fake, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_dot_workspace_size() >= 8
    assert lib.mtip_dot(fake, fake, 10, fake, 8, fake, nul) == _lib.MTIP_ERR_NOMEM
    assert lib.mtip_dot(fake, fake, -1, fake, 1 << 20, fake, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_cg_update(nul, fake, fake, fake, fake, fake, 4, nul) == _lib.MTIP_ERR_ARG
    assert lib.mtip_cg_update(nul, nul, nul, nul, nul, nul, 0, nul) == _lib.MTIP_OK
    assert lib.mtip_cg_direction(fake, fake, fake, nul, 4, 0, nul) == _lib.MTIP_ERR_ARG


def test_tile_rows_argument_is_checked():
    # mtip_match_partial / mtip_topk_merge take the tile height explicitly so
    # that both halves of one match use the same one
    lib = _lib.load()
    nul = ctypes.c_void_p()
    fake = ctypes.c_void_p(0x1000)    # never dereferenced
    rc = lib.mtip_topk_merge(1, fake, 1000, 10, 8, 192, fake, fake, nul)
    assert rc == _lib.MTIP_ERR_ARG and "tile_rows" in _lib.last_error()
    rc = lib.mtip_match_partial(1, fake, 64, nul, fake, 1000, fake, 64, nul, 10, 64, fake, 8, 64,
                                fake, 1 << 20, nul)
    assert rc == _lib.MTIP_ERR_ARG and "tile_rows" in _lib.last_error()
    # an explicit 128-row tile needs the larger partial buffer
    need128 = 10 * 8 * 8 * 8
    rc = lib.mtip_match_partial(1, fake, 64, nul, fake, 1000, fake, 64, nul, 10, 64, fake, 8, 128,
                                fake, need128 - 8, nul)
    assert rc == _lib.MTIP_ERR_NOMEM and "workspace" in _lib.last_error()
