"""The allocation-free top-K entry point (mtip_match_topk_static) against
mtip_match_topk, through the C ABI: two calls in a row, the second on
smaller inputs, in every precision; indices and values must be identical."""
import ctypes

import numpy as np
import pytest

from spinifel_amd import _lib

pytestmark = pytest.mark.gpu

LIMIT = 32 << 20      # MTIP_STATIC_SCRATCH_BYTES


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _operands(torch, engine, N_o, N_i, N_p, precision, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    ld = engine.padded(N_p)
    S = torch.zeros((N_o, ld), dtype=torch.float32)
    D = torch.zeros((N_i, ld), dtype=torch.float32)
    S[:, :N_p] = torch.rand((N_o, N_p), generator=g)
    D[:, :N_p] = torch.rand((N_i, N_p), generator=g)
    S, D = S.cuda(), D.cuda()
    Sop = engine.residual(S, precision, N_p)
    Dop = engine.residual(D, precision, N_p)
    return Sop, Dop, ld


def _call(lib, torch, fn, code, Sop, Dop, ld, N_o, N_i, N_p, K, alpha, stream, ws=None):
    idx = torch.full((N_i, K), -7, dtype=torch.int32, device="cuda")
    val = torch.full((N_i, K), -7.0, dtype=torch.float32, device="cuda")
    args = [code, _vp(Sop.data), ld, _vp(Sop.scale), _vp(Sop.sumsq), N_o, _vp(Dop.data), ld,
            _vp(Dop.scale), N_i, N_p, _vp(alpha), K]
    if ws is not None:
        args += [_vp(ws), ws.numel()]
    args += [_vp(idx), _vp(val), stream]
    _lib.check(fn(*args), fn.__name__)
    return idx, val


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp8"])
def test_static_entry_point_equals_workspace_entry_point(precision):
    import torch
    from spinifel_amd import engine

    engine.require_gpu()
    lib = _lib.load()
    code = _lib.PRECISIONS[precision]
    alpha = torch.ones(1, dtype=torch.float64, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    K = 1 if precision == "fp32" else 8
    # first call large enough for the 256-row tile (bf16 / fp8: 12 x 18 tiles
    # of 256), second smaller
    for N_o, N_i, N_p, seed in ((3000, 4400, 576, 1), (700, 333, 192, 2)):
        print(f"no-alloc {precision}: N_o={N_o} N_i={N_i} N_p={N_p} K={K}", flush=True)
        Sop, Dop, ld = _operands(torch, engine, N_o, N_i, N_p, precision, seed)
        ws = torch.empty(max(1, engine.workspace_bytes(N_o, N_i, K, precision)),
                         dtype=torch.uint8, device="cuda")
        i_ref, v_ref = _call(lib, torch, lib.mtip_match_topk, code, Sop, Dop, ld, N_o, N_i, N_p,
                             K, alpha, stream, ws)
        i_st, v_st = _call(lib, torch, lib.mtip_match_topk_static, code, Sop, Dop, ld, N_o, N_i,
                           N_p, K, alpha, stream)
        torch.cuda.synchronize()
        assert torch.equal(i_st, i_ref)
        assert torch.equal(v_st.view(torch.int32), v_ref.view(torch.int32))
        assert int(i_st.min()) >= 0 and int(i_st.max()) < N_o


def test_static_entry_point_rejects_sizes_beyond_its_scratch():
    import torch
    from spinifel_amd import engine

    engine.require_gpu()
    lib = _lib.load()
    # 10,000 x ceil(10,000 / 128) x 8 keys of 8 bytes: 50.6 MB > 32 MiB (FP32 uses 128 rows)
    N_o = N_i = 10000
    assert N_i * -(-N_o // 128) * 8 * 8 > LIMIT
    fake = ctypes.c_void_p(0x1000)    # rejected before any GPU work
    rc = lib.mtip_match_topk_static(0, fake, 16384, None, fake, N_o, fake, 16384, None, N_i,
                                    16384, fake, 8, fake, fake, None)
    assert rc == _lib.MTIP_ERR_ARG and "static scratch" in _lib.last_error()
