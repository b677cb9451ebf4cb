"""sbx_csr_build_noalloc (no device allocation, static scratch) against sbx_csr_build, through the C ABI.

Two calls in a row, the second with fewer items and keys, so that the first call's contents left in the
shared scratch must not leak into the second; keys shared by many items exercise the long-segment path.
Integers must match bit for bit.  The size limits are refused with SBX_ERR_ARG before any work.
"""
import ctypes as C

import numpy as np
import pytest

from softbox_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from softbox_amd.context import Context

    c = Context(0)
    yield c
    c.close()


def _keys(seed, n_items, n_keys):
    g = np.random.default_rng(seed)
    key = g.integers(-2, n_keys, n_items).astype(np.int32)
    key[g.random(n_items) < 0.05] = n_keys // 3            # one key shared by ~5 % of the items
    key[g.random(n_items) < 0.01] = n_keys - 1
    return key


def _build(ctx, fn, key_dev, n_items, n_keys):
    import torch

    off = torch.full((n_keys + 1,), -5, dtype=torch.int32, device="cuda:0")
    lst = torch.full((max(1, n_items),), -5, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    rc = fn(ctx.h, C.c_void_p(key_dev.data_ptr()), n_items, n_keys, C.c_void_p(off.data_ptr()),
            C.c_void_p(lst.data_ptr()))
    ctx.check(rc)
    ctx.sync()
    return off.cpu().numpy(), lst.cpu().numpy()


@pytest.mark.parametrize("sizes", [[(300000, 250000), (4000, 1500)], [(2097152, 4194304), (1000, 3)]])
def test_noalloc_equals_csr_build_twice(ctx, sizes):
    import torch

    for call, (n_items, n_keys) in enumerate(sizes):
        print(f"noalloc call {call}: {n_items} items, {n_keys} keys", flush=True)
        key = _keys(call + n_items, n_items, n_keys)
        kd = torch.from_numpy(key).to("cuda:0")
        want_off, want_lst = _build(ctx, ctx.lib.sbx_csr_build, kd, n_items, n_keys)
        got_off, got_lst = _build(ctx, ctx.lib.sbx_csr_build_noalloc, kd, n_items, n_keys)
        assert np.array_equal(got_off, want_off), f"offsets differ (call {call})"
        m = int(want_off[-1])
        assert m == int((key >= 0).sum())
        assert np.array_equal(got_lst[:m], want_lst[:m]), f"list differs (call {call})"
        assert np.all(got_lst[m:] == -5)                    # nothing written past the total


def test_noalloc_refuses_sizes_beyond_its_scratch(ctx):
    import torch

    kd = torch.zeros(4, dtype=torch.int32, device="cuda:0")
    off = torch.zeros(8, dtype=torch.int32, device="cuda:0")
    rc = ctx.lib.sbx_csr_build_noalloc(ctx.h, C.c_void_p(kd.data_ptr()), _lib.SBX_NOALLOC_MAX_ITEMS + 1, 4,
                                       C.c_void_p(off.data_ptr()), C.c_void_p(kd.data_ptr()))
    assert rc == _lib.SBX_ERR_ARG
    assert b"exceed" in ctx.lib.sbx_last_error(ctx.h)
    rc = ctx.lib.sbx_csr_build_noalloc(ctx.h, C.c_void_p(kd.data_ptr()), 4, _lib.SBX_NOALLOC_MAX_KEYS + 1,
                                       C.c_void_p(off.data_ptr()), C.c_void_p(kd.data_ptr()))
    assert rc == _lib.SBX_ERR_ARG
    ctx.clear_error()
