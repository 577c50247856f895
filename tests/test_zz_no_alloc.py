"""ctdm_unwrap_dct_noalloc (the allocation-free unwrap entry point) against ctdm_unwrap_dct: bit-identical
outputs on two calls in a row, the second smaller, for each flag set, called through the C ABI; sizes
beyond the fixed scratch are rejected; the calls leave device memory untouched."""
import ctypes as C

import numpy as np
import pytest
import torch

import ctdm_amd

pytestmark = pytest.mark.gpu


def _call(ctx, name, *args):
    st = getattr(ctdm_amd.lib(), name)(ctx.h, *args)
    assert st == ctdm_amd.OK, ctdm_amd.lib().ctdm_last_error(ctx.h).decode()


@pytest.mark.parametrize("flags", [0, ctdm_amd.UNWRAP_SINCOS | ctdm_amd.UNWRAP_CONGRUENT, ctdm_amd.UNWRAP_REFCOMPAT])
def test_noalloc_equals_allocating_entry_point(ctx, flags):
    for batch, M in ((60, 256), (5, 128)):  # the second call is smaller and finds the first one's contents
        w = torch.from_numpy(ctdm_amd.synth_wrapped(40 + M, batch, M, M)).cuda()
        ref = torch.empty_like(w)
        got = torch.full_like(w, float("nan"))
        _call(ctx, "ctdm_unwrap_dct", w.data_ptr(), ref.data_ptr(), batch, M, flags)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        _call(ctx, "ctdm_unwrap_dct_noalloc", w.data_ptr(), got.data_ptr(), batch, M, flags)
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0
        print(f"flags {flags:#x} batch {batch} M {M}: max|d| {(got - ref).abs().max().item():.1e}")
        assert torch.equal(got, ref), (batch, M)


def test_noalloc_rejects_batches_beyond_the_scratch(ctx):
    M = 256
    batch = (1 << 22) // (M * M) + 1
    w = torch.from_numpy(ctdm_amd.synth_wrapped(9, batch, M, M)).cuda()
    out = torch.empty_like(w)
    st = ctdm_amd.lib().ctdm_unwrap_dct_noalloc(ctx.h, w.data_ptr(), out.data_ptr(), batch, M, 0)
    assert st == ctdm_amd.EINVAL
    assert b"exceeds the fixed scratch" in ctdm_amd.lib().ctdm_last_error(ctx.h)
    # the largest accepted batch runs
    ok = (1 << 22) // (M * M)
    got = ctx.unwrap_dct_noalloc(w[:ok], out[:ok])
    torch.cuda.synchronize()
    assert torch.equal(got, ctx.unwrap_dct(w[:ok]))
