"""The entry point that allocates nothing, oct_mlp_gradient_noalloc, called
through the C ABI and checked against oct_mlp_gradient (the allocating
entry point with the same kernels): bit-identical gradient and loss on two
calls in a row, the second with fewer rows, and a size past the fixed
scratch refused before anything is launched."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _params(seed):
    import torch

    g = np.random.default_rng(seed)
    p = np.zeros(6203, np.float32)  # OCT_MLP_PARAMS
    p[0:29] = g.normal(0, 0.5, 29)  # feature mean
    p[29:58] = g.uniform(0.5, 2.0, 29)  # feature std
    p[58:] = g.normal(0, math.sqrt(2 / 64), 6203 - 58)
    return torch.from_numpy(p).cuda()


def test_mlp_gradient_noalloc_matches_allocating_entry_point():
    import torch

    import octohip

    tns = octohip.TNS()
    ctx = tns._enter()
    lib = tns._lib
    P = _params(3)
    for call, n in enumerate((5000, 1200)):
        print(f"noalloc call {call}: n={n}", flush=True)
        g = torch.Generator(device="cuda")
        g.manual_seed(10 + call)
        X = torch.randn(n, 29, generator=g, device="cuda")
        T = torch.rand(n, generator=g, device="cuda") * 5
        g_ref = torch.empty(6145, device="cuda")
        l_ref = torch.empty(1, device="cuda")
        tns._check(lib.oct_mlp_gradient(ctx, _ptr(P), _ptr(X), _ptr(T), n, _ptr(g_ref), _ptr(l_ref)))
        g_new = torch.full((6145,), float("nan"), device="cuda")
        l_new = torch.full((1,), float("nan"), device="cuda")
        tns._check(lib.oct_mlp_gradient_noalloc(ctx, _ptr(P), _ptr(X), _ptr(T), n, _ptr(g_new), _ptr(l_new)))
        torch.cuda.synchronize()
        assert torch.isfinite(g_new).all() and torch.isfinite(l_new).all()
        assert torch.equal(g_new.view(torch.int32), g_ref.view(torch.int32)), f"call {call}: gradient bits differ"
        assert torch.equal(l_new.view(torch.int32), l_ref.view(torch.int32)), f"call {call}: loss bits differ"
    big = 8193
    X = torch.zeros(big, 29, device="cuda")
    T = torch.zeros(big, device="cuda")
    out = torch.zeros(6145, device="cuda")
    rc = lib.oct_mlp_gradient_noalloc(ctx, _ptr(P), _ptr(X), _ptr(T), big, _ptr(out), None)
    assert rc == -5  # OCT_ERR_CAPACITY
    assert "8192" in lib.oct_last_error(ctx).decode()
