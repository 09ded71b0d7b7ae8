"""spm_ngrad_noalloc (no device allocation; its work buffers are one static
device array per GPU) against spm_ngrad, the entry point with a context, on
two calls in a row, the second at a smaller N, in both precisions."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(spm, torch, N, prec, phi, shape=0):
    p = spm._params(N, prec, 26.0, 1.0, -0.0061, 0.30, 0.3, shape, 0)
    grad = torch.zeros_like(phi)
    nrm = ctypes.c_double(-1.0)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = spm.lib().spm_ngrad_noalloc(ctypes.byref(p), ctypes.c_void_p(phi.data_ptr()),
                                     ctypes.c_void_p(grad.data_ptr()), ctypes.byref(nrm), s)
    return rc, grad, nrm.value


@pytest.mark.parametrize("prec", [0, 1])
def test_ngrad_noalloc_matches_ngrad(spm, seeds, prec):
    import torch

    for N, seed in ((64, 3), (32, 4)):  # the second call smaller than the first
        print("ngrad_noalloc: N=%d precision=%d seed=%d" % (N, prec, seed), flush=True)
        dt = torch.float64 if prec == spm.FP64 else torch.float32
        phi = torch.as_tensor(seeds.gaussian(N, 0.1, seed)).to(dt).cuda()
        with spm.Grid(N, prec) as g:
            ref, ref_nrm = g.ngrad(phi)
        rc, grad, nrm = _noalloc(spm, torch, N, prec, phi)
        assert rc == 0, spm.lib().spm_last_error()
        torch.cuda.synchronize()
        assert torch.equal(grad, ref), (grad - ref).abs().max().item()
        assert nrm == ref_nrm


def test_ngrad_noalloc_rejects_what_does_not_fit(spm, seeds):
    import torch

    phi = torch.zeros((8, 8, 8), dtype=torch.float64, device="cuda")
    rc, _, _ = _noalloc(spm, torch, 256, spm.FP64, phi)  # checked before the field is touched
    assert rc == -1 and b"scratch" in spm.lib().spm_last_error()
    rc, _, _ = _noalloc(spm, torch, 8, spm.FP64, phi, shape=-1)
    assert rc == -1
    rc, _, _ = _noalloc(spm, torch, 8, spm.FP64, phi, shape=99)
    assert rc == -1
    # and a good call after the rejected ones
    phi = torch.as_tensor(seeds.gaussian(16, 0.1, 5)).cuda()
    rc, grad, nrm = _noalloc(spm, torch, 16, spm.FP64, phi)
    assert rc == 0
    with spm.Grid(16, spm.FP64) as g:
        ref, ref_nrm = g.ngrad(phi)
    assert torch.equal(grad, ref) and nrm == ref_nrm
