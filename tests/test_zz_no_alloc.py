"""The no-allocation framelet gradient (fl_grad_framelet_noalloc) against
fl_grad_framelet through the C ABI: same gradient bit for bit and the same
loss, on two calls in a row (the second on a smaller volume, which reuses the
static scratch array the first one left), and a volume beyond the static
scratch is refused before any device work."""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _p(t):
    return C.c_void_p(t.data_ptr())


def _both(L, st, dim, kind, level, shape, seed, dev):
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.randn(shape, generator=g).to(dev)
    nb = L.fl_framelet_num_bands(dim, kind)
    nx, ny, nz = (shape[-1], shape[-2], shape[0] if dim == 3 else 1)
    work = torch.empty((level * nb,) + tuple(shape), device=dev)
    g_ref = torch.full(shape, 0.25, device=dev)        # accumulated into
    g_new = g_ref.clone()
    loss_ref = C.c_double(0.0)
    from framelab_amd import _lib
    _lib.check(L.fl_grad_framelet(dim, kind, level, nx, ny, nz, _p(x), 1e-3, 1e-6, 0, _p(g_ref), _p(work),
                                  C.byref(loss_ref), st))
    d_loss = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    _lib.check(L.fl_grad_framelet_noalloc(dim, kind, level, nx, ny, nz, _p(x), 1e-3, 1e-6, 0, _p(g_new), _p(work),
                                          _p(d_loss), st))
    torch.cuda.synchronize()
    return g_ref, g_new, loss_ref.value, float(d_loss.item())


def test_noalloc_framelet_gradient(gpu):
    from framelab_amd import _lib
    L = _lib.lib()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    cases = [(3, _lib.FL_FRAME_LINEAR, 2, (40, 48, 56), 1), (3, _lib.FL_FRAME_LINEAR, 1, (16, 20, 24), 2),
             (2, _lib.FL_FRAME_CUBIC, 1, (96, 80), 3)]
    for dim, kind, level, shape, seed in cases:
        _say(f"case: no-alloc framelet gradient dim={dim} kind={kind} level={level} shape={shape}")
        g_ref, g_new, l_ref, l_new = _both(L, st, dim, kind, level, shape, seed, gpu)
        assert torch.equal(g_new, g_ref)
        assert l_new == l_ref
    # beyond the static scratch: refused on the host, nothing launched
    big = (_lib.FL_NOALLOC_PARTIALS - 1) * 256 + 256
    rc = L.fl_grad_framelet_noalloc(3, _lib.FL_FRAME_LINEAR, 1, big, 1, 1, C.c_void_p(8), 1e-3, 1e-6, 0,
                                    C.c_void_p(8), C.c_void_p(8), None, st)
    assert rc == _lib.FL_ERR_INVALID
    assert b"too large" in L.fl_last_error_message()
