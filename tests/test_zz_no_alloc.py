"""tfsm_ranker_gradient_noalloc (the entry point that allocates nothing: its
sums live in a fixed device array) against tfsm_ranker_gradient on the same
rows, through the C ABI: two calls in a row, the second with fewer rows, so
a sum left from the first call would show.  Inputs and outputs are device
buffers owned by the caller (torch tensors here)."""
import ctypes

import numpy as np
import pytest

import tfsm
from conftest import workload_path

pytestmark = pytest.mark.gpu


def test_noalloc_gradient_matches(workdir):
    import torch
    g = tfsm.Graph(workload_path("gnm2k", 1, workdir), 80)
    ctx = tfsm.Context(g)
    rng = np.random.default_rng(31)
    w = rng.normal(0, 0.3, tfsm.RANKER_PARAMS).astype(np.float32)
    ctx.ranker_set(w)
    lib = tfsm.gpu_lib()
    for n in (20000, 3000):
        X = rng.normal(0.3, 0.6, size=(n, tfsm.RANKER_FEATURES)).astype(np.float32)
        X[:, 23] = 1.0
        y = (rng.random(n) < 0.3).astype(np.float32)
        ref_g, ref_loss = ctx.ranker_gradient(X, y)
        dx = torch.from_numpy(X).to("cuda")
        dy = torch.from_numpy(y).to("cuda")
        dg = torch.full((tfsm.RANKER_PARAMS,), float("nan"), dtype=torch.float32, device="cuda")
        dl = torch.full((1,), float("nan"), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()            # the library works on its own stream
        rc = lib.tfsm_ranker_gradient_noalloc(ctx.h, ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()),
                                              n, ctypes.c_void_p(dg.data_ptr()), ctypes.c_void_p(dl.data_ptr()))
        assert rc == 0, tfsm.gpu_lib().tfsm_last_error(ctx.h)
        got = dg.cpu().numpy()
        loss = float(dl.cpu().numpy()[0])
        # both are FP32 sums of the same products in atomic order
        tol = 1e-4 * np.abs(ref_g).max()
        assert np.all(np.abs(got - ref_g) <= tol + 1e-4 * np.abs(ref_g)), n
        assert loss == pytest.approx(ref_loss, rel=1e-4), n
    # sizes beyond the stated maximum are an error, and nothing is launched
    rc = lib.tfsm_ranker_gradient_noalloc(ctx.h, ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()),
                                          (1 << 31) + 1, ctypes.c_void_p(dg.data_ptr()), None)
    assert rc == -4
    ctx.close()
    g.close()
