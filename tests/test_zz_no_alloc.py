"""mmca_adjoint_step_noalloc, the adjoint step whose workspace is the
library's fixed device array (for callers that cannot allocate), against
mmca_adjoint_step with a caller workspace, through the C ABI: bit-identical
outputs on two calls in a row, the second on a smaller block (so the array
holds the first call's values where the second block's halo lies), and a
clear error for a block beyond the array's size."""
import ctypes as C

import numpy as np
import pytest
import torch

from mumatscale_amd import abi
from mumatscale_amd.lib import lib
from mumatscale_amd.workloads import seeded_grains_config

pytestmark = pytest.mark.gpu


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _inputs(n, ny, nz, seed):
    cfg = seeded_grains_config(n=n, nseeds=2, seed=4, ny=ny, nz=nz)
    params = cfg.params(check_negative_ce=False)
    shape = (nz + 2, ny + 2, n + 2)
    g = torch.Generator().manual_seed(seed)
    dev = torch.device("cuda")
    codes = torch.randint(0, 5, shape, generator=g, dtype=torch.uint8).to(dev)
    ce = (1.0 + 9.0 * torch.rand(shape, generator=g, dtype=torch.float64)).to(dev)
    cl = (1.0 + 9.0 * torch.rand(shape, generator=g, dtype=torch.float64)).to(dev)
    fs = torch.rand(shape, generator=g, dtype=torch.float64).to(dev)
    grads = [torch.randn(shape, generator=g, dtype=torch.float64).to(dev) for _ in range(3)]
    return params, codes, (ce, cl, fs), grads


def _run(params, codes, state, grads, noalloc):
    L = lib()
    out = [torch.full_like(grads[0], 7.0) for _ in range(3)]
    g_param = torch.zeros(3, dtype=torch.float64, device=grads[0].device)
    if noalloc:
        st = L.mmca_adjoint_step_noalloc(C.byref(params), 5, _ptr(codes), *(_ptr(t) for t in state),
                                         *(_ptr(t) for t in grads), *(_ptr(t) for t in out), _ptr(g_param), None)
    else:
        nb = C.c_int64(0)
        assert L.mmca_adjoint_workspace_bytes(C.byref(params), C.byref(nb)) == abi.OK
        ws = torch.zeros((nb.value + 7) // 8, dtype=torch.float64, device=grads[0].device)
        st = L.mmca_adjoint_step(C.byref(params), 5, _ptr(codes), *(_ptr(t) for t in state),
                                 *(_ptr(t) for t in grads), *(_ptr(t) for t in out), _ptr(g_param), _ptr(ws), None)
    assert st == abi.OK
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in out] + [g_param.cpu().numpy()]


def test_noalloc_adjoint_step_matches_workspace_variant():
    for n, ny, nz, seed in ((48, 40, 30, 1), (20, 16, 12, 2)):
        print("no-alloc adjoint step: block %dx%dx%d" % (n, ny, nz), flush=True)
        params, codes, state, grads = _inputs(n, ny, nz, seed)
        ref = _run(params, codes, state, grads, noalloc=False)
        got = _run(params, codes, state, grads, noalloc=True)
        for a, b in zip(ref, got):
            assert np.array_equal(a.view(np.int64), b.view(np.int64))
        assert np.any(ref[3] != 0.0)


def test_noalloc_adjoint_step_rejects_blocks_beyond_its_array():
    params, codes, state, grads = _inputs(20, 16, 12, 3)
    big = abi.Params()
    C.pointer(big)[0] = params
    big.nx, big.ny, big.nz = 200, 200, 200
    big.gdim[0], big.gdim[1], big.gdim[2] = 200, 200, 200
    g_param = torch.zeros(3, dtype=torch.float64, device="cuda")
    st = lib().mmca_adjoint_step_noalloc(C.byref(big), 5, _ptr(codes), *(_ptr(t) for t in state),
                                         *(_ptr(t) for t in grads), *(_ptr(t) for t in grads), _ptr(g_param), None)
    assert st == abi.ERR_ARG
