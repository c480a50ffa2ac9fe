"""empm_nn_sqdist_noalloc (the nearest-neighbour search of the training loss
with its split workspace in a library-owned device array) against
empm_nn_sqdist with a caller workspace, through the C-ABI: two calls in a
row on the same GPU, the second smaller, each bit-identical to the
allocating entry point (same kernels, same split plan), and both against a
float64 brute force away from float32 near-ties."""
import ctypes as C

import numpy as np
import pytest
import torch

from empm_amd import _lib
from empm_amd._lib import check

pytestmark = pytest.mark.gpu


def _search(lib, a, b, noalloc):
    n, m = a.shape[0], b.shape[0]
    dist = torch.empty(n, dtype=torch.float32, device=a.device)
    idx = torch.empty(n, dtype=torch.int32, device=a.device)
    stream = torch.cuda.current_stream(a.device).cuda_stream
    if noalloc:
        check(lib.empm_nn_sqdist_noalloc(n, a.data_ptr(), m, b.data_ptr(), dist.data_ptr(), idx.data_ptr(), stream))
    else:
        ws_bytes = int(lib.empm_nn_sqdist_workspace(n, m))
        ws = torch.empty(max(ws_bytes // 4, 1), dtype=torch.int32, device=a.device)
        check(lib.empm_nn_sqdist(n, a.data_ptr(), m, b.data_ptr(), dist.data_ptr(), idx.data_ptr(), ws.data_ptr(),
                                 C.c_size_t(ws_bytes), stream))
    return dist, idx


def test_noalloc_nearest_neighbour_equals_the_workspace_call_twice(dev):
    lib = _lib.load()
    rng = np.random.Generator(np.random.PCG64(90))
    for n, m in ((20000, 30000), (3000, 5000)):
        print(f"[noalloc nn] n={n} m={m} splits-workspace={int(lib.empm_nn_sqdist_workspace(n, m))} B", flush=True)
        assert lib.empm_nn_sqdist_workspace(n, m) > 0  # the split (scratch-using) path runs
        a_np = rng.random((n, 3)).astype(np.float32)
        b_np = rng.random((m, 3)).astype(np.float32)
        a = torch.as_tensor(a_np, device=dev)
        b = torch.as_tensor(b_np, device=dev)
        d1, i1 = _search(lib, a, b, noalloc=True)
        d0, i0 = _search(lib, a, b, noalloc=False)
        torch.cuda.synchronize()
        assert torch.equal(d1, d0) and torch.equal(i1, i0)
        # float64 brute force on a sample of queries, where the runner-up is
        # not a float32 near-tie
        q = rng.choice(n, 400, replace=False)
        D = ((a_np[q, None, :].astype(np.float64) - b_np[None, :, :].astype(np.float64)) ** 2).sum(-1)
        srt = np.sort(D, axis=1)
        clear = (srt[:, 1] - srt[:, 0]) > 1e-6 * np.maximum(srt[:, 0], 1e-12)
        ii = i1.cpu().numpy()[q]
        assert np.all(ii[clear] == np.argmin(D, axis=1)[clear])
        assert np.allclose(d1.cpu().numpy()[q], srt[:, 0], rtol=1e-5, atol=1e-9)


def test_noalloc_nearest_neighbour_argument_errors(dev):
    lib = _lib.load()
    a = torch.zeros((4, 3), device=dev)
    d = torch.empty(4, device=dev)
    i = torch.empty(4, dtype=torch.int32, device=dev)
    assert lib.empm_nn_sqdist_noalloc(4, a.data_ptr(), 0, None, d.data_ptr(), i.data_ptr(), None) == \
        _lib.EMPM_E_INVALID_ARG
    assert lib.empm_nn_sqdist_noalloc(0, None, 0, None, None, None, None) == 0
