"""ctas_ring_filter_dev_noalloc: the ring filter's device entry point with
its column means in a fixed device array instead of a per-call allocation,
checked through the C ABI against ctas_ring_filter_dev (the allocating
entry point) and the oracle, on two calls in a row, the second smaller."""
import ctypes

import numpy as np
import pytest

import ctas_amd as ca
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if ca.device_count() < 1:
        pytest.skip("no HIP device")


def test_ring_filter_noalloc_matches_allocating_entry_point(torch_cuda):
    torch = torch_cuda
    L = ca.lib()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(23)
    # the second call is smaller: the array keeps the first call's means
    # beyond the second's S*w, which must not matter
    for S, N, w, box in ((6, 120, 1500, 5), (2, 33, 257, 2)):
        print(f"noalloc ring filter: S={S} N={N} w={w} box={box}", flush=True)
        P = (rng.random((S, N, w)) + np.linspace(0.0, 2.0, w)[None, None, :]).astype(np.float32)
        a = torch.from_numpy(P).cuda()
        b = a.clone()
        torch.cuda.synchronize()
        assert L.ctas_ring_filter_dev(ctypes.c_void_p(a.data_ptr()), S, N, w, box, sp) == 0
        assert L.ctas_ring_filter_dev_noalloc(ctypes.c_void_p(b.data_ptr()), S, N, w, box, sp) == 0
        torch.cuda.synchronize()
        got, ref_dev = b.cpu().numpy(), a.cpu().numpy()
        assert np.array_equal(got, ref_dev)
        ref = np.stack([oracle.ring_filter(P[s], box) for s in range(S)])
        assert np.array_equal(got, ref)


def test_ring_filter_noalloc_rejects_sizes_beyond_its_array(torch_cuda):
    torch = torch_cuda
    L = ca.lib()
    d = torch.zeros((2, 8), dtype=torch.float32, device="cuda")
    # 257 slices x 4096 columns = 1,052,672 means > 1,048,576: rejected
    # before anything is enqueued (the pointer is never read)
    rc = L.ctas_ring_filter_dev_noalloc(ctypes.c_void_p(d.data_ptr()), 257, 2, 4096, 3, None)
    assert rc == ca.CTAS_E_ARG
    assert b"exceed" in L.ctas_last_error()
    # box 0 is a no-op, as in the allocating entry point
    assert L.ctas_ring_filter_dev_noalloc(ctypes.c_void_p(d.data_ptr()), 2, 1, 8, 0, None) == 0
