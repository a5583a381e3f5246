"""wf_bvh_build_device_noalloc against wf_bvh_build_device, through the C-ABI:
the same nodes and indices, byte for byte, on two builds in a row (the
second on a smaller mesh, so stale scratch contents from the first would
show), at the size limit, and meshes above the limit are refused."""
import ctypes as C

import numpy as np
import pytest

from wavefrontpathtracing_amd import scenes
from wavefrontpathtracing_amd import _lib as L

import oracle_api as O

pytestmark = pytest.mark.gpu

LIMIT = 65536


def _build(entry, d_pos, n, bins, single, stream):
    import torch
    d_nodes = torch.full((2 * n, 8), -1, dtype=torch.int32, device="cuda")
    d_idx = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    opt = L.BuildOptions(bins, int(single))
    cnt = C.c_uint32()
    torch.cuda.synchronize()
    st = getattr(L.lib(), entry)(C.c_void_p(d_pos.data_ptr()), n, C.byref(opt), C.c_void_p(d_nodes.data_ptr()), 2 * n,
                                 C.c_void_p(d_idx.data_ptr()), C.byref(cnt), C.c_void_p(stream))
    return st, cnt.value, d_nodes, d_idx


@pytest.mark.parametrize("sizes,bins,single", [((LIMIT, 1000), 16, False), ((5000, 64), 8, True), ((2000, 1500), 64, False)])
def test_noalloc_build_equals_the_allocating_build(sizes, bins, single):
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    for k, n in enumerate(sizes):
        pos9 = O.deindex(*scenes.random_soup(100 + k + n, n))
        d_pos = torch.from_numpy(pos9).cuda()
        s_a, c_a, n_a, i_a = _build("wf_bvh_build_device", d_pos, n, bins, single, stream)
        s_b, c_b, n_b, i_b = _build("wf_bvh_build_device_noalloc", d_pos, n, bins, single, stream)
        assert s_a == 0 and s_b == 0, (s_a, s_b)
        assert c_a == c_b
        assert torch.equal(n_a[:c_a], n_b[:c_b]) and torch.equal(i_a, i_b), (n, bins, single)
        if n <= 5000:
            nodes, idx = O.blas_build(pos9, bins, single)
            assert n_b[:c_b].cpu().numpy().tobytes() == nodes.tobytes() and i_b.cpu().numpy().tobytes() == idx.tobytes()


def test_noalloc_build_refuses_meshes_above_the_limit():
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    n = LIMIT + 1
    d_pos = torch.zeros((n, 9), dtype=torch.float32, device="cuda")
    st, _, _, _ = _build("wf_bvh_build_device_noalloc", d_pos, n, 8, False, stream)
    assert st == 2   # WF_ERROR_CAPACITY
