"""pdxg_search_device_noalloc (no workspace, no device allocation) against pdxg_search_device on
the same device buffers: identical ids, float and integer distances, counts and byte counts on two
calls in a row, the second with fewer queries; free device memory unchanged across a call; a call
beyond its fixed scratch is rejected with a clear error and the next call still works."""
import numpy as np
import pytest

from pdx_amd import pdxfile as pf
from tests.pdx_testdata import index_file

pytestmark = pytest.mark.gpu


def _outputs(torch, dev, nq, K):
    return (torch.full((nq, K), -1, dtype=torch.int32, device=dev), torch.zeros((nq, K), device=dev),
            torch.zeros((nq, K), dtype=torch.int32, device=dev), torch.zeros(nq, dtype=torch.int32, device=dev),
            torch.zeros(nq, dtype=torch.int64, device=dev))


def _host(torch, outs):
    torch.cuda.synchronize()
    return [t.cpu().numpy().copy() for t in outs]


@pytest.mark.parametrize("kind,norm", [("pdx_tree_u8", True), ("pdx_tree_f32", False), ("pdx_f32", False),
                                       ("pdx_u8", True)])
def test_noalloc_equals_search_device(kind, norm):
    import torch

    from pdx_amd import _capi

    path, x, _ = index_file(kind, 4000, 256, 40, normalize=norm)
    g = _capi.GpuIndex(pf.load(path))
    rng = np.random.default_rng(41)
    dev = torch.device("cuda", 0)
    sp = torch.cuda.current_stream().cuda_stream
    K, NP = 20, 12
    for nq in (96, 33):  # two calls in a row, the second with a smaller input
        print(f"{kind} nq={nq}", flush=True)
        qa = np.ascontiguousarray(x[rng.integers(0, len(x), nq)]
                                  + np.float32(0.05) * rng.standard_normal((nq, 256), dtype=np.float32))
        qd = torch.from_numpy(qa).to(dev)
        kd = torch.from_numpy(rng.choice([1, 10, 20], nq).astype(np.int32)).to(dev)
        nd = torch.from_numpy(rng.integers(1, NP + 1, nq).astype(np.int32)).to(dev)

        def args(o):
            return (qd.data_ptr(), nq, kd.data_ptr(), nd.data_ptr(), K, NP, K, o[0].data_ptr(), o[1].data_ptr(),
                    o[2].data_ptr(), o[3].data_ptr(), o[4].data_ptr())

        ws = _capi.Workspace(g, nq, K, NP)
        ref = _outputs(torch, dev, nq, K)
        ws.search(*args(ref), stream=sp)
        want = _host(torch, ref)
        ws.close()
        for rep in range(2):
            out = _outputs(torch, dev, nq, K)
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            _capi.search_device_noalloc(g, *args(out), stream=sp)
            got = _host(torch, out)
            free1 = torch.cuda.mem_get_info()[0]
            for a, b, name in zip(got, want, ("ids", "dists", "int dists", "counts", "bytes")):
                np.testing.assert_array_equal(a, b, err_msg=f"{name} nq={nq} rep={rep}")
            if rep == 1:
                assert free1 == free0, "the no-allocation search changed free device memory"


def test_noalloc_rejects_calls_beyond_its_scratch():
    import torch

    from pdx_amd import _capi

    path, x, _ = index_file("pdx_tree_u8", 4000, 256, 40, normalize=True)
    g = _capi.GpuIndex(pf.load(path))
    dev = torch.device("cuda", 0)
    nq, K = 40000, 10   # normalised + rotated copies alone: 2 x 40000 x 256 x 4 B > 64 MiB
    qd = torch.zeros((nq, 256), device=dev)
    out = _outputs(torch, dev, nq, K)
    with pytest.raises(_capi.PdxgError) as e:
        _capi.search_device_noalloc(g, qd.data_ptr(), nq, 0, 0, K, 8, K, out[0].data_ptr(), out[1].data_ptr(),
                                    0, out[3].data_ptr(), k_all=K, nprobe_all=8)
    assert e.value.status == 3 and "scratch" in str(e.value)
    # a call that fits still runs after the rejection
    q = torch.from_numpy(np.ascontiguousarray(x[:8])).to(dev)
    _capi.search_device_noalloc(g, q.data_ptr(), 8, 0, 0, K, 8, K, out[0].data_ptr(), out[1].data_ptr(), 0,
                                out[3].data_ptr(), k_all=K, nprobe_all=8)
    torch.cuda.synchronize()
    assert (out[3][:8].cpu().numpy() == K).all()
