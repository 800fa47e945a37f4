"""bbrt_trace_device_noalloc: the trace with its scratch in the library's fixed per-GPU array.

Checked against bbrt_trace_device (the same trace with scratch the handle
allocates), record for record and counter for counter, bit-exact:
* two calls in a row, the second with a smaller input (the array keeps the
  first call's contents; the counts, counters and queue bins must be cleared);
* the grouped queue (its bins live in the same array) and the learned-free
  reference Newton mode;
* ray counts beyond the array are refused on the host with an error, as are
  batch handles and the flags that need the host to wait.
"""
import numpy as np
import pytest
import torch

from billardballraytracing3d_amd import _lib as L
from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _outs(n, K):
    return dict(status=torch.full((n,), 9, dtype=torch.int8, device=DEV),
                err_kind=torch.full((n,), 9, dtype=torch.int8, device=DEV),
                bounces=torch.full((n,), -9, dtype=torch.int32, device=DEV),
                err_patch=torch.full((n,), -9, dtype=torch.int32, device=DEV),
                final_od=torch.zeros((n, 6), dtype=torch.float64, device=DEV),
                hit_patch=torch.full((n, K), -7, dtype=torch.int32, device=DEV),
                hit_xyz=torch.full((n, K, 3), 0.5, dtype=torch.float64, device=DEV),
                newton_iters=torch.zeros((n, K), dtype=torch.uint8, device=DEV))


def _ptrs(outs):
    return {k: v.data_ptr() for k, v in outs.items()}


@pytest.fixture(scope="module")
def sinai():
    sc, o, d, mb = S.sinai_cone(n_rays=1 << 15)
    return sc, o, d, mb, T.GpuScene(sc)


@pytest.mark.parametrize("mode,flags", [(L.NEWTON_CONVERGED, 0), (L.NEWTON_REF, L.FLAG_GROUP_RAYS)])
def test_noalloc_matches_allocating_trace(sinai, mode, flags):
    sc, o, d, mb, gs = sinai
    K = 12
    st = torch.cuda.Stream(DEV)
    for n in (o.shape[0], 4099):   # the second call is smaller and reuses the array the first one filled
        print(f"noalloc parity: n={n} mode={mode} flags={flags}", flush=True)
        to, td = torch.from_numpy(o[:n]).to(DEV), torch.from_numpy(d[:n]).to(DEV)
        ref = _outs(n, K)
        c_ref, _ = gs.trace_device(to.data_ptr(), td.data_ptr(), n, mb, _ptrs(ref), mode=mode, record_k=K,
                                   flags=flags)
        assert c_ref[L.CTR["bounces"]] > 0
        got = _outs(n, K)
        ctr = torch.full((L.NCOUNTERS,), -1, dtype=torch.int64, device=DEV)
        torch.cuda.synchronize()
        gs.trace_device_noalloc(to.data_ptr(), td.data_ptr(), n, mb, _ptrs(got), counters_ptr=ctr.data_ptr(),
                                mode=mode, record_k=K, flags=flags, stream=st.cuda_stream)
        st.synchronize()
        for k in ref:
            assert torch.equal(got[k], ref[k]), (n, k)
        np.testing.assert_array_equal(ctr.cpu().numpy(), c_ref, err_msg=f"counters n={n}")


def test_noalloc_refusals(sinai):
    sc, o, d, mb, gs = sinai
    n_big = L.NOALLOC_SCRATCH_BYTES // 48   # more rays than the array holds (56 bytes of scratch per ray)
    # real buffers of the full size: the refusal happens on the host before any launch
    to = torch.zeros((n_big, 3), dtype=torch.float64, device=DEV)
    td = torch.zeros((n_big, 3), dtype=torch.float64, device=DEV)
    torch.cuda.synchronize()
    with pytest.raises(L.BBRTError, match="fixed scratch"):
        gs.trace_device_noalloc(to.data_ptr(), td.data_ptr(), n_big, mb, {}, record_k=0)
    for fl in (L.FLAG_GRAPH, L.FLAG_WAVEFRONT, L.FLAG_TIME_KERNELS):
        with pytest.raises(L.BBRTError):
            gs.trace_device_noalloc(to.data_ptr(), td.data_ptr(), 16, mb, {}, record_k=0, flags=fl)
    torch.cuda.synchronize()
    gb = T.GpuScene.batch([S.two_spheres_scene(), S.two_spheres_scene()])
    with pytest.raises(L.BBRTError, match="one-scene"):
        gb.trace_device_noalloc(to.data_ptr(), td.data_ptr(), 16, mb, {}, record_k=0)
    # an empty trace writes zero counters and touches nothing else
    ctr = torch.full((L.NCOUNTERS,), -1, dtype=torch.int64, device=DEV)
    gs.trace_device_noalloc(to.data_ptr(), td.data_ptr(), 0, mb, {}, counters_ptr=ctr.data_ptr(), record_k=0)
    torch.cuda.synchronize()
    assert int(ctr.abs().sum().item()) == 0
