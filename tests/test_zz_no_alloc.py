"""The allocation-free network timestep (lue_amd_network_kinematic_wave_f32_noalloc) through the C ABI: its
results equal lue_amd_network_kinematic_wave_f32's bit for bit on two calls in a row (the second on a
smaller network, so the shared device array holds the first call's contents), and a network beyond its fixed
buffer is rejected with an error."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fw():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lue_amd import framework

    return framework


def _ptrs(*tensors):
    return (ctypes.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])


def _timestep(lib, name, net, q_cur, q_lat, alpha, out, stream):
    fn = getattr(lib, name)
    return fn(net._handle, _ptrs(q_cur), _ptrs(q_lat), _ptrs(alpha), 0.0, None, 0.6, 3600.0, None, 100.0,
              _ptrs(out), stream)


def _catchment(fw, seed, rows, cols):
    from lue_amd import synthetic as syn

    dev = torch.device("cuda")
    ldd = fw.PartitionedArray(syn.ldd_device(seed, rows, cols, dev))
    q_cur = syn.uniform_device(seed, syn.STREAM_CURRENT, 0.0, 5.0, (rows, cols), torch.float32, dev)
    q_lat = syn.uniform_device(seed, syn.STREAM_LATERAL, 0.0, 1e-4, (rows, cols), torch.float32, dev)
    alpha = syn.uniform_device(seed, syn.STREAM_ALPHA, 1.5, 2.5, (rows, cols), torch.float32, dev)
    return ldd, q_cur, q_lat, alpha


def test_noalloc_timestep_equals_the_network_timestep(fw):
    from lue_amd import _lib

    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    for seed, (rows, cols) in ((41, (700, 650)), (42, (300, 260))):
        ldd, q_cur, q_lat, alpha = _catchment(fw, seed, rows, cols)
        net = fw.FlowNetwork(ldd)
        want = torch.full((rows, cols), -1.0, dtype=torch.float32, device="cuda")
        got = torch.full((rows, cols), -2.0, dtype=torch.float32, device="cuda")
        _lib.check(_timestep(lib, "lue_amd_network_kinematic_wave_f32", net, q_cur, q_lat, alpha, want, stream))
        _lib.check(_timestep(lib, "lue_amd_network_kinematic_wave_f32_noalloc", net, q_cur, q_lat, alpha, got,
                             stream))
        torch.cuda.synchronize()
        assert torch.equal(got, want), (rows, cols)
        assert bool(torch.isfinite(got).all())
        # the per-call entry point's result, for a third opinion
        per_call = fw.kinematic_wave(ldd, *[fw.PartitionedArray(t) for t in (q_cur, q_lat, alpha)], 0.6, 3600.0,
                                     100.0)
        torch.cuda.synchronize()
        assert torch.equal(per_call.data, got)
        net.close()


def test_noalloc_timestep_rejects_a_network_beyond_its_buffer(fw):
    from lue_amd import _lib

    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    ldd, q_cur, q_lat, alpha = _catchment(fw, 43, 2600, 2600)
    net = fw.FlowNetwork(ldd)
    out = torch.full((2600, 2600), -1.0, dtype=torch.float32, device="cuda")
    status = _timestep(lib, "lue_amd_network_kinematic_wave_f32_noalloc", net, q_cur, q_lat, alpha, out, stream)
    assert status == _lib.STATUS_INVALID_ARGUMENT
    assert b"bytes of per-call buffers" in lib.lue_amd_last_error()
    torch.cuda.synchronize()
    assert bool((out == -1.0).all())  # nothing ran
    net.close()
