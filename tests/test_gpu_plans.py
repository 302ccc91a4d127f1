"""Plans: one network operation with fixed rasters captured into a HIP graph (lue_amd_plan_*). Every replay
is checked against the same operation run eagerly, with the program's own eager work and synchronisations
between the replays, and with the rasters' contents changed between replays."""
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


def _catchment(fw, seed, rows, cols, dtype=torch.float32):
    from lue_amd import synthetic as syn

    dev = torch.device("cuda")
    ldd = fw.PartitionedArray(syn.ldd_device(seed, rows, cols, dev))

    def field(stream, lo, hi):
        return fw.PartitionedArray(syn.uniform_device(seed, stream, lo, hi, (rows, cols), dtype, dev))

    return ldd, field


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_kinematic_wave_plan_replays_equal_eager(fw, dtype):
    from lue_amd import synthetic as syn

    ldd, field = _catchment(fw, 51, 700, 650, dtype)
    q_cur, lateral, alpha = field(syn.STREAM_CURRENT, 0.0, 5.0), field(syn.STREAM_LATERAL, 0.0, 1e-4), \
        field(syn.STREAM_ALPHA, 1.5, 2.5)
    net = fw.FlowNetwork(ldd)
    plan = net.plan_kinematic_wave(q_cur, lateral, alpha, 0.6, 3600.0, 100.0)
    info = plan.info()
    assert info["launches"] >= 3 and info["buffer_bytes"] > 0
    for replay in range(4):
        # new lateral inflow in the same raster: the graph reads it at run time
        lateral.data.copy_(syn.uniform_device(60 + replay, syn.STREAM_LATERAL, 0.0, 1e-4, lateral.shape, dtype))
        want = net.kinematic_wave(q_cur, lateral, alpha, 0.6, 3600.0, 100.0)
        torch.cuda.synchronize()
        got = plan.run()
        torch.cuda.synchronize()
        assert torch.equal(got.data, want.data), replay
        _ = (q_cur.data * 2.0).sum().item()  # eager work between replays
    plan.close()


def test_timestep_loop_of_two_plans(fw):
    """Double-buffered state: plan A (q0 -> q1) and plan B (q1 -> q0) alternate; equal to per-call steps."""
    from lue_amd import synthetic as syn

    ldd, field = _catchment(fw, 52, 513, 400)
    alpha = field(syn.STREAM_ALPHA, 1.5, 2.5)
    lateral = field(syn.STREAM_LATERAL, 0.0, 1e-4)
    q0 = field(syn.STREAM_CURRENT, 0.0, 5.0)
    q1 = fw.PartitionedArray(torch.zeros_like(q0.data))
    net = fw.FlowNetwork(ldd)
    a = net.plan_kinematic_wave(q0, lateral, alpha, 0.6, 3600.0, 100.0, out=q1)
    b = net.plan_kinematic_wave(q1, lateral, alpha, 0.6, 3600.0, 100.0, out=q0)
    ref = fw.PartitionedArray(q0.data.clone())
    for step in range(6):
        lateral.data.copy_(syn.uniform_device(70 + step, syn.STREAM_LATERAL, 0.0, 1e-4, lateral.shape))
        ref = fw.kinematic_wave(ldd, ref, lateral, alpha, 0.6, 3600.0, 100.0)
        (a if step % 2 == 0 else b).run()
        torch.cuda.synchronize()
        state = q1 if step % 2 == 0 else q0
        assert torch.equal(state.data, ref.data), step
    a.close()
    b.close()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_accu_plans_equal_eager(fw, oracle_mod, dtype):
    from lue_amd import synthetic as syn

    ldd, field = _catchment(fw, 53, 300, 410, dtype)
    material = field(syn.STREAM_MATERIAL, 0.0, 1.0)
    net = fw.FlowNetwork(ldd)
    plan = net.plan_accu(material)
    for replay in range(3):
        material.data.mul_(1.5)
        want = net.accu(material)
        torch.cuda.synchronize()
        got = plan.run()
        torch.cuda.synchronize()
        assert torch.equal(got.data, want.data), replay
    want64 = oracle_mod.accu(ldd.numpy(), material.numpy().astype(np.float64), None, np.float64)
    rel = np.abs(plan.outputs.numpy() - want64) / np.maximum(np.abs(want64), 1e-30)
    assert rel.max() <= 1e-6


def test_batch_threshold_plan_of_40_catchments(fw):
    """A batch of 40 catchments (its argument table does not fit a kernel argument) as one graph launch."""
    from lue_amd import synthetic as syn

    shapes = syn.batch_shapes(54, 40, 48, 200)
    dev = torch.device("cuda")
    ldds, mats, thrs = [], [], []
    for i, (r, c) in enumerate(shapes):
        ldds.append(fw.PartitionedArray(syn.ldd_device(54 + i, r, c, dev)))
        mats.append(fw.PartitionedArray(syn.uniform_device(54 + i, syn.STREAM_MATERIAL, 0.0, 1.0, (r, c))))
        thrs.append(fw.PartitionedArray(syn.uniform_device(54 + i, syn.STREAM_THRESHOLD, 0.0, 0.5, (r, c))))
    net = fw.FlowNetwork(ldds)
    plan = net.plan_accu_threshold(mats, thrs)
    for replay in range(3):
        for m in mats:
            m.data.add_(0.25)
        want_out, want_rem = net.accu_threshold(mats, thrs)
        torch.cuda.synchronize()
        got_out, got_rem = plan.run()
        torch.cuda.synchronize()
        for i in range(len(shapes)):
            assert torch.equal(got_out[i].data, want_out[i].data), (replay, i)
            assert torch.equal(got_rem[i].data, want_rem[i].data), (replay, i)
    plan.close()
