"""Batches of uneven catchments (accu_threshold over many catchments, autotuned tile shape) on the GPU."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def fw():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lue_amd import framework

    return framework


def test_batch_shapes_follow_the_distribution():
    from lue_amd import synthetic as syn

    shapes = syn.batch_shapes(7, 2000, 64, 2048)
    sides = np.array([np.sqrt(r * c) for r, c in shapes])
    aspects = np.array([c / r for r, c in shapes])
    assert sides.min() >= 60 and sides.max() <= 2100
    assert 0.45 <= aspects.min() and aspects.max() <= 2.1
    # log-uniform: the median side is near the geometric mean of the bounds
    assert abs(np.log(np.median(sides)) - 0.5 * (np.log(64) + np.log(2048))) < 0.2


@pytest.mark.gpu
def test_autotuned_batch_accu_threshold_matches_oracle(fw, oracle_mod):
    """A batch of catchments with sides log-uniform on 48..320 and aspect U[0.5, 2], each with its own
    seeded DEM; accu_threshold on the autotuned network equals each catchment routed alone and the oracle."""
    from lue_amd import autotune
    from lue_amd import synthetic as syn

    seed = 20261205
    shapes = syn.batch_shapes(seed, 24, 48, 320)
    ldds, mats, thrs, host = [], [], [], []
    for i, (r, c) in enumerate(shapes):
        s = seed + 1 + i
        ldd = syn.ldd_host(s, r, c)
        mat = syn.uniform_host(s, syn.STREAM_MATERIAL, 0.0, 1.0, (r, c))
        thr = syn.uniform_host(s, syn.STREAM_THRESHOLD, 0.0, 0.5, (r, c))
        host.append((ldd, mat, thr))
        ldds.append(fw.PartitionedArray.from_numpy(ldd))
        mats.append(fw.PartitionedArray.from_numpy(mat))
        thrs.append(fw.PartitionedArray.from_numpy(thr))
    net, table, _ = autotune.tune(ldds, lambda n: n.accu_threshold(mats, thrs))
    assert len(table) >= 2 and net.info()["catchments"] == len(shapes)
    outs, rems = net.accu_threshold(mats, thrs)
    shape = net.tile_shape
    torch.cuda.synchronize()
    for i in (0, 5, 11, 23):
        alone = fw.FlowNetwork(ldds[i], tile_shape=shape)
        o, r = alone.accu_threshold(mats[i], thrs[i])
        torch.cuda.synchronize()
        assert np.array_equal(o.numpy(), outs[i].numpy(), equal_nan=True)
        assert np.array_equal(r.numpy(), rems[i].numpy(), equal_nan=True)
    for i, (ldd, mat, thr) in enumerate(host):
        w_out, w_rem = oracle_mod.accu_threshold(ldd, mat.astype(np.float64), thr.astype(np.float64), None,
                                                 np.float64)
        scale = np.abs(w_out) + np.abs(w_rem) + 1e-30
        assert np.max(np.abs(outs[i].numpy() - w_out) / scale) <= 1e-6
        assert np.max(np.abs(rems[i].numpy() - w_rem) / scale) <= 1e-6
