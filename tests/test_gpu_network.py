"""Flow networks (recorded sweep schedule, batches of catchments) on the GPU against the oracle."""
import numpy as np
import pytest

from conftest import BETA32, assert_equal_nan

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fw():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lue_amd import framework

    return framework


def pa(fw, a, ps=None):
    return fw.PartitionedArray.from_numpy(a, ps)


def rel_err(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    assert (np.isnan(got) == np.isnan(want)).all(), "no-data cells differ"
    err = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
    return np.nan_to_num(err, nan=0.0)


def catchment(seed, rows, cols):
    from lue_amd import synthetic as syn

    ldd = syn.ldd_host(seed, rows, cols)
    material = syn.uniform_host(seed, syn.STREAM_MATERIAL, 0.0, 1.0, (rows, cols))
    threshold = syn.uniform_host(seed, syn.STREAM_THRESHOLD, 0.0, 0.5, (rows, cols))
    q_cur = syn.uniform_host(seed, syn.STREAM_CURRENT, 0.0, 5.0, (rows, cols))
    q_lat = syn.uniform_host(seed, syn.STREAM_LATERAL, 0.0, 1e-4, (rows, cols))
    alpha = syn.uniform_host(seed, syn.STREAM_ALPHA, 1.5, 2.5, (rows, cols))
    return dict(ldd=ldd, material=material, threshold=threshold, q_cur=q_cur, q_lat=q_lat, alpha=alpha)


@pytest.mark.parametrize("tile_shape", [(32, 32), (64, 32), (64, 64), (32, 64)])
@pytest.mark.parametrize("schedule", ["chains", "rounds"])
def test_network_matches_oracle(fw, oracle_mod, tile_shape, schedule):
    c = catchment(11, 333, 257)
    ldd = pa(fw, c["ldd"])
    net = fw.FlowNetwork(ldd, tile_shape=tile_shape, schedule=schedule)
    info = net.info()
    assert info["cells"] == 333 * 257
    if schedule == "rounds":
        assert info["pending_cells"] == 0 and info["levels"] == 0
    else:
        assert info["pending_cells"] > 0 and info["levels"] >= 1 and info["rounds"] == 0
    out = net.accu(pa(fw, c["material"]))
    want = oracle_mod.accu(c["ldd"], c["material"].astype(np.float64), None, np.float64)
    torch.cuda.synchronize()
    assert rel_err(out.numpy(), want).max() <= 1e-6
    ones = net.accu(1.0)
    torch.cuda.synchronize()
    assert_equal_nan(ones.numpy(), oracle_mod.accu(c["ldd"], 1.0, None, np.float32))
    kw_args = (pa(fw, c["q_cur"]), pa(fw, c["q_lat"]), pa(fw, c["alpha"]), 0.6, 3600.0, 100.0)
    if tile_shape == (64, 64):
        # the kinematic wave's staged inputs do not fit a 64x64 tile in LDS
        from lue_amd import _lib

        with pytest.raises(_lib.RoutingError, match="does not fit"):
            net.kinematic_wave(*kw_args)
    else:
        q = net.kinematic_wave(*kw_args)
        want_q = oracle_mod.kinematic_wave(c["ldd"], *[c[k].astype(np.float64) for k in ("q_cur", "q_lat", "alpha")],
                                           BETA32, 3600.0, 100.0, None, np.float64)
        torch.cuda.synchronize()
        assert rel_err(q.numpy(), want_q).max() <= 1e-6
    out_t, rem_t = net.accu_threshold(pa(fw, c["material"]), pa(fw, c["threshold"]))
    w_out, w_rem = oracle_mod.accu_threshold(c["ldd"], c["material"].astype(np.float64),
                                             c["threshold"].astype(np.float64), None, np.float64)
    torch.cuda.synchronize()
    scale = np.abs(w_out) + np.abs(w_rem) + 1e-30
    assert np.max(np.abs(out_t.numpy() - w_out) / scale) <= 1e-6
    assert np.max(np.abs(rem_t.numpy() - w_rem) / scale) <= 1e-6
    net.close()


def test_network_replay_is_deterministic_and_matches_per_call(fw):
    from lue_amd import _lib

    c = catchment(12, 300, 400)
    ldd = pa(fw, c["ldd"])
    net = fw.FlowNetwork(ldd, tile_shape=(32, 32))  # the tile shape of per-call kinematic_wave (float)
    args = [pa(fw, c[k]) for k in ("q_cur", "q_lat", "alpha")]
    a = net.kinematic_wave(*args, 0.6, 3600.0, 100.0)
    b = net.kinematic_wave(*args, 0.6, 3600.0, 100.0)
    torch.cuda.synchronize()
    assert np.array_equal(a.numpy(), b.numpy())
    # per-call entry point with the same schedule parameters gives the same bits
    _lib.set_sweep_options("chains")
    per_call = fw.kinematic_wave(ldd, *args, 0.6, 3600.0, 100.0)
    assert np.array_equal(a.numpy(), per_call.numpy())


def test_batch_equals_each_alone(fw, oracle_mod):
    """Catchments of different shapes in one network: every result equals the catchment routed alone."""
    shapes = [(64, 200), (300, 90), (129, 129), (40, 41), (257, 300)]
    cs = [catchment(100 + i, *s) for i, s in enumerate(shapes)]
    ldds = [pa(fw, c["ldd"]) for c in cs]
    batch = fw.FlowNetwork(ldds, tile_shape=(32, 32))
    assert batch.info()["catchments"] == len(shapes)
    mats = [pa(fw, c["material"]) for c in cs]
    thrs = [pa(fw, c["threshold"]) for c in cs]
    outs, rems = batch.accu_threshold(mats, thrs)
    kin = batch.kinematic_wave([pa(fw, c["q_cur"]) for c in cs], [pa(fw, c["q_lat"]) for c in cs],
                               [pa(fw, c["alpha"]) for c in cs], 0.6, 3600.0, 100.0)
    torch.cuda.synchronize()
    for i, c in enumerate(cs):
        alone = fw.FlowNetwork(ldds[i], tile_shape=(32, 32))
        o, r = alone.accu_threshold(mats[i], thrs[i])
        k = alone.kinematic_wave(pa(fw, c["q_cur"]), pa(fw, c["q_lat"]), pa(fw, c["alpha"]), 0.6, 3600.0, 100.0)
        torch.cuda.synchronize()
        assert np.array_equal(o.numpy(), outs[i].numpy(), equal_nan=True)
        assert np.array_equal(r.numpy(), rems[i].numpy(), equal_nan=True)
        assert np.array_equal(k.numpy(), kin[i].numpy(), equal_nan=True)
        w_out, w_rem = oracle_mod.accu_threshold(c["ldd"], c["material"].astype(np.float64),
                                                 c["threshold"].astype(np.float64), None, np.float64)
        scale = np.abs(w_out) + np.abs(w_rem) + 1e-30
        assert np.max(np.abs(outs[i].numpy() - w_out) / scale) <= 1e-6


@pytest.mark.parametrize("schedule", ["chains", "rounds"])
def test_network_rejects_cycles(fw, schedule):
    from lue_amd import _lib

    ldd = np.full((100, 100), 2, dtype=np.uint8)
    ldd[50, 50], ldd[50, 51], ldd[51, 51], ldd[51, 50] = 6, 2, 4, 8
    with pytest.raises(_lib.InvalidNetworkError):
        fw.FlowNetwork(pa(fw, ldd), schedule=schedule)


@pytest.mark.parametrize("kind", ["capacity", "trigger"])
@pytest.mark.parametrize("schedule", ["chains", "rounds"])
def test_network_capacity_trigger(fw, oracle_mod, kind, schedule):
    """accu_capacity / accu_trigger on recorded networks (a batch of two catchments), FP32 and FP64, against
    the oracle; the batch equals each catchment alone."""
    cs = [catchment(400 + i, *s) for i, s in enumerate([(200, 170), (96, 260)])]
    ldds = [pa(fw, c["ldd"]) for c in cs]
    batch = fw.FlowNetwork(ldds, schedule=schedule)
    for dtype in (np.float32, np.float64):
        mats = [pa(fw, c["material"].astype(dtype)) for c in cs]
        params = [pa(fw, (c["threshold"] * 4.0).astype(dtype)) for c in cs]  # capacity / trigger U[0, 2)
        outs, rems = getattr(batch, f"accu_{kind}")(mats, params)
        torch.cuda.synchronize()
        for i, c in enumerate(cs):
            alone = fw.FlowNetwork(ldds[i], schedule=schedule)
            o, r = getattr(alone, f"accu_{kind}")(mats[i], params[i])
            torch.cuda.synchronize()
            assert np.array_equal(o.numpy(), outs[i].numpy(), equal_nan=True)
            assert np.array_equal(r.numpy(), rems[i].numpy(), equal_nan=True)
            w_out, w_rem = getattr(oracle_mod, f"accu_{kind}")(
                c["ldd"], c["material"].astype(dtype).astype(np.float64),
                (c["threshold"] * 4.0).astype(dtype).astype(np.float64), None, np.float64)
            scale = np.abs(w_out) + np.abs(w_rem) + 1e-30
            tol = 1e-6 if dtype == np.float32 else 1e-12
            assert (np.isnan(outs[i].numpy()) == np.isnan(w_out)).all()
            assert np.nanmax(np.abs(outs[i].numpy() - w_out) / scale) <= tol
            assert np.nanmax(np.abs(rems[i].numpy() - w_rem) / scale) <= tol


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_lane_walk_matches_oracle(fw, oracle_mod, dtype):
    """The lane schedule of the chain levels (lue_amd_network_set_lane_limit): work items up to the limit
    walked one per lane, cell after cell. Every operation against the oracle at the tolerances of the
    default schedule, with a catchment alone equal, bitwise, to the same catchment in a batch."""
    from lue_amd import _lib

    shapes = [(333, 257), (700, 410), (96, 520)]
    cs = [catchment(640 + i, *s) for i, s in enumerate(shapes)]
    ldds = [pa(fw, c["ldd"]) for c in cs]
    batch = fw.FlowNetwork(ldds)
    alone = fw.FlowNetwork(ldds[1])
    for net in (batch, alone):
        assert net.info()["lane_limit"] == 0
        net.set_lane_limit(1 << 20)  # every item of a level not on the wide path
        assert net.info()["lane_limit"] == 1 << 20
    tol = 1e-6 if dtype == np.float32 else 1e-12
    mats = [pa(fw, c["material"].astype(dtype)) for c in cs]
    params = [pa(fw, c["threshold"].astype(dtype)) for c in cs]
    outs = batch.accu(mats)
    torch.cuda.synchronize()
    assert np.array_equal(outs[1].numpy(), alone.accu(mats[1]).numpy(), equal_nan=True)
    for i, c in enumerate(cs):
        want = oracle_mod.accu(c["ldd"], c["material"].astype(dtype).astype(np.float64), None, np.float64)
        assert rel_err(outs[i].numpy(), want).max() <= tol
    for kind in ("threshold", "capacity", "fraction", "trigger"):
        p = params if kind != "trigger" else [pa(fw, (c["threshold"] * 4.0).astype(dtype)) for c in cs]
        o_b, r_b = getattr(batch, f"accu_{kind}")(mats, p)
        o_a, r_a = getattr(alone, f"accu_{kind}")(mats[1], p[1])
        torch.cuda.synchronize()
        assert np.array_equal(o_b[1].numpy(), o_a.numpy(), equal_nan=True)
        assert np.array_equal(r_b[1].numpy(), r_a.numpy(), equal_nan=True)
        for i, c in enumerate(cs):
            w_out, w_rem = getattr(oracle_mod, f"accu_{kind}")(
                c["ldd"], c["material"].astype(dtype).astype(np.float64),
                p[i].numpy().astype(np.float64), None, np.float64)
            scale = np.abs(w_out) + np.abs(w_rem) + 1e-30
            assert np.nanmax(np.abs(o_b[i].numpy() - w_out) / scale) <= tol, kind
            assert np.nanmax(np.abs(r_b[i].numpy() - w_rem) / scale) <= tol, kind
    qs = [[pa(fw, c[k].astype(dtype)) for k in ("q_cur", "q_lat", "alpha")] for c in cs]
    q_b = batch.kinematic_wave([q[0] for q in qs], [q[1] for q in qs], [q[2] for q in qs], 0.6, 3600.0, 100.0)
    q_a = alone.kinematic_wave(*qs[1], 0.6, 3600.0, 100.0)
    torch.cuda.synchronize()
    assert np.array_equal(q_b[1].numpy(), q_a.numpy(), equal_nan=True)
    beta = BETA32 if dtype == np.float32 else 0.6
    for i, c in enumerate(cs):
        want_q = oracle_mod.kinematic_wave(c["ldd"], *[c[k].astype(dtype).astype(np.float64)
                                                       for k in ("q_cur", "q_lat", "alpha")],
                                           beta, 3600.0, 100.0, None, np.float64)
        assert rel_err(q_b[i].numpy(), want_q).max() <= (1e-6 if dtype == np.float32 else 1e-12)
    with pytest.raises(_lib.RoutingError):
        batch.set_lane_limit(-1)
    batch.close()
    alone.close()


def test_kinematic_wave_scalar_and_raster_beta_dx_agree(fw):
    """FP32 kinematic wave: beta and the channel length as scalars (the pending cells' inputs staged in 16 B,
    beta and c1 taken from the scalars) or as rasters holding the same values (the inputs staged whole) give
    the same bits, on a network and through the per-call entry point."""
    c = catchment(71, 700, 410)
    shape = c["ldd"].shape
    ldd = pa(fw, c["ldd"])
    net = fw.FlowNetwork(ldd)
    assert net.info()["pending_cells"] > 0
    q = [pa(fw, c[k]) for k in ("q_cur", "q_lat", "alpha")]
    beta_r = pa(fw, np.full(shape, np.float32(0.6), np.float32))
    dx_r = pa(fw, np.full(shape, np.float32(100.0), np.float32))
    scalars = net.kinematic_wave(*q, 0.6, 3600.0, 100.0).numpy()
    for beta, dx in ((beta_r, dx_r), (0.6, dx_r), (beta_r, 100.0)):
        got = net.kinematic_wave(*q, beta, 3600.0, dx).numpy()
        assert np.array_equal(got, scalars, equal_nan=True)
    per_call = fw.kinematic_wave(ldd, *q, beta_r, 3600.0, dx_r).numpy()
    assert np.array_equal(per_call, fw.kinematic_wave(ldd, *q, 0.6, 3600.0, 100.0).numpy(), equal_nan=True)
    net.close()
