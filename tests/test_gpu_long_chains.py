"""Long chains: levels whose longest work item has at least 2048 positions run one work item per workgroup
(chains_wide.hpp). Networks made of a few very long flow paths, against exact counts and the oracle."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

WIDE_MIN_POSITIONS = 2048


@pytest.fixture(scope="module")
def fw():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lue_amd import framework

    return framework


def row_ldd(cols):
    """One row flowing east, off the array at its end: one chain of almost every cell."""
    return np.full((1, cols), 6, dtype=np.uint8)


def serpentine_ldd(rows, cols):
    """Rows flowing east and west in turn, joined at the ends: one flow path through every cell, leaving
    the array from the last row."""
    ldd = np.empty((rows, cols), dtype=np.uint8)
    for r in range(rows):
        east = r % 2 == 0
        ldd[r, :] = 6 if east else 4
        if r < rows - 1:
            ldd[r, cols - 1 if east else 0] = 2  # down to the next row
    return ldd


def path_order(ldd):
    """Cells in flow order along the single path (start: the cell nothing drains into)."""
    rows, cols = ldd.shape
    order = []
    for r in range(rows):
        cs = range(cols) if r % 2 == 0 else range(cols - 1, -1, -1)
        order.extend((r, c) for c in cs)
    return order


def has_wide_level(net):
    return any(level["longest_item"] >= WIDE_MIN_POSITIONS for level in net.level_info())


@pytest.mark.parametrize("make", [lambda: row_ldd(12000), lambda: serpentine_ldd(64, 4096)], ids=["row", "serpentine"])
def test_long_chain_accu_counts(fw, make):
    ldd_np = make()
    ldd = fw.PartitionedArray.from_numpy(ldd_np)
    net = fw.FlowNetwork(ldd)
    assert has_wide_level(net)
    got = net.accu(1.0, dtype=torch.float64).numpy()
    want = np.zeros(ldd_np.shape)
    if ldd_np.shape[0] == 1:
        want[0, :] = np.arange(1, ldd_np.shape[1] + 1)
    else:
        for k, cell in enumerate(path_order(ldd_np)):
            want[cell] = k + 1
    np.testing.assert_array_equal(got, want)
    got32 = fw.accu(ldd, 1.0).numpy()  # per call: the same network schedule
    np.testing.assert_array_equal(got32, want.astype(np.float32))


@pytest.mark.parametrize("dry", [False, True], ids=["wet", "with_dry_cells"])
def test_long_chain_kinematic_wave_against_oracle(fw, oracle_mod, dry):
    rng = np.random.default_rng(41)
    ldd_np = serpentine_ldd(48, 2048)
    shape = ldd_np.shape
    q_cur = rng.uniform(0.0, 5.0, shape)
    q_lat = rng.uniform(0.0, 1e-4, shape)
    alpha = rng.uniform(1.5, 2.5, shape)
    if dry:
        # a few dry cells: their groups take the one-wavefront solver (LUE's per-cell iteration)
        q_cur[rng.random(shape) < 1e-4] = 0.0
    ldd = fw.PartitionedArray.from_numpy(ldd_np)
    net = fw.FlowNetwork(ldd)
    assert has_wide_level(net)
    for dtype, tol in ((np.float64, 1e-12), (np.float32, 1e-5)):
        args = [fw.PartitionedArray.from_numpy(v.astype(dtype)) for v in (q_cur, q_lat, alpha)]
        got = net.kinematic_wave(*args, dtype(0.6), dtype(3600.0), dtype(100.0)).numpy().astype(np.float64)
        want = oracle_mod.kinematic_wave(ldd_np, q_cur.astype(dtype), q_lat.astype(dtype), alpha.astype(dtype),
                                         dtype(0.6), dtype(3600.0), dtype(100.0), None, np.float64)
        rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
        print(f"long chain kinematic wave {np.dtype(dtype).name} dry={dry}: max rel {rel.max():.3e}", flush=True)
        assert rel.max() <= tol


def test_long_chain_accu_against_oracle(fw, oracle_mod):
    rng = np.random.default_rng(43)
    ldd_np = serpentine_ldd(32, 4096)
    material = rng.random(ldd_np.shape)
    ldd = fw.PartitionedArray.from_numpy(ldd_np)
    got = fw.accu(ldd, fw.PartitionedArray.from_numpy(material)).numpy()
    want = oracle_mod.accu(ldd_np, material, None, np.float64)
    assert (np.abs(got - want) / np.maximum(np.abs(want), 1e-30)).max() <= 1e-12


SPLIT_OPS = {
    "threshold": (0.0, 0.05),  # inflow U[0,1]; the parameter range of each operation
    "capacity": (0.0, 2.0),
    "fraction": (0.0, 1.0),
}


@pytest.mark.parametrize("op", sorted(SPLIT_OPS))
def test_long_chain_split_ops_against_oracle(fw, oracle_mod, op):
    """Threshold, capacity and fraction on one flow path of 131 072 cells (its chain runs on the wide path):
    FP64 within 1e-12 and FP32 within 1e-6 of the material entering each cell, against the FP64 oracle."""
    rng = np.random.default_rng(47)
    ldd_np = serpentine_ldd(32, 4096)
    material = rng.random(ldd_np.shape)
    lo, hi = SPLIT_OPS[op]
    param = rng.uniform(lo, hi, ldd_np.shape)
    ldd = fw.PartitionedArray.from_numpy(ldd_np)
    net = fw.FlowNetwork(ldd)
    assert has_wide_level(net)
    w_out, w_rem = getattr(oracle_mod, f"accu_{op}")(ldd_np, material, param, None, np.float64)
    scale = np.abs(w_out) + np.abs(w_rem) + 1e-30
    for dtype, tol in ((np.float64, 1e-12), (np.float32, 1e-6)):
        args = [fw.PartitionedArray.from_numpy(v.astype(dtype)) for v in (material, param)]
        out, rem = getattr(net, f"accu_{op}")(*args)
        err = max(np.max(np.abs(out.numpy() - w_out) / scale), np.max(np.abs(rem.numpy() - w_rem) / scale))
        print(f"long chain {op} {np.dtype(dtype).name}: max error {err:.3e} of the entering material", flush=True)
        assert err <= tol


@pytest.mark.parametrize("op", sorted(SPLIT_OPS) + ["accu"])
def test_wide_path_equals_one_wavefront_path(fw, op):
    """A path of 1 500 cells runs on the one-wavefront kernel alone and on the wide path in a batch with a
    path of 3 000 cells (its chain is in the same level): the same bits either way."""
    rng = np.random.default_rng(53)
    short_np = row_ldd(1500)
    long_np = row_ldd(3000)
    short = fw.PartitionedArray.from_numpy(short_np)
    long_ = fw.PartitionedArray.from_numpy(long_np)
    alone = fw.FlowNetwork(short)
    batch = fw.FlowNetwork([short, long_])
    assert not has_wide_level(alone) and has_wide_level(batch)
    for dtype in (np.float64, np.float32):
        mats = [fw.PartitionedArray.from_numpy(rng.random(a.shape).astype(dtype)) for a in (short_np, long_np)]
        if op == "accu":
            got = batch.accu(mats)[0].numpy()
            want = alone.accu(mats[0]).numpy()
            assert np.array_equal(got, want, equal_nan=True)
            continue
        lo, hi = SPLIT_OPS[op]
        params = [fw.PartitionedArray.from_numpy(rng.uniform(lo, hi, a.shape).astype(dtype))
                  for a in (short_np, long_np)]
        outs, rems = getattr(batch, f"accu_{op}")(mats, params)
        out, rem = getattr(alone, f"accu_{op}")(mats[0], params[0])
        assert np.array_equal(outs[0].numpy(), out.numpy(), equal_nan=True)
        assert np.array_equal(rems[0].numpy(), rem.numpy(), equal_nan=True)


def test_wide_threshold_changes_the_schedule_not_the_results(fw):
    """The wave schedule (lue_amd_network_set_wide_threshold): every level on one wavefront per work item
    (threshold 0) gives the bits of the default schedule, whose last levels run wide."""
    from lue_amd import _lib

    rng = np.random.default_rng(59)
    ldd_np = serpentine_ldd(32, 4096)
    ldd = fw.PartitionedArray.from_numpy(ldd_np)
    net = fw.FlowNetwork(ldd)
    assert has_wide_level(net) and net.info()["wide_threshold"] == WIDE_MIN_POSITIONS
    a = fw.PartitionedArray.from_numpy(rng.random(ldd_np.shape).astype(np.float32))
    b = fw.PartitionedArray.from_numpy(rng.uniform(0.0, 0.05, ldd_np.shape).astype(np.float32))
    q = [fw.PartitionedArray.from_numpy(rng.uniform(lo, hi, ldd_np.shape).astype(np.float32))
         for lo, hi in ((0.0, 5.0), (0.0, 1e-4), (1.5, 2.5))]

    def results():
        out = [net.accu(a).numpy()]
        for op in ("threshold", "capacity", "fraction"):
            out.extend(x.numpy() for x in getattr(net, f"accu_{op}")(a, b))
        out.append(net.kinematic_wave(*q, 0.6, 3600.0, 100.0).numpy())
        return out

    wide = results()
    net.set_wide_threshold(0)
    assert net.info()["wide_threshold"] == 0
    narrow = results()
    for w, n in zip(wide, narrow):
        assert np.array_equal(w, n, equal_nan=True)
    with pytest.raises(_lib.RoutingError):
        net.set_wide_threshold(-1)


def test_tile_queue_gives_the_same_results(fw):
    """The first round as a persistent launch fed by a tile work queue (lue_amd_network_set_tile_queue) gives
    the bits of one workgroup per tile, for a batch of synthetic catchments and for a plan."""
    from lue_amd import synthetic as syn

    shapes = [(300, 520), (96, 64), (700, 410)]
    ldds = [fw.PartitionedArray.from_numpy(syn.ldd_host(61 + i, r, c)) for i, (r, c) in enumerate(shapes)]
    mats = [fw.PartitionedArray.from_numpy(syn.uniform_host(61 + i, syn.STREAM_MATERIAL, 0.0, 1.0, s))
            for i, s in enumerate(shapes)]
    thrs = [fw.PartitionedArray.from_numpy(syn.uniform_host(61 + i, syn.STREAM_THRESHOLD, 0.0, 0.5, s))
            for i, s in enumerate(shapes)]
    qs = [[fw.PartitionedArray.from_numpy(syn.uniform_host(61 + i, stream, lo, hi, s))
           for stream, lo, hi in ((syn.STREAM_CURRENT, 0.0, 5.0), (syn.STREAM_LATERAL, 0.0, 1e-4),
                                  (syn.STREAM_ALPHA, 1.5, 2.5))] for i, s in enumerate(shapes)]
    net = fw.FlowNetwork(ldds)

    def results():
        out = [x.numpy() for x in net.accu(mats)]
        o, r = net.accu_threshold(mats, thrs)
        out += [x.numpy() for x in o] + [x.numpy() for x in r]
        out += [x.numpy() for x in net.kinematic_wave([q[0] for q in qs], [q[1] for q in qs], [q[2] for q in qs],
                                                      0.6, 3600.0, 100.0)]
        plan = net.plan_accu(mats)
        plan.run()
        out += [x.numpy() for x in plan.outputs]
        plan.close()
        return out

    per_tile = results()
    net.set_tile_queue(True)
    assert net.info()["tile_queue"] == 1
    queued = results()
    for a, b in zip(per_tile, queued):
        assert np.array_equal(a, b, equal_nan=True)
