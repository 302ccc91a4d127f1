"""GPU parity of the HIP routing path (through the C ABI) against LUE's golden vectors and the oracle."""
import numpy as np
import pytest

import reference_models as rm
from conftest import BETA32, assert_close_boost, assert_equal_nan, golden_arg, load_golden, raster

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fw():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lue_amd import framework

    return framework


def pa(fw, array, partition_shape=None):
    return fw.PartitionedArray.from_numpy(array, partition_shape)


def arg(fw, value, partition_shape):
    return pa(fw, value, partition_shape) if isinstance(value, np.ndarray) else value


# ---------------------------------------------------------------------------------------------------
# LUE's own test vectors

@pytest.mark.parametrize("case", load_golden("inflow_count"), ids=lambda c: c["name"])
def test_inflow_count_golden(fw, case):
    ldd = pa(fw, raster(case["ldd"], np.uint8), case["partition_shape"])
    got = fw.inflow_count(ldd).numpy()
    assert_equal_nan(got, raster(case["expected"]["count"]["raster"], np.uint8))


ROUTING_CASES = [c for op in ("accu", "accu_threshold", "accu_fraction", "accu_capacity", "accu_trigger",
                              "kinematic_wave")
                 for c in load_golden(op) if c["kind"] != "iterate_to_new_discharge"]
# second argument of the two-output operations
SPLIT_SECOND = {"accu_threshold": "threshold", "accu_fraction": "fraction", "accu_capacity": "capacity",
                "accu_trigger": "trigger"}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", ROUTING_CASES, ids=lambda c: f"{c['kind']}-{c['name']}")
def test_routing_golden(fw, case, dtype):
    ps = case["partition_shape"]
    ldd = pa(fw, raster(case["ldd"], np.uint8), ps)
    args = {k: arg(fw, golden_arg(v, dtype), ps) for k, v in case["args"].items()}
    kind = case["kind"]
    if kind == "accu":
        got = {"outflow": fw.accu(ldd, args["inflow"])}
    elif kind in SPLIT_SECOND:
        out, rem = getattr(fw, kind)(ldd, args["inflow"], args[SPLIT_SECOND[kind]])
        got = {"outflow": out, "remainder": rem}
    else:
        shape = ldd.shape
        for key in ("current_outflow", "inflow"):
            if not isinstance(args[key], fw.PartitionedArray):
                args[key] = pa(fw, np.full(shape, args[key], dtype), ps)
        got = {"outflow": fw.kinematic_wave(ldd, args["current_outflow"], args["inflow"], args["alpha"],
                                            args["beta"], args["time_step_duration"], args["channel_length"])}
    got = {k: v.numpy() for k, v in got.items()}
    if case.get("property") == "all_positive":
        assert (got["outflow"] > 0).all()
    for key, exp in case["expected"].items():
        want = raster(exp["raster"], dtype)
        if exp["compare"] == "equal":
            assert_equal_nan(got[key], want)
        else:
            assert_close_boost(got[key], want, 1e-6)


NEWTON_CASES = [c for c in load_golden("kinematic_wave") if c["kind"] == "iterate_to_new_discharge"]


@pytest.mark.parametrize("case", NEWTON_CASES, ids=lambda c: c["name"])
def test_iterate_to_new_discharge_golden(fw, case):
    dtype = torch.float64 if case["element"] == "double" else torch.float32
    args = np.array([c["args"] for c in case["calls"]], dtype=np.float64)
    cols = [torch.tensor(args[:, k], dtype=dtype, device="cuda") for k in range(7)]
    got = fw.iterate_to_new_discharge(*cols).cpu().numpy()
    want = np.array([c["want"] for c in case["calls"]])
    assert_close_boost(got, want, 1e-6)


# ---------------------------------------------------------------------------------------------------
# Random networks against the oracle (and brute force), several partition shapes

SHAPES = [((9, 9), (3, 3)), ((37, 53), (10, 20)), ((130, 70), (64, 64)), ((200, 300), (50, 60))]


def rel_err(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    both_nan = np.isnan(got) & np.isnan(want)
    assert (np.isnan(got) == np.isnan(want)).all(), "no-data cells differ"
    err = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
    return np.where(both_nan, 0.0, err)


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("shape,pshape", SHAPES)
def test_inflow_count_random(fw, oracle_mod, seed, shape, pshape):
    rng = np.random.default_rng(seed)
    ldd = rm.random_ldd(rng, *shape, nodata_fraction=0.05, pit_fraction=0.02, flat_fraction=0.3)
    got = fw.inflow_count(pa(fw, ldd, pshape)).numpy()
    assert_equal_nan(got, oracle_mod.inflow_count(ldd, pshape))


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("shape,pshape", SHAPES)
def test_accu_random(fw, oracle_mod, seed, shape, pshape):
    rng = np.random.default_rng(10 + seed)
    ldd = rm.random_ldd(rng, *shape, nodata_fraction=0.03, pit_fraction=0.02, flat_fraction=0.3)
    material = rng.random(shape)
    material[rng.random(shape) < 0.01] = np.nan
    material[rng.random(shape) < 0.01] = -1.0
    want = oracle_mod.accu(ldd, material, pshape, np.float64)
    got64 = fw.accu(pa(fw, ldd, pshape), pa(fw, material, pshape)).numpy()
    assert rel_err(got64, want).max() <= 1e-12
    got32 = fw.accu(pa(fw, ldd, pshape), pa(fw, material.astype(np.float32), pshape)).numpy()
    want32in = oracle_mod.accu(ldd, material.astype(np.float32).astype(np.float64), pshape, np.float64)
    assert rel_err(got32, want32in).max() <= 1e-6
    # upstream cell count, exact
    ones = fw.accu(pa(fw, ldd, pshape), 1.0).numpy()
    assert_equal_nan(ones, rm.accu_by_walking(ldd, 1.0).astype(np.float32))


@pytest.mark.parametrize("kind", ["threshold", "fraction", "capacity", "trigger"])
@pytest.mark.parametrize("seed", range(2))
@pytest.mark.parametrize("shape,pshape", SHAPES)
def test_split_random(fw, oracle_mod, kind, seed, shape, pshape):
    rng = np.random.default_rng(20 + seed)
    ldd = rm.random_ldd(rng, *shape, nodata_fraction=0.03, flat_fraction=0.3)
    material = rng.random(shape)
    param = rng.random(shape) * {"threshold": 0.5, "fraction": 1.0, "capacity": 3.0, "trigger": 2.0}[kind]
    param[rng.random(shape) < 0.01] = np.nan
    fn = getattr(fw, f"accu_{kind}")
    ofn = getattr(oracle_mod, f"accu_{kind}")
    want_out, want_rem = ofn(ldd, material, param, pshape, np.float64)
    out, rem = fn(pa(fw, ldd, pshape), pa(fw, material, pshape), pa(fw, param, pshape))
    # out = total - threshold cancels: compare against the magnitude of the total material
    scale = np.nan_to_num(np.abs(want_out) + np.abs(want_rem)) + 1e-30
    assert (np.isnan(out.numpy()) == np.isnan(want_out)).all()
    assert np.nanmax(np.abs(out.numpy() - want_out) / scale) <= 1e-12
    assert np.nanmax(np.abs(rem.numpy() - want_rem) / scale) <= 1e-12


def kinwave_inputs(rng, shape, dtype):
    q_cur = (rng.random(shape) * 5).astype(dtype)
    q_lat = (rng.random(shape) * 1e-4).astype(dtype)
    alpha = (1.5 + rng.random(shape)).astype(dtype)
    return q_cur, q_lat, alpha


@pytest.mark.parametrize("seed", range(2))
@pytest.mark.parametrize("shape,pshape", SHAPES)
def test_kinematic_wave_random(fw, oracle_mod, seed, shape, pshape):
    rng = np.random.default_rng(30 + seed)
    ldd = rm.random_ldd(rng, *shape, nodata_fraction=0.02, flat_fraction=0.3)
    q_cur, q_lat, alpha = kinwave_inputs(rng, shape, np.float64)
    want = oracle_mod.kinematic_wave(ldd, q_cur, q_lat, alpha, 0.6, 3600.0, 100.0, pshape, np.float64)
    got64 = fw.kinematic_wave(pa(fw, ldd, pshape), pa(fw, q_cur, pshape), pa(fw, q_lat, pshape),
                              pa(fw, alpha, pshape), 0.6, 3600.0, 100.0).numpy()
    assert rel_err(got64, want).max() <= 1e-12
    # FP32 inputs: compare with the FP64 oracle on the same (FP32-representable) inputs
    f32 = [a.astype(np.float32) for a in (q_cur, q_lat, alpha)]
    want32in = oracle_mod.kinematic_wave(ldd, *[a.astype(np.float64) for a in f32], BETA32, 3600.0, 100.0, pshape,
                                         np.float64)
    got32 = fw.kinematic_wave(pa(fw, ldd, pshape), *[pa(fw, a, pshape) for a in f32], 0.6, 3600.0,
                              100.0).numpy()
    assert rel_err(got32, want32in).max() <= 1e-6


@pytest.mark.parametrize("beta,dt,dx", [(0.6, 3600.0, 100.0), (0.3, 60.0, 10.0), (1.4, 600.0, 250.0)])
def test_kinematic_wave_f32_wide_range(fw, oracle_mod, beta, dt, dx):
    """Discharges from 1e-8 to 1e3 (log-uniform): a power through the FP32 log2 of the whole argument loses
    |beta log2 Q| 2^-24 there, and LUE's 10 Newton steps stop short of the root in some cells (a tiny
    current discharge under a large inflow); both paths still follow LUE (FP32 within 1e-6)."""
    rng = np.random.default_rng(int(beta * 10) + int(dt))
    shape = (90, 110)
    ldd = rm.random_ldd(rng, *shape, nodata_fraction=0.02, flat_fraction=0.3)
    q_cur = np.exp(rng.uniform(np.log(1e-8), np.log(1e3), shape)).astype(np.float32)
    q_lat = np.exp(rng.uniform(np.log(1e-9), np.log(1e-3), shape)).astype(np.float32)
    alpha = rng.uniform(0.5, 5.0, shape).astype(np.float32)
    beta = float(np.float32(beta))  # as the FP32 path receives it
    want = oracle_mod.kinematic_wave(ldd, *[a.astype(np.float64) for a in (q_cur, q_lat, alpha)], beta, dt, dx,
                                     (45, 55), np.float64)
    got = fw.kinematic_wave(pa(fw, ldd, (45, 55)), *[pa(fw, a, (45, 55)) for a in (q_cur, q_lat, alpha)], beta, dt,
                            dx).numpy()
    assert rel_err(got, want).max() <= 1e-6
    # FP64 path on the same values: LUE's iteration, including where its 10 steps stop short of the root
    got64 = fw.kinematic_wave(pa(fw, ldd, (45, 55)),
                              *[pa(fw, a.astype(np.float64), (45, 55)) for a in (q_cur, q_lat, alpha)], beta, dt,
                              dx).numpy()
    assert rel_err(got64, want).max() <= 1e-10


def test_kinematic_wave_linear_closed_form(fw):
    rng = np.random.default_rng(5)
    ldd = rm.random_ldd(rng, 40, 50, flat_fraction=0.3)
    q_cur, q_lat, alpha = kinwave_inputs(rng, ldd.shape, np.float64)
    want = rm.kinematic_wave_linear(ldd, q_cur, q_lat, alpha, 3600.0, 100.0)
    got = fw.kinematic_wave(pa(fw, ldd), pa(fw, q_cur), pa(fw, q_lat), pa(fw, alpha), 1.0, 3600.0, 100.0)
    np.testing.assert_allclose(got.numpy(), want, rtol=1e-12)


def test_cycle_is_reported(fw):
    from lue_amd import _lib

    ldd = np.full((70, 70), 2, dtype=np.uint8)
    ldd[10, 10], ldd[10, 11], ldd[11, 11], ldd[11, 10] = 6, 2, 4, 8  # 2x2 loop
    with pytest.raises(_lib.InvalidNetworkError):
        fw.accu(pa(fw, ldd), 1.0)


def test_shape_mismatch_raises(fw):
    ldd = pa(fw, np.full((8, 8), 2, dtype=np.uint8), (4, 4))
    with pytest.raises(ValueError, match="Shapes differ"):
        fw.accu(ldd, pa(fw, np.ones((8, 9), np.float32), (4, 4)))


# ---------------------------------------------------------------------------------------------------
# Workload-shaped inputs (synthetic DEM) and the device generators

def test_device_generators_match_host(fw):
    from lue_amd import synthetic

    ldd = synthetic.ldd_device(20261203, 300, 700).cpu().numpy()
    assert np.array_equal(ldd, synthetic.ldd_host(20261203, 300, 700))
    u = synthetic.uniform_device(3, 1, 0.0, 5.0, (300, 700)).cpu().numpy()
    assert np.array_equal(u, synthetic.uniform_host(3, 1, 0.0, 5.0, (300, 700), np.float32))
    u64 = synthetic.uniform_device(3, 2, 1.5, 2.5, (30, 70), torch.float64).cpu().numpy()
    assert np.array_equal(u64, synthetic.uniform_host(3, 2, 1.5, 2.5, (30, 70), np.float64))


def test_config0_accu_1000(fw, oracle_mod):
    """Plain accu on one 1000 x 1000 partition of the synthetic DEM, material f32 U[0, 1)."""
    from lue_amd import synthetic

    ldd = synthetic.ldd_host(20261203, 1000, 1000)
    material = synthetic.uniform_host(20261203, synthetic.STREAM_MATERIAL, 0.0, 1.0, (1000, 1000))
    got = fw.accu(pa(fw, ldd), pa(fw, material)).numpy()
    want64 = oracle_mod.accu(ldd, material.astype(np.float64), None, np.float64)
    want32 = oracle_mod.accu(ldd, material, None, np.float32)  # LUE's own float arithmetic
    assert rel_err(got, want64).max() <= 1e-6
    assert rel_err(want32, want64).max() > rel_err(got, want64).max()  # compensation pays off
    ones = fw.accu(pa(fw, ldd), 1.0).numpy()
    assert ones[-1].sum() == 1000 * 1000  # every cell drains to the bottom row


def test_config1_kinematic_wave_2048(fw, oracle_mod):
    """kinematic_wave on the synthetic DEM (2048^2, 1024^2 partitions): FP32 path vs FP64 oracle."""
    from lue_amd import synthetic

    n = 2048
    seed = 20261204
    ldd = synthetic.ldd_host(seed, n, n)
    q_cur = synthetic.uniform_host(seed, synthetic.STREAM_CURRENT, 0.0, 5.0, (n, n), np.float64)
    q_lat = synthetic.uniform_host(seed, synthetic.STREAM_LATERAL, 0.0, 1e-4, (n, n), np.float64)
    alpha = synthetic.uniform_host(seed, synthetic.STREAM_ALPHA, 1.5, 2.5, (n, n), np.float64)
    ps = (1024, 1024)
    want = oracle_mod.kinematic_wave(ldd, q_cur, q_lat, alpha, 0.6, 3600.0, 100.0, ps, np.float64)
    got64 = fw.kinematic_wave(pa(fw, ldd, ps), pa(fw, q_cur, ps), pa(fw, q_lat, ps), pa(fw, alpha, ps), 0.6,
                              3600.0, 100.0).numpy()
    assert rel_err(got64, want).max() <= 1e-12
    f32 = [a.astype(np.float32) for a in (q_cur, q_lat, alpha)]
    want32in = oracle_mod.kinematic_wave(ldd, *[a.astype(np.float64) for a in f32], BETA32, 3600.0, 100.0, ps,
                                         np.float64)
    got32 = fw.kinematic_wave(pa(fw, ldd, ps), *[pa(fw, a, ps) for a in f32], 0.6, 3600.0, 100.0).numpy()
    err = rel_err(got32, want32in)
    assert err.max() <= 1e-5, err.max()


# ---------------------------------------------------------------------------------------------------
# Schedules after the first round of tiles (heavy-path chains, tile rounds): same results

POLICIES = {"chains": "chains", "rounds": "rounds"}


@pytest.fixture()
def policy_lib(fw):
    from lue_amd import _lib

    yield _lib
    _lib.set_sweep_options("chains")


@pytest.mark.parametrize("policy", list(POLICIES))
def test_tail_policy_accu(fw, oracle_mod, policy_lib, policy):
    from lue_amd import synthetic

    policy_lib.set_sweep_options(POLICIES[policy])
    n0, n1 = 700, 900
    ldd = synthetic.ldd_host(99, n0, n1)
    material = synthetic.uniform_host(99, synthetic.STREAM_MATERIAL, 0.0, 1.0, (n0, n1))
    ones = fw.accu(pa(fw, ldd), 1.0).numpy()
    assert_equal_nan(ones, oracle_mod.accu(ldd, 1.0, None, np.float32))
    got = fw.accu(pa(fw, ldd), pa(fw, material)).numpy()
    want = oracle_mod.accu(ldd, material.astype(np.float64), None, np.float64)
    assert rel_err(got, want).max() <= 1e-6
    got64 = fw.accu(pa(fw, ldd), pa(fw, material.astype(np.float64))).numpy()
    assert rel_err(got64, want).max() <= 1e-12


@pytest.mark.parametrize("policy", list(POLICIES))
def test_tail_policy_split_and_kinematic_wave(fw, oracle_mod, policy_lib, policy):
    from lue_amd import synthetic

    policy_lib.set_sweep_options(POLICIES[policy])
    n0, n1 = 600, 500
    seed = 7
    ldd = synthetic.ldd_host(seed, n0, n1)
    shape = (n0, n1)
    material = synthetic.uniform_host(seed, synthetic.STREAM_MATERIAL, 0.0, 1.0, shape, np.float64)
    thr = synthetic.uniform_host(seed, synthetic.STREAM_THRESHOLD, 0.0, 0.5, shape, np.float64)
    out, rem = fw.accu_threshold(pa(fw, ldd), pa(fw, material), pa(fw, thr))
    w_out, w_rem = oracle_mod.accu_threshold(ldd, material, thr, None, np.float64)
    scale = np.abs(w_out) + np.abs(w_rem) + 1e-30
    assert np.max(np.abs(out.numpy() - w_out) / scale) <= 1e-12
    assert np.max(np.abs(rem.numpy() - w_rem) / scale) <= 1e-12
    q_cur = synthetic.uniform_host(seed, synthetic.STREAM_CURRENT, 0.0, 5.0, shape, np.float32)
    q_lat = synthetic.uniform_host(seed, synthetic.STREAM_LATERAL, 0.0, 1e-4, shape, np.float32)
    alpha = synthetic.uniform_host(seed, synthetic.STREAM_ALPHA, 1.5, 2.5, shape, np.float32)
    want = oracle_mod.kinematic_wave(ldd, *[x.astype(np.float64) for x in (q_cur, q_lat, alpha)], 0.6, 3600.0,
                                     100.0, None, np.float64)
    want32 = oracle_mod.kinematic_wave(ldd, *[x.astype(np.float64) for x in (q_cur, q_lat, alpha)], BETA32, 3600.0,
                                       100.0, None, np.float64)
    got32 = fw.kinematic_wave(pa(fw, ldd), pa(fw, q_cur), pa(fw, q_lat), pa(fw, alpha), 0.6, 3600.0, 100.0).numpy()
    assert rel_err(got32, want32).max() <= 1e-6
    got64 = fw.kinematic_wave(pa(fw, ldd), *[pa(fw, x.astype(np.float64)) for x in (q_cur, q_lat, alpha)], 0.6,
                              3600.0, 100.0).numpy()
    assert rel_err(got64, want).max() <= 1e-12


@pytest.mark.parametrize("policy", list(POLICIES))
def test_cycles_are_reported_with_every_policy(fw, policy_lib, policy):
    from lue_amd import _lib

    policy_lib.set_sweep_options(POLICIES[policy])
    ldd = np.full((150, 140), 2, dtype=np.uint8)
    ldd[0, 0], ldd[0, 1], ldd[1, 1], ldd[1, 0] = 6, 2, 4, 8      # cycle without entry
    ldd[80, 70], ldd[80, 71], ldd[81, 71], ldd[81, 70] = 6, 2, 4, 8  # cycle fed from upstream
    with pytest.raises(_lib.InvalidNetworkError):
        fw.accu(pa(fw, ldd), 1.0)
    ldd[0, 0], ldd[0, 1], ldd[1, 1], ldd[1, 0] = 2, 2, 2, 2
    with pytest.raises(_lib.InvalidNetworkError):
        fw.accu(pa(fw, ldd), 1.0)
    ldd[80, 70], ldd[80, 71], ldd[81, 71], ldd[81, 70] = 2, 2, 2, 2
    ones = fw.accu(pa(fw, ldd), 1.0).numpy()
    assert ones[-1].sum() == ldd.size


def test_per_call_network_cache_follows_ldd_changes(fw, oracle_mod):
    """Per-call operations keep their recorded schedule, keyed by a fingerprint of the ldd bytes."""
    from lue_amd import _lib, synthetic

    ldd = synthetic.ldd_host(21, 300, 310)
    material = synthetic.uniform_host(21, synthetic.STREAM_MATERIAL, 0.0, 1.0, (300, 310))
    arr = pa(fw, ldd)
    mat = pa(fw, material)
    first = fw.accu(arr, mat).numpy()   # records
    second = fw.accu(arr, mat).numpy()  # replays the kept schedule
    assert np.array_equal(first, second)
    # change the ldd in place: the bottom half drains to pits
    arr.data[150:, :] = 5
    changed = fw.accu(arr, mat).numpy()
    want = oracle_mod.accu(arr.numpy(), material.astype(np.float64), None, np.float64)
    assert rel_err(changed, want).max() <= 1e-6
    try:
        _lib.set_network_cache(0)
        assert np.array_equal(fw.accu(arr, mat).numpy(), changed)
    finally:
        _lib.set_network_cache(4)


@pytest.mark.gpu
def test_scratch_blocks_reused_across_streams_and_trimmed(fw):
    """Per-call workspaces come from cached device blocks: runs on two streams that take turns on the same
    blocks, and runs after the idle blocks are freed, give the same bits."""
    import torch

    from lue_amd import _lib, synthetic

    ldd = synthetic.ldd_host(33, 520, 480)
    q_cur = synthetic.uniform_host(33, synthetic.STREAM_CURRENT, 0.0, 5.0, (520, 480))
    q_lat = synthetic.uniform_host(33, synthetic.STREAM_LATERAL, 0.0, 1e-4, (520, 480))
    alpha = synthetic.uniform_host(33, synthetic.STREAM_ALPHA, 1.5, 2.5, (520, 480))
    net = fw.FlowNetwork(pa(fw, ldd))
    args = [pa(fw, a) for a in (q_cur, q_lat, alpha)]
    want = net.kinematic_wave(*args, 0.6, 3600.0, 100.0).numpy()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = []
    torch.cuda.synchronize()
    for i in range(6):
        with torch.cuda.stream(streams[i % 2]):
            outs.append(net.kinematic_wave(*args, 0.6, 3600.0, 100.0))
    torch.cuda.synchronize()
    for o in outs:
        assert np.array_equal(o.numpy(), want)
    _lib.trim_scratch()
    assert np.array_equal(net.kinematic_wave(*args, 0.6, 3600.0, 100.0).numpy(), want)


def test_per_call_operations_are_asynchronous(fw):
    """A per-call kinematic_wave returns before its kernels finish (the stream is still busy), and the next
    operation, which reads its output, is ordered after it: the same bits as with the host waiting between."""
    from lue_amd import synthetic as syn

    n = 8192
    dev = torch.device("cuda")
    ldd = fw.PartitionedArray(syn.ldd_device(20261209, n, n, dev), (1024, 1024))
    q0, lateral, alpha = [fw.PartitionedArray(syn.uniform_device(20261209, s, lo, hi, (n, n)), (1024, 1024))
                          for s, lo, hi in ((1, 0.0, 5.0), (2, 0.0, 1e-4), (3, 1.5, 2.5))]
    r1 = fw.kinematic_wave(ldd, q0, lateral, alpha, 0.6, 3600.0, 100.0)  # records the ldd's network
    torch.cuda.synchronize()
    r2 = fw.kinematic_wave(ldd, r1, lateral, alpha, 0.6, 3600.0, 100.0)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    q1 = fw.kinematic_wave(ldd, q0, lateral, alpha, 0.6, 3600.0, 100.0)
    pending = not stream.query()
    q2 = fw.kinematic_wave(ldd, q1, lateral, alpha, 0.6, 3600.0, 100.0)
    torch.cuda.synchronize()
    assert pending
    assert torch.equal(q1.data, r1.data) and torch.equal(q2.data, r2.data)


def test_per_call_accu_on_an_unaligned_ldd_view(fw, oracle_mod):
    """An ldd view one byte off the 4-byte grid (rows not a multiple of 4 long) works per call, through the
    Python front end (network of the array) and through the C ABI (fingerprint cache), as in LUE."""
    import ctypes

    from lue_amd import _lib, synthetic

    rows, cols = 37, 41
    ldd = synthetic.ldd_host(44, rows, cols)
    material = synthetic.uniform_host(44, synthetic.STREAM_MATERIAL, 0.0, 1.0, (rows, cols))
    flat = torch.empty(rows * cols + 1, dtype=torch.uint8, device="cuda")
    view = flat[1:].view(rows, cols)
    view.copy_(torch.from_numpy(ldd))
    arr = fw.PartitionedArray(view)
    assert arr.data.data_ptr() % 4 != 0
    want = oracle_mod.accu(ldd, material.astype(np.float64), None, np.float64)
    got = fw.accu(arr, pa(fw, material)).numpy()
    assert rel_err(got, want).max() <= 1e-6
    out = torch.empty((rows, cols), dtype=torch.float32, device="cuda")
    mat = torch.from_numpy(material).cuda()
    lib = _lib.load()
    for _ in range(2):  # record, then replay from the fingerprint cache
        _lib.check(lib.lue_amd_accu_f32(arr.data.data_ptr(), rows, cols, mat.data_ptr(), ctypes.c_float(0.0),
                                        out.data_ptr(), torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), got)
