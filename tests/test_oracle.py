"""Pin the oracle: LUE's own test vectors, independent brute-force models and closed forms (CPU only)."""
import math

import numpy as np
import pytest

import reference_models as rm
from conftest import (MATERIAL_DTYPES, assert_close_boost, assert_equal_nan, golden_arg, load_golden, material_raster,
                      raster)


# ---------------------------------------------------------------------------------------------------
# LUE's golden vectors (tests/golden/*.json, extracted from source/framework/algorithm/test/*.cpp)

@pytest.mark.parametrize("case", load_golden("inflow_count"), ids=lambda c: c["name"])
def test_inflow_count_golden(oracle_mod, case):
    ldd = raster(case["ldd"], np.uint8)
    got = oracle_mod.inflow_count(ldd, case["partition_shape"])
    assert_equal_nan(got, raster(case["expected"]["count"]["raster"], np.uint8))


# second argument of the two-output operations
SPLIT_SECOND = {"accu_threshold": "threshold", "accu_fraction": "fraction", "accu_capacity": "capacity",
                "accu_trigger": "trigger"}


def _run_golden(oracle_mod, case, dtype):
    ldd = raster(case["ldd"], np.uint8)
    args = {k: golden_arg(v, dtype) for k, v in case["args"].items()}
    ps = case["partition_shape"]
    kind = case["kind"]
    if kind == "accu":
        return {"outflow": oracle_mod.accu(ldd, args["inflow"], ps, dtype)}
    if kind in SPLIT_SECOND:
        fn = getattr(oracle_mod, kind)
        out, rem = fn(ldd, args["inflow"], args[SPLIT_SECOND[kind]], ps, dtype)
        return {"outflow": out, "remainder": rem}
    if kind == "kinematic_wave":
        shape = ldd.shape
        qc = args["current_outflow"]
        ql = args["inflow"]
        qc = qc if isinstance(qc, np.ndarray) else np.full(shape, qc, dtype)
        ql = ql if isinstance(ql, np.ndarray) else np.full(shape, ql, dtype)
        return {"outflow": oracle_mod.kinematic_wave(ldd, qc, ql, args["alpha"], args["beta"],
                                                     args["time_step_duration"], args["channel_length"], ps,
                                                     dtype)}
    raise AssertionError(kind)


ROUTING_CASES = [c for op in ("accu", "accu_threshold", "accu_fraction", "accu_capacity", "accu_trigger",
                              "kinematic_wave")
                 for c in load_golden(op) if c["kind"] != "iterate_to_new_discharge"]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", ROUTING_CASES, ids=lambda c: f"{c['kind']}-{c['name']}")
def test_routing_golden(oracle_mod, case, dtype):
    got = _run_golden(oracle_mod, case, dtype)
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
def test_iterate_to_new_discharge_golden(oracle_mod, case):
    dtype = np.float64 if case["element"] == "double" else np.float32
    for call in case["calls"]:
        got = oracle_mod.iterate_to_new_discharge(*call["args"], dtype=dtype)
        assert_close_boost(np.array([got]), np.array([call["want"]]), 1e-6)


# ---------------------------------------------------------------------------------------------------
# Independent models on random networks (brute force; different algorithms than LUE's router)

SHAPES = [((7, 9), (7, 9)), ((12, 17), (5, 4)), ((16, 16), (4, 8)), ((13, 11), (3, 3))]


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("shape,pshape", SHAPES)
def test_inflow_count_brute_force(oracle_mod, seed, shape, pshape):
    rng = np.random.default_rng(seed)
    ldd = rm.random_ldd(rng, *shape, nodata_fraction=0.1, pit_fraction=0.05, flat_fraction=0.3)
    assert_equal_nan(oracle_mod.inflow_count(ldd, pshape), rm.inflow_count(ldd))


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("shape,pshape", SHAPES)
def test_accu_brute_force(oracle_mod, seed, shape, pshape):
    rng = np.random.default_rng(100 + seed)
    ldd = rm.random_ldd(rng, *shape, nodata_fraction=0.05, pit_fraction=0.05, flat_fraction=0.3)
    # small integers: every order of summation is exact, so the comparison is exact
    material = rng.integers(0, 5, size=shape).astype(np.float64)
    material[rng.random(shape) < 0.03] = np.nan
    material[rng.random(shape) < 0.03] = -1.0  # outside the domain -> no-data downstream
    want = rm.accu_by_walking(ldd, material)
    for dtype in (np.float32, np.float64):
        assert_equal_nan(oracle_mod.accu(ldd, material.astype(dtype), pshape, dtype), want.astype(dtype))
    # accu(ldd, 1) is the upstream cell count, including the cell itself
    ones = rm.accu_by_walking(ldd, 1.0)
    assert_equal_nan(oracle_mod.accu(ldd, 1.0, pshape, np.float64), ones)


@pytest.mark.parametrize("kind", ["threshold", "fraction", "capacity", "trigger"])
@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("shape,pshape", SHAPES)
def test_split_brute_force(oracle_mod, kind, seed, shape, pshape):
    rng = np.random.default_rng(200 + seed)
    ldd = rm.random_ldd(rng, *shape, nodata_fraction=0.05, flat_fraction=0.3)
    material = rng.random(shape)
    param = rng.random(shape) * {"threshold": 0.5, "fraction": 1.0, "capacity": 3.0, "trigger": 2.0}[kind]
    param[rng.random(shape) < 0.03] = np.nan
    want_out, want_rem = rm.route_split(ldd, material, param, kind)
    fn = getattr(oracle_mod, f"accu_{kind}")
    out, rem = fn(ldd, material, param, pshape, np.float64)
    np.testing.assert_allclose(out, want_out, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(rem, want_rem, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("shape,pshape", SHAPES)
def test_kinematic_wave_linear_closed_form(oracle_mod, seed, shape, pshape):
    """beta = 1 makes the kinematic-wave equation linear; the Newton solve must hit the closed form."""
    rng = np.random.default_rng(300 + seed)
    ldd = rm.random_ldd(rng, *shape, flat_fraction=0.3)
    q_cur = rng.random(shape) * 5
    q_cur[rng.random(shape) < 0.1] = 0.0
    q_lat = rng.random(shape) * 1e-3
    alpha = 1.5 + rng.random(shape)
    want = rm.kinematic_wave_linear(ldd, q_cur, q_lat, alpha, 3600.0, 100.0)
    got = oracle_mod.kinematic_wave(ldd, q_cur, q_lat, alpha, 1.0, 3600.0, 100.0, pshape, np.float64)
    np.testing.assert_allclose(got, want, rtol=1e-12)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-9), (np.float32, 2e-4)])
def test_newton_residual(oracle_mod, dtype, tol):
    """Converged roots satisfy f(Q) = 0 (kinematic_wave.hpp:101-116) to the Newton stop criterion."""
    rng = np.random.default_rng(7)
    for _ in range(300):
        q_up, q_cur = rng.random(2) * 10
        q_lat = rng.random() * 1e-3
        alpha = 0.5 + 5.5 * rng.random()
        beta = 0.5 + 1.5 * rng.random()
        dt, dx = 1 + 99 * rng.random(), 1 + 99 * rng.random()
        q = oracle_mod.iterate_to_new_discharge(q_up, q_cur, q_lat, alpha, beta, dt, dx, dtype=dtype)
        res = rm.kinematic_residual(q, q_up, q_cur, q_lat, alpha, beta, dt, dx)
        assert abs(res) < tol, (q_up, q_cur, q_lat, alpha, beta, dt, dx, q, res)


def test_newton_random_input_nonnegative(oracle_mod):
    """ALGT/kinematic_wave_test.cpp:316-373: random inputs, extraction included, never negative."""
    rng = np.random.default_rng(11)
    for dtype in (np.float32, np.float64):
        for _ in range(2000):
            q_up, q_cur = rng.random(2) * 1000
            alpha, beta = 0.5 + 5.5 * rng.random(), 0.5 + 1.5 * rng.random()
            dt, dx = 1 + 99 * rng.random(), 1 + 99 * rng.random()
            q_lat = (rng.random() * 2000 - 1000) / dx
            q = oracle_mod.iterate_to_new_discharge(q_up, q_cur, q_lat, alpha, beta, dt, dx, dtype=dtype)
            assert q >= 0 and math.isfinite(q)


def test_partitioning_does_not_change_integer_results(oracle_mod):
    rng = np.random.default_rng(5)
    ldd = rm.random_ldd(rng, 40, 37, nodata_fraction=0.02, flat_fraction=0.4)
    ref = oracle_mod.accu(ldd, 1.0, (40, 37), np.float64)
    for ps in [(1, 1), (3, 5), (8, 8), (40, 1), (13, 37)]:
        assert_equal_nan(oracle_mod.accu(ldd, 1.0, ps, np.float64), ref)


def test_cycle_is_reported(oracle_mod):
    ldd = np.array([[6, 2], [8, 4]], dtype=np.uint8)  # 2x2 loop
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.accu(ldd, 1.0, (2, 2), np.float64)


def test_planted_mistake_is_caught(oracle_mod):
    """A wrong neighbour offset in a model of the router shows up against the walking model."""
    rng = np.random.default_rng(9)
    ldd = rm.random_ldd(rng, 10, 10, flat_fraction=0.3)
    want = rm.accu_by_walking(ldd, 1.0)
    saved = dict(rm.OFFSETS)
    try:
        rm.OFFSETS[3] = (1, 0)  # SE decoded as S
        wrong = rm.accu_by_walking(ldd, 1.0)
    finally:
        rm.OFFSETS.clear()
        rm.OFFSETS.update(saved)
    if (ldd == 3).any():
        assert not np.array_equal(np.nan_to_num(wrong), np.nan_to_num(want))
    assert_equal_nan(oracle_mod.accu(ldd, 1.0, (4, 4), np.float64), want)


D8_CASES = load_golden("d8_flow_direction")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", D8_CASES, ids=lambda c: c["name"])
def test_d8_flow_direction_golden(oracle_mod, case, dtype):
    got = oracle_mod.d8_flow_direction(raster(case["elevation"], dtype), dtype)
    assert np.array_equal(got, raster(case["expected"]["flow_direction"]["raster"], np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("seed", range(3))
def test_d8_flow_direction_brute_force(oracle_mod, dtype, seed):
    """Against the vectorised model on DEMs with ties (integer heights), flats, no-data and negative heights."""
    rng = np.random.default_rng(700 + seed)
    shape = (57, 83)
    dem = rng.integers(-3, 6, shape).astype(dtype)
    dem[rng.random(shape) < 0.05] = np.nan
    dem[:, 40:] += rng.random((shape[0], shape[1] - 40)).astype(dtype)
    assert np.array_equal(oracle_mod.d8_flow_direction(dem, dtype), rm.d8_flow_direction(dem))


FLOW_NEIGHBOUR_CASES = [c for op in ("upstream", "downstream") for c in load_golden(op)]


@pytest.mark.parametrize("dtype", MATERIAL_DTYPES, ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("case", FLOW_NEIGHBOUR_CASES, ids=lambda c: f"{c['kind']}-{c['name']}")
def test_flow_neighbour_golden(oracle_mod, case, dtype):
    """LUE's upstream / downstream cases, through every default material element type (the cases' values are
    small integers; their no-data follows the element type)."""
    ldd = raster(case["ldd"], np.uint8)
    got = getattr(oracle_mod, case["kind"])(ldd, material_raster(case["material"], dtype))
    want = material_raster(case["expected"]["material"]["raster"], dtype)
    assert np.array_equal(got, want, equal_nan=np.dtype(dtype).kind == "f")


def _random_material(rng, shape, dtype):
    dtype = np.dtype(dtype)
    if dtype.kind == "f":
        m = (rng.random(shape) * 10 - 2).astype(dtype)  # negative values too: no domain check here
    else:
        info = np.iinfo(dtype)
        m = rng.integers(info.min, info.max, size=shape, dtype=dtype, endpoint=True)  # sums wrap
    m[rng.random(shape) < 0.05] = rm.material_no_data(dtype)
    return m


@pytest.mark.parametrize("dtype", MATERIAL_DTYPES, ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("seed", range(3))
def test_flow_neighbour_brute_force(oracle_mod, dtype, seed):
    """upstream / downstream against the push / pull models on random ldds (no-data, pits, codes that leave the
    array, unknown codes) and random material (no-data, wrapping integer sums, float order of addition)."""
    rng = np.random.default_rng(800 + seed)
    shape = (23, 31)
    ldd = rm.random_ldd(rng, *shape, nodata_fraction=0.05, pit_fraction=0.05, flat_fraction=0.3)
    ldd[rng.random(shape) < 0.02] = 0  # no direction: downstream takes the cell's own material
    material = _random_material(rng, shape, dtype)
    exact = np.dtype(dtype).kind != "f"
    up = oracle_mod.upstream(ldd, material)
    with np.errstate(over="ignore", invalid="ignore"):
        want_up = rm.upstream_by_pushing(ldd, material)
    if exact:
        assert np.array_equal(up, want_up)
    else:
        assert np.array_equal(up, want_up, equal_nan=True)  # same order of addition
    down = oracle_mod.downstream(ldd, material)
    assert np.array_equal(down, rm.downstream_by_pulling(ldd, material), equal_nan=not exact)


def test_upstream_of_ones_is_the_inflow_count(oracle_mod):
    """upstream(ldd, 1) counts the upstream neighbours: inflow_count (definition/inflow_count.hpp:670-672)."""
    rng = np.random.default_rng(17)
    ldd = rm.random_ldd(rng, 40, 37, nodata_fraction=0.05, flat_fraction=0.3)
    count = oracle_mod.upstream(ldd, np.ones(ldd.shape, np.int32))
    want = rm.inflow_count(ldd)
    valid = ldd != 255
    assert np.array_equal(count[valid], want[valid].astype(np.int32))


@pytest.mark.parametrize("case", load_golden("downstream_distance"), ids=lambda c: c["name"])
def test_downstream_distance_golden(oracle_mod, case):
    ldd = raster(case["ldd"], np.uint8)
    want = raster(case["expected"]["distance"]["raster"], np.float32)
    for value_policies in (True, False):  # the case has no no-data ldd: both policies agree
        got = oracle_mod.downstream_distance(ldd, case["cell_size"], np.float32, value_policies)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_downstream_distance_geometry(oracle_mod, dtype):
    """The distance to the downstream cell is the length of the ldd's offset times the cell size (Euclidean,
    rounded in the element type); no-data ldd: NaN under value policies, 0 otherwise."""
    rng = np.random.default_rng(31)
    ldd = rng.integers(0, 12, (19, 23)).astype(np.uint8)
    ldd[rng.random(ldd.shape) < 0.1] = 255
    cell_size = 2.5
    got = oracle_mod.downstream_distance(ldd, cell_size, dtype, True)
    for code in range(256):
        sel = ldd == code
        if not sel.any():
            continue
        if code == 255:
            assert np.isnan(got[sel]).all()
        elif code in rm.OFFSETS:
            dr, dc = rm.OFFSETS[code]
            length = np.sqrt(dtype(dr * dr + dc * dc)) * dtype(cell_size) if dr and dc else dtype(cell_size)
            assert (got[sel] == length).all()
        else:
            assert (got[sel] == 0).all()
    plain = oracle_mod.downstream_distance(ldd, cell_size, dtype, False)
    assert (plain[ldd == 255] == 0).all()


# ---------------------------------------------------------------------------------------------------
# accu_info: LUE's golden classes, a closed-form model of the classes (reference_models.accu_info_by_rule)
# and what a single partition reduces the classes to

@pytest.mark.parametrize("case", load_golden("accu_info"), ids=lambda c: c["name"])
def test_accu_info_golden(oracle_mod, case):
    ldd = raster(case["ldd"], np.uint8)
    got = oracle_mod.accu_info(ldd, case["partition_shape"])
    np.testing.assert_array_equal(got, raster(case["expected"]["cell_class"]["raster"], np.uint8))


def test_accu_info_golden_matches_the_rule():
    for case in load_golden("accu_info"):
        ldd = raster(case["ldd"], np.uint8)
        classes, alternative = rm.accu_info_by_rule(ldd, case["partition_shape"])
        np.testing.assert_array_equal(classes, raster(case["expected"]["cell_class"]["raster"], np.uint8))
        assert not alternative.any()


def _accu_info_cases(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        rows, cols = (int(v) for v in rng.integers(1, 36, size=2))
        ldd = rm.random_ldd(rng, rows, cols, nodata_fraction=float(rng.choice([0.0, 0.05])),
                            pit_fraction=float(rng.choice([0.0, 0.03])), flat_fraction=0.3)
        yield ldd, (int(rng.integers(1, rows + 1)), int(rng.integers(1, cols + 1)))


def test_accu_info_against_the_rule(oracle_mod):
    """The oracle replays the router's events; the rule derives each class from the cell's place. They agree
    on every cell whose last arrival LUE fixes; elsewhere the oracle's class is one of the two LUE can give."""
    ambiguous = 0
    for ldd, ps in _accu_info_cases(150, 21):
        got = oracle_mod.accu_info(ldd, ps)
        classes, alternative = rm.accu_info_by_rule(ldd, ps)
        fixed = alternative == 0
        np.testing.assert_array_equal(got[fixed], classes[fixed], err_msg=f"partition shape {ps}")
        free = ~fixed
        assert ((got[free] == classes[free]) | (got[free] == alternative[free])).all()
        ambiguous += int(free.sum())
    assert ambiguous > 0  # the case LUE leaves open does occur in these networks


def test_accu_info_single_partition(oracle_mod):
    """With one partition nothing crosses partitions: ridges (no upstream neighbour), sinks, cells that
    flow off the array (partition output) and intra-partition stream cells only."""
    rng = np.random.default_rng(4)
    ldd = rm.random_ldd(rng, 30, 23, nodata_fraction=0.03, pit_fraction=0.02, flat_fraction=0.3)
    got = oracle_mod.accu_info(ldd, ldd.shape)
    count = rm.inflow_count(ldd)
    want = np.full(ldd.shape, 12, dtype=np.uint8)
    rows, cols = ldd.shape
    for r in range(rows):
        for c in range(cols):
            code = int(ldd[r, c])
            if code == rm.ND:
                want[r, c] = 255
            elif count[r, c] == 0:
                want[r, c] = 11
            elif code not in rm.OFFSETS:
                want[r, c] = 17
            else:
                dr, dc = rm.OFFSETS[code]
                if not (0 <= r + dr < rows and 0 <= c + dc < cols):
                    want[r, c] = 13
    np.testing.assert_array_equal(got, want)


def test_accu_info_planted_mistake_is_caught(oracle_mod):
    """A rule that forgets the last-partition-absorbs-the-remainder grid differs from the oracle."""
    rng = np.random.default_rng(8)
    ldd = rm.random_ldd(rng, 23, 17, flat_fraction=0.3)
    ps = (5, 4)  # 23 = 4 * 5 + 3, 17 = 4 * 4 + 1: last partitions absorb the remainder
    saved = rm.partition_index
    try:
        rm.partition_index = lambda extent, part: np.arange(extent) // part
        wrong, _ = rm.accu_info_by_rule(ldd, ps)
    finally:
        rm.partition_index = saved
    right, alternative = rm.accu_info_by_rule(ldd, ps)
    got = oracle_mod.accu_info(ldd, ps)
    fixed = alternative == 0
    np.testing.assert_array_equal(got[fixed], right[fixed])
    assert not np.array_equal(wrong, right)
