"""lue_amd.pcraster (the PCRaster-style names of the routing operations) on the GPU: the calls of the
reference's source/framework/python/test/pcraster/routing_operation_test.py, each result checked against the
oracle (the reference test only calls them)."""
import numpy as np
import pytest

import reference_models as rm

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SHAPE = (60, 40)
PARTITION = (10, 10)


@pytest.fixture(scope="module")
def env():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lue_amd import framework as lfr
    from lue_amd import pcraster as lpr

    rng = np.random.default_rng(5)
    ldd = rm.random_ldd(rng, *SHAPE, flat_fraction=0.3)
    material = (rng.random(SHAPE) * 3).astype(np.float32)
    return lfr, lpr, ldd, material


def pa(lfr, array):
    return lfr.PartitionedArray.from_numpy(array, PARTITION)


def test_accu_family(env, oracle_mod):
    lfr, lpr, ldd, material = env
    d_ldd, d_material = pa(lfr, ldd), pa(lfr, material)
    # FP32 sums against the FP64 oracle on the same inputs (the GPU's FP32 order of addition is its own)
    np.testing.assert_allclose(lpr.accuflux(d_ldd, d_material).numpy(),
                               oracle_mod.accu(ldd, material.astype(np.float64), None, np.float64), rtol=1e-6)
    for name, param in (("capacity", 0.5 * material), ("fraction", np.full(SHAPE, 0.5, np.float32)),
                        ("threshold", 0.5 * material), ("trigger", 0.5 * material)):
        oracle_op = getattr(oracle_mod, f"accu_{name}")
        m64 = material.astype(np.float64)
        for second in (pa(lfr, param), float(param.flat[0]) if name == "fraction" else None):
            if second is None:
                continue
            second64 = param.astype(np.float64) if isinstance(second, lfr.PartitionedArray) else second
            want_pair = oracle_op(ldd, m64, second64, None, np.float64)
            flux = getattr(lpr, f"accu{name}flux")(d_ldd, d_material, second).numpy()
            state = getattr(lpr, f"accu{name}state")(d_ldd, d_material, second).numpy()
            np.testing.assert_allclose(flux, want_pair[0], rtol=1e-6)
            np.testing.assert_allclose(state, want_pair[1], rtol=1e-6)
    # non-spatial material
    flux = lpr.accuthresholdflux(d_ldd, 1.5, pa(lfr, 0.5 * material)).numpy()
    want = oracle_mod.accu_threshold(ldd, 1.5, (0.5 * material).astype(np.float64), None, np.float64)[0]
    np.testing.assert_allclose(flux, want, rtol=1e-6)


def test_catchmenttotal(env, oracle_mod):
    lfr, lpr, ldd, material = env
    got = lpr.catchmenttotal(pa(lfr, material), pa(lfr, ldd)).numpy()
    want = oracle_mod.accu_threshold(ldd, material.astype(np.float64), 0.0, None, np.float64)[1]
    np.testing.assert_allclose(got, want, rtol=1e-6)


def test_kinematic(env, oracle_mod):
    lfr, lpr, ldd, material = env
    d_ldd = pa(lfr, ldd)
    spatial = pa(lfr, material + 0.5)
    # FP32 results against the FP64 oracle on the same (FP32-rounded) inputs, within 1e-5 relative
    got = lpr.kinematic(d_ldd, spatial, spatial, spatial, spatial, 1.5, 1.5, spatial).numpy()
    q = (material + 0.5).astype(np.float64)
    want = oracle_mod.kinematic_wave(ldd, q, q, q, q, 1.5, q, None, np.float64)
    np.testing.assert_allclose(got, want, rtol=1e-5)
    got = lpr.kinematic(d_ldd, spatial, spatial, 1.5, 0.6, 1.5, 1.5, 1.5).numpy()
    want = oracle_mod.kinematic_wave(ldd, q, q, 1.5, float(np.float32(0.6)), 1.5, 1.5, None, np.float64)
    np.testing.assert_allclose(got, want, rtol=1e-5)


def test_upstream_downstream(env, oracle_mod):
    lfr, lpr, ldd, material = env
    d_ldd = pa(lfr, ldd)
    assert np.array_equal(lpr.upstream(d_ldd, pa(lfr, material)).numpy(), oracle_mod.upstream(ldd, material))
    ones = np.full(SHAPE, 2.5, np.float32)
    assert np.array_equal(lpr.upstream(d_ldd, 2.5).numpy(), oracle_mod.upstream(ldd, ones))
    for dtype in (np.uint8, np.int32, np.float32):
        values = (material * 10).astype(dtype)
        assert np.array_equal(lpr.downstream(d_ldd, pa(lfr, values)).numpy(), oracle_mod.downstream(ldd, values))


def test_downstreamdist(env, oracle_mod):
    lfr, lpr, ldd, _ = env
    lpr.configuration.cell_size = 2.0
    try:
        got = lpr.downstreamdist(pa(lfr, ldd)).numpy()
    finally:
        lpr.configuration.cell_size = 1.0
    assert got.dtype == np.float32
    assert np.array_equal(got, oracle_mod.downstream_distance(ldd, 2.0, np.float32, True))


def test_lddcreate(env, oracle_mod):
    lfr, lpr, _, material = env
    got = lpr.lddcreate(pa(lfr, material), 1e30, 1e30, 1e30, 1e30).numpy()
    assert np.array_equal(got, oracle_mod.d8_flow_direction(material, np.float32))


def test_value_scale_conversions(env, oracle_mod):
    """scalar() casts other element types to float32 with no-data -> NaN; ldd() casts to uint8."""
    lfr, lpr, ldd, material = env
    counts = (material * 10).astype(np.int32)
    counts[0, 0] = np.iinfo(np.int32).min  # no-data
    as_scalar = lpr.scalar(pa(lfr, counts)).numpy()
    assert np.isnan(as_scalar[0, 0])
    assert np.array_equal(as_scalar.ravel()[1:], counts.astype(np.float32).ravel()[1:])
    got = lpr.accuflux(pa(lfr, ldd.astype(np.int32)), pa(lfr, material)).numpy()
    np.testing.assert_allclose(got, oracle_mod.accu(ldd, material.astype(np.float64), None, np.float64), rtol=1e-6)
    for name in ("accutraveltimeflux", "kinwavestate", "lddrepair"):
        with pytest.raises(NotImplementedError):
            getattr(lpr, name)()
    with pytest.raises(NotImplementedError):
        lpr.accuflux("ldd.map", 1.0)
