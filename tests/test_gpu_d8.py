"""d8_flow_direction on the GPU (lue_amd_d8_flow_direction_*) against LUE's test vectors, the oracle and the
vectorised model of the reference's functor."""
import numpy as np
import pytest

import reference_models as rm
from conftest import load_golden, raster

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def fw():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from lue_amd import framework

    return framework


def pa(fw, array, partition_shape=None):
    return fw.PartitionedArray.from_numpy(array, partition_shape)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", load_golden("d8_flow_direction"), ids=lambda c: c["name"])
def test_d8_golden(fw, case, dtype):
    elevation = pa(fw, raster(case["elevation"], dtype), case["partition_shape"])
    got = fw.d8_flow_direction(elevation)
    assert got.partition_shape == elevation.partition_shape
    assert np.array_equal(got.numpy(), raster(case["expected"]["flow_direction"]["raster"], np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, 1), (1, 70), (33, 5), (32, 64), (65, 129), (300, 257), (1000, 777),
                                   (1, 4), (5, 8), (64, 256), (65, 260), (130, 1028), (129, 2048), (3, 4100)])
def test_d8_random_against_oracle(fw, oracle_mod, dtype, shape):
    """Ties (integer heights), no-data, negative heights, odd shapes: partial tiles and unaligned rows (the
    LDS-tile kernel), rows a multiple of 4 long (the strip kernel: partial strips, wavefront and workgroup
    borders inside a row)."""
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    dem = rng.integers(-3, 6, shape).astype(dtype)
    dem[rng.random(shape) < 0.05] = np.nan
    dem[:, shape[1] // 2:] += rng.random((shape[0], shape[1] - shape[1] // 2)).astype(dtype)
    got = fw.d8_flow_direction(pa(fw, dem)).numpy()
    assert np.array_equal(got, oracle_mod.d8_flow_direction(dem, dtype))
    assert np.array_equal(got, rm.d8_flow_direction(dem))


@pytest.mark.parametrize("shape", [(4099, 3001), (2051, 4100)])
def test_d8_synthetic_dem(fw, oracle_mod, shape):
    """The benchmark DEM (FP32): device generator equals the host one, and the ldd equals the oracle's; every
    cell of this DEM drains (no sinks but the ones the edges make)."""
    from lue_amd import synthetic as syn

    dem_d = syn.dem_device(20261208, *shape, torch.float32)
    dem_h = syn.dem_host(20261208, *shape).astype(np.float32)
    assert np.array_equal(dem_d.cpu().numpy(), dem_h)
    got = fw.d8_flow_direction(fw.PartitionedArray(dem_d)).numpy()
    assert np.array_equal(got, oracle_mod.d8_flow_direction(dem_h, np.float32))
    # the DEM falls by at least 0.1 per row: interior cells never are sinks
    assert (got[1:-1, 1:-1] != 5).all()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(37, 64), (37, 63)])
def test_d8_special_values(fw, oracle_mod, dtype, shape):
    """Infinities (inf - inf is NaN), signed zeros, subnormals and the halo value itself: the kernels reuse each
    difference negated for the opposite neighbour, which must give the drops the reference computes."""
    rng = np.random.default_rng(shape[1])
    info = np.finfo(dtype)
    pool = np.array([np.inf, -np.inf, 0.0, -0.0, info.tiny, info.smallest_subnormal, -info.smallest_subnormal,
                     1.0, 1.0 + info.eps, np.nan, info.max, -info.max], dtype=dtype)
    dem = pool[rng.integers(0, len(pool), shape)]
    got = fw.d8_flow_direction(pa(fw, dem)).numpy()
    assert np.array_equal(got, oracle_mod.d8_flow_direction(dem, dtype))
    with np.errstate(invalid="ignore", over="ignore"):
        assert np.array_equal(got, rm.d8_flow_direction(dem))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_d8_unaligned_rows_of_four(fw, oracle_mod, dtype):
    """Rows a multiple of 4 long but an elevation pointer off the 16-byte grid: the LDS-tile kernel takes over
    and gives the same ldd."""
    rows, cols = 37, 128
    rng = np.random.default_rng(3)
    dem = rng.integers(-3, 6, (rows, cols)).astype(dtype)
    dem[rng.random(dem.shape) < 0.05] = np.nan
    flat = torch.empty(rows * cols + 1, dtype=torch.float32 if dtype == np.float32 else torch.float64,
                       device="cuda")
    view = flat[1:].view(rows, cols)
    view.copy_(torch.from_numpy(dem))
    assert view.data_ptr() % 16 != 0
    got = fw.d8_flow_direction(fw.PartitionedArray(view)).numpy()
    assert np.array_equal(got, oracle_mod.d8_flow_direction(dem, dtype))


@pytest.mark.parametrize("offset", [1, 2, 3])
def test_d8_output_off_the_word_grid(fw, oracle_mod, offset):
    """An ldd output pointer 1 to 3 bytes past a 4-byte boundary (a uint8 view sliced [offset:]), written
    through the C ABI: the LDS-tile kernel stores whole words only where the address is word-aligned."""
    from lue_amd import _lib

    rows, cols = 45, 130
    rng = np.random.default_rng(17 + offset)
    dem = rng.integers(-3, 6, (rows, cols)).astype(np.float32)
    elevation = torch.from_numpy(dem).cuda()
    flat = torch.full((rows * cols + 8,), 77, dtype=torch.uint8, device="cuda")
    out = flat[offset:offset + rows * cols]
    assert out.data_ptr() % 4 == offset % 4
    _lib.check(_lib.load().lue_amd_d8_flow_direction_f32(elevation.data_ptr(), rows, cols, out.data_ptr(),
                                                        torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    got = flat.cpu().numpy()
    assert np.array_equal(got[offset:offset + rows * cols].reshape(rows, cols), oracle_mod.d8_flow_direction(dem, np.float32))
    assert (got[:offset] == 77).all() and (got[offset + rows * cols:] == 77).all()
