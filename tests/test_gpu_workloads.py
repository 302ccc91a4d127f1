"""Large-domain workload pieces: exact int64 upstream cell counts and resident timestep loops."""
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


def test_upstream_cell_count_is_exact(fw, oracle_mod):
    from lue_amd import synthetic as syn
    from lue_amd import workloads

    ldd = syn.ldd_host(31, 1500, 1300)
    got = workloads.upstream_cell_count(fw.PartitionedArray.from_numpy(ldd)).cpu().numpy()
    want = oracle_mod.accu(ldd, 1.0, (1024, 1024), np.float64)
    assert got.dtype == np.int64
    assert np.array_equal(got, want.astype(np.int64))


def test_upstream_cell_count_invariants_at_4096(fw):
    """Every cell counts itself plus the counts of the neighbours draining into it; all flow leaves
    through the bottom row, so the bottom row holds every cell once."""
    from lue_amd import synthetic as syn
    from lue_amd import workloads

    n = 4096
    ldd = syn.ldd_device(32, n, n)
    counts = workloads.upstream_cell_count(fw.PartitionedArray(ldd))
    assert int(counts[-1].sum()) == n * n
    padded = torch.nn.functional.pad(counts, (1, 1, 1, 1))
    lpad = torch.nn.functional.pad(ldd.to(torch.int32), (1, 1, 1, 1), value=255)
    expect = torch.ones_like(counts)
    # neighbour k at (dr, dc) drains here if its code points back (opposite direction)
    codes = {(-1, 0): 2, (-1, 1): 1, (0, 1): 4, (1, 1): 7, (1, 0): 8, (1, -1): 9, (0, -1): 6, (-1, -1): 3}
    for (dr, dc), code in codes.items():
        nb = padded[1 + dr:1 + dr + n, 1 + dc:1 + dc + n]
        nl = lpad[1 + dr:1 + dr + n, 1 + dc:1 + dc + n]
        expect += torch.where(nl == code, nb, torch.zeros_like(nb))
    assert torch.equal(counts, expect)


def test_large_domain_loop_small(fw):
    """The resident timestep loop at 2048^2 for 3 steps equals the per-call entry points step by step."""
    from lue_amd import synthetic as syn
    from lue_amd import workloads

    r = workloads.large_domain(seed=9, side=2048, steps=3, partition=512)
    assert r["outlet_count_sum"] == 2048 * 2048 and np.isfinite(r["final_discharge_sum"])
    dev = torch.device("cuda")
    ldd = fw.PartitionedArray(syn.ldd_device(9, 2048, 2048, dev), (512, 512))
    alpha = fw.PartitionedArray(syn.uniform_device(9, syn.STREAM_ALPHA, 1.5, 2.5, (2048, 2048)), (512, 512))
    q = fw.PartitionedArray(workloads.upstream_cell_count(ldd).to(torch.float32) * 1e-3, (512, 512))
    for step in range(3):
        lateral = fw.PartitionedArray(syn.uniform_device(10 + step, syn.STREAM_LATERAL, 0.0, 1e-4, (2048, 2048)),
                                      (512, 512))
        q = fw.kinematic_wave(ldd, q, lateral, alpha, 0.6, 3600.0, 100.0)
    assert float(q.data.double().sum()) == r["final_discharge_sum"]
