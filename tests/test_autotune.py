"""The tuner's choice (CPU): with timings fixed by a model, it keeps the fastest tile shape, wave schedule, lane
limit and first-round launch, sets the kept network to them and releases the others."""
import pytest

pytest.importorskip("torch")


class FakeNetwork:
    made = []

    def __init__(self, flow_directions, tile_shape, schedule):
        self.shape = tuple(tile_shape)
        self.threshold = 2048
        self.queue = False
        self.lane = 0
        self.closed = False
        FakeNetwork.made.append(self)

    def info(self):
        return {"wide_threshold": self.threshold, "lane_limit": self.lane}

    def set_wide_threshold(self, positions):
        self.threshold = positions

    def set_tile_queue(self, on):
        self.queue = bool(on)

    def set_lane_limit(self, positions):
        self.lane = positions

    def close(self):
        self.closed = True


def model_ms(net):
    # fastest: 64x32, threshold 4096, no queue, lane limit 64
    base = {(32, 32): 7.0, (64, 32): 6.0, (32, 64): 6.5, (64, 64): 9.0}[net.shape]
    return (base + {1024: 0.3, 2048: 0.2, 4096: 0.0, 0: 0.1}[net.threshold] + (0.5 if net.queue else 0.0) +
            {64: 0.0, 128: 0.02, 0: 0.4}[net.lane])


def test_tuner_keeps_the_fastest_combination(monkeypatch):
    from lue_amd import autotune

    FakeNetwork.made = []
    monkeypatch.setattr(autotune, "FlowNetwork", FakeNetwork)
    monkeypatch.setattr(autotune, "_time", lambda run, net, repeats: model_ms(net))
    net, table, schedule_table = autotune.tune([object()], lambda n: None)
    assert net.shape == (64, 32) and net.threshold == 4096 and net.queue is False and not net.closed
    assert net.lane == 64
    assert table == {(32, 32): 7.0, (64, 32): 6.0, (32, 64): 6.5, (64, 64): 9.0}
    assert len(schedule_table) == len(autotune.CANDIDATES) * (
        len(autotune.WIDE_THRESHOLDS) * len(autotune.TILE_QUEUE) + len(autotune.LANE_LIMITS) - 1)
    assert all(n.closed for n in FakeNetwork.made if n is not net)


def test_tuner_keeps_an_earlier_shape_when_later_ones_are_slower(monkeypatch):
    from lue_amd import autotune

    FakeNetwork.made = []
    monkeypatch.setattr(autotune, "FlowNetwork", FakeNetwork)
    monkeypatch.setattr(autotune, "_time", lambda run, net, repeats: model_ms(net) - (5.0 if net.shape == (32, 32) else 0))
    net, table, _ = autotune.tune([object()], lambda n: None)
    assert net.shape == (32, 32) and net.threshold == 4096 and not net.queue and net.lane == 64
