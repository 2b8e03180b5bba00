"""The batch autotuner's search (serving.autotune) on a stand-in batch whose round time
is a known function of the knobs: it tries grid and graph replay, then the virtual
blocks per draw, then the job order, keeps the fastest setting of each, leaves the
batch set to it, and stops trying once the batch has no active job."""
import pytest

from apirl_amd import hip, serving


class FakeBatch:
    """Round time = cost(grid, graph, grab, order); `rounds_left` rounds of work."""

    def __init__(self, cost, rounds_left=1000):
        self.cost = cost
        self.rounds_left = rounds_left
        self.state = {"grid": 2048, "graph": True, "grab": 2, "order": 0}
        self.tried = []
        self.clock = 0.0

    def set_launch(self, grid=0, use_graph=True):
        if grid:
            self.state["grid"] = grid
        self.state["graph"] = use_graph

    def set_schedule(self, grab=2, order=0):
        self.state["grab"] = grab
        self.state["order"] = order

    def synchronize(self):
        pass

    def step(self):
        self.tried.append(dict(self.state))
        self.clock += self.cost(**self.state)
        self.rounds_left -= 1
        return max(self.rounds_left, 0)


@pytest.fixture
def fake_clock(monkeypatch):
    holder = {}

    def perf_counter():
        return holder["batch"].clock

    monkeypatch.setattr(serving.time, "perf_counter", perf_counter)
    return holder


def test_autotune_keeps_the_fastest_setting_of_each_knob(fake_clock):
    def cost(grid, graph, grab, order):
        return (abs(grid - 1024) / 1024 + (0.0 if graph else 0.3) + abs(grab - 4) * 0.1 +
                (0.0 if order == hip.Batch.ORDER_LARGEST_FIRST else 0.05) + 1.0)
    b = FakeBatch(cost)
    fake_clock["batch"] = b
    choice = serving.autotune(b, rounds=2)
    assert (choice["grid"], choice["graph"], choice["grab"], choice["order"]) == \
        (1024, True, 4, hip.Batch.ORDER_LARGEST_FIRST)
    assert b.state == {"grid": 1024, "graph": True, "grab": 4,
                       "order": hip.Batch.ORDER_LARGEST_FIRST}
    assert choice["round_s"] == pytest.approx(1.0)
    # 8 grid / graph trials, then 4 grab and 2 order trials, 2 rounds each
    assert len(b.tried) == 2 * (8 + 4 + 2)


def test_autotune_stops_when_the_batch_is_done(fake_clock):
    b = FakeBatch(lambda **k: 1.0 + (0.0 if k["grid"] == 512 else 0.5), rounds_left=5)
    fake_clock["batch"] = b
    choice = serving.autotune(b, rounds=2)
    assert len(b.tried) == 6  # three trials; the third ends the batch's work
    assert choice["grid"] == 512 and b.state["grid"] == 512
