"""Host logic of the serving loop on the CPU, with a stand-in for the batch
(no device calls): admission at tick boundaries, background builds, departures,
batch merges and the frame bookkeeping of every request."""
import numpy as np
import pytest

import serving


class FakeBatch:
    """Records what the loop asks of a batch."""
    made = []

    def __init__(self, reqs):
        self.reqs = list(reqs)
        self.closed = False
        self.frames = 0
        self.copied = []
        FakeBatch.made.append(self)

    def step_async(self, n):
        self.frames += n

    def sync(self):
        pass

    def set_colliders(self, *a, **k):
        pass

    def set_attached(self, *a, **k):
        pass

    def shapes_device(self, *a, **k):
        pass

    def set_constraints(self, *a, **k):
        pass

    def set_self_collision(self, *a, **k):
        pass

    def set_ground(self, *a, **k):
        pass

    def positions(self, k):
        return None

    def copy_object(self, dst, src, slot):
        assert src.reqs[slot] is self.reqs[dst]
        self.copied.append(self.reqs[dst].id)

    def close(self):
        self.closed = True


class Script:
    ramp = 5

    def colliders(self, f):
        return [], [], []

    def attached(self, f):
        return None


@pytest.fixture
def fake(monkeypatch):
    FakeBatch.made = []
    monkeypatch.setattr(serving, "build_batch", lambda reqs, lib=None, tuning=None: FakeBatch(reqs))
    return FakeBatch


def make_requests(count, per_tick, sizes=1000):
    out = []
    for j in range(count):
        scene = {"pos": np.zeros((sizes + 10 * j, 3), np.float32)}
        out.append(serving.Request(j, scene, {"table": 4096}, Script(), life=5 + j % 7, arrival_tick=j // per_tick))
    return out


def test_every_request_runs_its_frames(fake):
    reqs = make_requests(30, 3)
    srv = serving.Server(max_segments=3)
    srv.run(reqs, max_ticks=500)
    srv.close()
    for r in reqs:
        assert r.done is not None and r.frame == r.life
        assert r.started >= r.arrival and r.done == r.started + r.life - 1
    assert srv.merges > 0
    assert max(rec["segments"] for rec in srv.log) <= 10   # merges keep the batch count bounded
    # every live object moved by a merge was copied over
    copied = sum(len(b.copied) for b in fake.made)
    assert copied > 0


def test_arrivals_of_one_tick_share_a_build(fake):
    reqs = make_requests(12, 4)
    srv = serving.Server(max_segments=100, waste_limit=1.0)
    srv.run(reqs, max_ticks=200)
    srv.close()
    admits = [b for b in fake.made if not b.copied]
    # 12 requests arriving 4 per tick: 3 admission batches of 4
    assert [len(b.reqs) for b in admits] == [4, 4, 4]


def test_hbm_budget_holds_requests_back(fake):
    reqs = make_requests(6, 6, sizes=100_000)
    one = serving.hbm_bytes(reqs[-1].n, reqs[-1].table)
    srv = serving.Server(hbm_budget=2.5 * one, max_segments=100)
    srv.run(reqs, max_ticks=500)
    srv.close()
    assert all(r.done is not None for r in reqs)
    # never more than two objects live at once
    assert max(rec["live_objects"] for rec in srv.log) <= 2
    assert max(r.started - r.arrival for r in reqs) > 0
