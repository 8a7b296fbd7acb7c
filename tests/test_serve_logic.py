"""Host logic of the serving loop and the request generator, on the CPU with a fake device.

The fake ensemble finishes each admitted solve after a fixed number of steps and enforces the
slot count and the memory budget, so admission order, backpressure, release and latency
accounting are exercised without a GPU.
"""
import numpy as np
import pytest

import exaca_amd as ea
from exaca_amd import decks, serve


def test_request_distribution(tmp_path):
    reqs = decks.process_map_requests(400, tmp_path, rate=50.0)
    sides = np.array([[r.nx, r.ny, r.nz] for r in reqs]).ravel()
    assert sides.min() >= 32 and sides.max() <= 256
    logs = np.log(sides)
    assert abs(logs.mean() - 0.5 * (np.log(32) + np.log(256))) < 0.1  # log-uniform centre
    for attr, lo, hi in [("G", 1e5, 1e7), ("R", 1e4, 1e7), ("density", 1e11, 1e14)]:
        v = np.array([getattr(r, attr) for r in reqs])
        assert v.min() >= lo and v.max() <= hi
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 1 / 50) < 0.004  # Poisson at 50/s
    assert len({r.seed for r in reqs}) == len(reqs)
    (tmp_path / "c").mkdir()
    cubes = decks.process_map_requests(20, tmp_path / "c", cubes=True)
    assert all(r.nx == r.ny == r.nz for r in cubes)
    d = ea.host_state(reqs[0].deck)
    assert (d["nx"], d["ny"], d["nz"]) == (reqs[0].nx + 2, reqs[0].ny + 2, reqs[0].nz + 2)


class FakePrepared:
    def __init__(self, deck, seed):
        self.deck, self.seed = deck, seed
        self.cells = 10
        self.device_bytes = 100
        self.closed = False

    def close(self):
        self.closed = True


class FakeEnsemble:
    log = []

    fail = set()
    forever = set()

    def __init__(self, slots, arena_bytes):
        self.slots, self.free = slots, arena_bytes
        self.occ = {}
        FakeEnsemble.log = []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        pass

    def admit(self, p):
        free = [s for s in range(self.slots) if s not in self.occ]
        if not free or self.free < p.device_bytes:
            return None
        s = free[0]
        self.free -= p.device_bytes
        self.occ[s] = [p, 120 + (p.seed % 3) * 50]  # steps left
        FakeEnsemble.log.append(("admit", p.seed, s))
        return s

    def advance(self, steps):
        for v in self.occ.values():
            if v[0].seed not in self.forever:
                v[1] -= steps

    def poll_errors(self):
        st = np.full(self.slots, -1, dtype=np.int32)
        err = np.zeros(self.slots, dtype=np.int32)
        for s, (p, left) in self.occ.items():
            st[s] = 1000 if left <= 0 else 0
            if p.seed in self.fail:  # the device stopped this domain with an error flag
                st[s], err[s] = 50, 8
        return st, err

    def autotune(self, k):
        return 32

    def field(self, slot, kind, n):
        p = self.occ[slot][0]
        return np.full(n, p.seed + kind, dtype=np.int32)

    def release(self, slot):
        p = self.occ.pop(slot)[0]
        self.free += p.device_bytes
        FakeEnsemble.log.append(("release", p.seed, slot))


def _requests(n):
    return [decks.Request(i, f"deck{i}", 1000 + i, 8, 8, 8, 1e6, 1e6, 1e13, arrival=0.0) for i in range(n)]


@pytest.mark.parametrize("slots,budget", [(2, 10_000), (8, 250)])
def test_continuous_batching_respects_slots_and_budget(monkeypatch, slots, budget):
    monkeypatch.setattr(serve, "Ensemble", FakeEnsemble)
    monkeypatch.setattr(serve, "Prepared", FakePrepared)
    reqs = _requests(7)
    out = serve.run_served(reqs, slots=slots, arena_bytes=budget, realtime=False, autotune=False)
    assert sorted(out["digests"]) == list(range(7))
    resident, peak = set(), 0
    for ev, seed, slot in FakeEnsemble.log:
        if ev == "admit":
            resident.add(seed)
        else:
            resident.discard(seed)
        peak = max(peak, len(resident))
    assert peak <= min(slots, budget // 100)
    admits = [seed for ev, seed, _ in FakeEnsemble.log if ev == "admit"]
    assert admits == sorted(admits)  # admitted in arrival order
    assert out["p99_s"] >= out["p50_s"] >= 0


def test_every_slot_can_be_resident(monkeypatch):
    """The host-init cap bounds requests prepared and waiting for admission only: with room for
    them, all slots fill even when the cap is far below the slot count."""
    monkeypatch.setattr(serve, "Ensemble", FakeEnsemble)
    monkeypatch.setattr(serve, "Prepared", FakePrepared)
    monkeypatch.setattr(FakeEnsemble, "forever", {1000 + i for i in range(12)})  # first 12 never end alone

    class Ens(FakeEnsemble):
        def advance(self, steps):
            super().advance(steps)
            if len(self.occ) == self.slots:  # all resident: let everything finish
                FakeEnsemble.forever = set()

    monkeypatch.setattr(serve, "Ensemble", Ens)
    out = serve.run_served(_requests(16), slots=12, arena_bytes=10_000, realtime=False, max_prepared=3,
                           autotune=True, tune_at=12)
    assert out["peak_resident"] == 12
    assert out["blocks_per_slot"] == 32  # the autotuner ran once the tuning point was reached
    assert sorted(out["digests"]) == list(range(16)) and not out["failed"]


def test_device_error_is_a_failed_request(monkeypatch):
    monkeypatch.setattr(serve, "Ensemble", FakeEnsemble)
    monkeypatch.setattr(serve, "Prepared", FakePrepared)
    monkeypatch.setattr(FakeEnsemble, "fail", {1002})
    out = serve.run_served(_requests(5), slots=2, arena_bytes=10_000, realtime=False, autotune=False)
    assert out["failed"] == [2]
    assert out["digests"][2] == "error:8"
    assert all(not out["digests"][k].startswith("error") for k in (0, 1, 3, 4))


def test_step_bound_ends_a_stuck_loop(monkeypatch):
    monkeypatch.setattr(serve, "Ensemble", FakeEnsemble)
    monkeypatch.setattr(serve, "Prepared", FakePrepared)
    monkeypatch.setattr(FakeEnsemble, "forever", {1001})
    with pytest.raises(RuntimeError, match="exceeded 5000 steps"):
        serve.run_served(_requests(3), slots=2, arena_bytes=10_000, realtime=False, autotune=False,
                         max_steps=5000)


def test_digest_identifies_results():
    a = serve.digest(5000, np.arange(10), np.arange(10) * 3)
    assert a == serve.digest(5000, np.arange(10), np.arange(10) * 3)
    assert a != serve.digest(5001, np.arange(10), np.arange(10) * 3)
    b = np.arange(10) * 3
    b[4] += 1
    assert a != serve.digest(5000, np.arange(10), b)


def test_ensemble_refuses_domains_beyond_one_rebuild_round(tmp_path):
    """An ensemble slot covers its domain's list rebuild in one round of workgroups (at most 128
    units of 4 x 65536 cells): a larger domain is refused at preparation, on the host."""
    from conftest import c_deck

    ok = ea.Prepared(c_deck(tmp_path / "ok.txt", 256, density=1))
    assert ok.cells == 260 * 260 * 258  # walls and the ghost ring in x, y; walls in z
    ok.close()
    with pytest.raises(ea.ExacaError, match="too large for an ensemble"):
        ea.Prepared(c_deck(tmp_path / "big.txt", 330, density=1))
