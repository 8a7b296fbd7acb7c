"""GPU: batched ensembles and continuous batching give every domain exactly its plain-path result.

Each domain (its own size, G, R, nucleation density and random seed) is run alone through the
plain path (exaca_create_seeded + exaca_run_layer), then inside an ensemble; CellType, GrainID,
final cycle and counters must be identical. The seed-0 domain is also the oracle's own case.
"""
import numpy as np
import pytest

import exaca_amd as ea
import oracle as orc
from conftest import DATA, c_deck

pytestmark = pytest.mark.gpu

CASES = [  # n, nx, ny, G, R, density, fraction, seed
    (16, 16, 16, 1e6, 1e6, 2000, 0.05, 0),
    (20, 14, 26, 3e6, 1e5, 800, 0.08, 11),
    (12, 12, 12, 2e5, 2e6, 4000, 0.10, 12),
    (24, 24, 24, 1e7, 1e6, 300, 0.04, 13),
    (18, 22, 9, 5e5, 5e4, 1500, 0.06, 14),
    (10, 30, 30, 1e6, 3e6, 0, 0.03, 15),
]


def _decks(tmp_path):
    out = []
    for k, (n, nx, ny, G, R, dens, frac, seed) in enumerate(CASES):
        deck = c_deck(tmp_path / f"e{k}.txt", n, nx=nx, ny=ny, G=G, R=R, density=dens, fraction=frac)
        out.append((deck, seed))
    return out


def _alone(deck, seed):
    with ea.Solver(deck, seed=seed) as sol:
        cycle, _ = sol.run()
        return cycle, sol.field(ea.FIELD_CELLTYPE), sol.field(ea.FIELD_GRAINID), sol.query()


def _same(slot_result, alone, tag):
    cyc, ct, gid, q = slot_result
    acyc, act, agid, aq = alone
    assert cyc == acyc, f"{tag}: final cycle {cyc} vs alone {acyc}"
    assert np.array_equal(ct, act), f"{tag}: CellType differs"
    assert np.array_equal(gid, agid), f"{tag}: GrainID differs"
    assert q["captures"] == aq["captures"] and q["nn"] == aq["nn"] and q["frontier"] == aq["frontier"], tag


def test_seed_zero_plain_path_is_the_oracle_case(tmp_path):
    deck, seed = _decks(tmp_path)[0]
    assert seed == 0
    cycle, ct, gid, _ = _alone(deck, 0)
    o = orc.OracleSim(deck, DATA)
    ocycle, _ = o.run()
    assert cycle == ocycle
    assert np.array_equal(ct, o.field(orc.CELLTYPE)) and np.array_equal(gid, o.field(orc.GRAINID))


@pytest.mark.parametrize("blocks,stage,tune", [(16, None, False), (64, None, False), (8, 2, False),
                                              (64, None, True)])
def test_batched_ensemble_matches_each_domain_alone(tmp_path, monkeypatch, blocks, stage, tune):
    """stage: the apply pass's LDS candidate stage narrowed to that many entries (its overflow
    list then carries almost every candidate); tune: the online autotuner picks the workgroups
    per slot on the live ensemble after the first chunk."""
    if stage:
        monkeypatch.setenv("EXACA_CAND_STAGE", str(stage))
    decks = _decks(tmp_path)
    alone = [_alone(d, s) for d, s in decks]
    preps = [ea.Prepared(d, seed=s) for d, s in decks]
    with ea.Ensemble(slots=8, arena_bytes=1 << 30) as ens:
        ens.set_blocks(blocks)
        slots = [ens.admit(p) for p in preps]
        assert None not in slots
        for it in range(10000):
            ens.advance(50)
            if tune and it == 0:
                assert ens.autotune(1) in (16, 32, 64, 128, 256)
            st, err = ens.poll_errors()
            assert not err.any(), err
            if all(st[s] > 0 for s in slots):
                break
        for k, s in enumerate(slots):
            q = ens.query(s)
            assert q["error"] == 0
            res = (q["final_cycle"], ens.field(s, ea.FIELD_CELLTYPE, preps[k].cells),
                   ens.field(s, ea.FIELD_GRAINID, preps[k].cells), q)
            _same(res, alone[k], f"domain {k}")


def test_continuous_batching_admits_into_freed_slots(tmp_path):
    """Three slots, six domains: new domains enter as others finish; each result equals its own
    plain-path run, whatever it shared the GPU with."""
    decks = _decks(tmp_path)
    alone = [_alone(d, s) for d, s in decks]
    preps = [ea.Prepared(d, seed=s) for d, s in decks]
    pending = list(range(len(preps)))
    running = {}
    results = {}
    with ea.Ensemble(slots=3, arena_bytes=1 << 30) as ens:
        ens.set_blocks(32)
        for _ in range(20000):
            while pending:
                slot = ens.admit(preps[pending[0]])
                if slot is None:
                    break
                running[slot] = pending.pop(0)
            if not running:
                break
            ens.advance(50)
            st = ens.poll()
            for slot in [s for s in running if st[s] > 0]:
                k = running.pop(slot)
                q = ens.query(slot)
                results[k] = (q["final_cycle"], ens.field(slot, ea.FIELD_CELLTYPE, preps[k].cells),
                              ens.field(slot, ea.FIELD_GRAINID, preps[k].cells), q)
                ens.release(slot)
    assert sorted(results) == list(range(len(preps)))
    for k in results:
        _same(results[k], alone[k], f"domain {k}")


def test_admission_respects_the_memory_budget(tmp_path):
    decks = _decks(tmp_path)
    preps = [ea.Prepared(d, seed=s) for d, s in decks]
    need = preps[0].device_bytes
    with ea.Ensemble(slots=8, arena_bytes=need + need // 2) as ens:
        assert ens.admit(preps[0]) is not None
        assert ens.admit(preps[0]) is None  # no room for a second copy


@pytest.mark.parametrize("k", range(1, len(CASES)))
def test_seeded_instance_matches_the_oracle(tmp_path, k):
    """An ensemble instance's own random streams (mt19937_64(seed) for the substrate and the sites,
    default_random_engine(seed % 2147483646 + 1) for the undercoolings): the plain path with that
    seed equals the oracle initialised with the same streams, cell for cell."""
    deck, seed = _decks(tmp_path)[k]
    assert seed != 0
    cycle, ct, gid, q = _alone(deck, seed)
    o = orc.OracleSim(deck, DATA, seed=seed)
    ocycle, _ = o.run()
    d = o.dims()
    assert cycle == ocycle
    assert (q["nn"], q["captures"], q["solidified"], q["frontier"]) == (d["nn"], d["captures"], d["solidified"],
                                                                      d["frontier"])
    assert np.array_equal(ct, o.field(orc.CELLTYPE)) and np.array_equal(gid, o.field(orc.GRAINID))


def test_process_map_ensemble_and_serving_256(tmp_path):
    """The batched-ensemble and serving workloads in full: 256 process-map requests (sides
    log-uniform in [32, 256], G, R and nucleation density log-uniform, one seed each). Solved all
    at once (every instance resident together, the autotuner choosing workgroups per slot on the
    live mix), served as they arrive (Poisson 50 req/s, continuous batching under a 200 GiB
    budget) and one after another on the plain path: every request's digest (final cycle,
    CellType, GrainID) must be the same three times. The four smallest instances are also run on
    the oracle with their own seeds."""
    from exaca_amd import decks, serve

    reqs = decks.process_map_requests(256, tmp_path, rate=50.0)
    print("[ensemble] 256 requests, batched", flush=True)
    batched = serve.run_served(reqs, realtime=False)
    assert batched["peak_resident"] == 256 and not batched["failed"]
    assert batched["blocks_per_slot"] in (16, 32, 64, 128, 256)
    print("[ensemble] served, Poisson 50 req/s", flush=True)
    served = serve.run_served(reqs, realtime=True)
    assert not served["failed"]
    print("[ensemble] sequential plain path", flush=True)
    seq = serve.run_sequential(reqs)
    for r in reqs:
        assert batched["digests"][r.index] == seq["digests"][r.index] == served["digests"][r.index], r.index
    for r in sorted(reqs, key=lambda r: r.nx * r.ny * r.nz)[:4]:
        print(f"[ensemble] oracle check of request {r.index} ({r.nx}x{r.ny}x{r.nz}, seed {r.seed})", flush=True)
        o = orc.OracleSim(r.deck, DATA, seed=r.seed)
        ocycle, _ = o.run()
        assert serve.digest(ocycle, o.field(orc.CELLTYPE), o.field(orc.GRAINID)) == seq["digests"][r.index]
        o.close()
