"""The benchmark's full 800^3 deck (804 x 804 x 802 cells) on the GPU, checked through properties
that hold at any size, since the oracle cannot step a domain this large:

* conservation: cells leave the liquid only by capture or nucleation, so the liquid count drops by
  exactly captures + nucleation events; solid cells appear only through the solidification
  counter; active cells are what is left;
* no cell returns to the liquid, walls never change, and every cell that left the liquid carries a
  nonzero GrainID;
* determinism: a second run of the same deck gives the same CellType and GrainID bytes (the
  claim resolution makes the result independent of the order in which lanes run).

The identities were checked against the oracle on the 100^3 and 40^3 dense-nucleation decks
(`captures + nn` equals the drop of the liquid count, the solid count equals `solidified`).
"""
import hashlib

import numpy as np
import pytest

import exaca_amd as ea
from conftest import EXAMPLES

pytestmark = pytest.mark.gpu

WALL, SOLID, ACTIVE, LIQUID = 0, 1, 2, 5
DECK = EXAMPLES / "Inp_DirSolidification_800.txt"
CYCLES = 3000


def _run(deck):
    with ea.Solver(deck) as sol:
        cycle, _ = sol.run(max_cycles=CYCLES)
        q = sol.query()
        ct = sol.field(ea.FIELD_CELLTYPE)
        gid = sol.field(ea.FIELD_GRAINID)
    return cycle, q, ct, gid


def test_800_full_size_conservation_and_determinism():
    print(f"[800^3] initial state of {DECK.name}", flush=True)
    _, ct0 = ea.host_state(DECK, ea.FIELD_CELLTYPE)
    print(f"[800^3] first run, {CYCLES} steps", flush=True)
    cycle, q, ct1, gid1 = _run(DECK)
    assert cycle == CYCLES
    assert (q["X"], q["Y"], q["nz"]) == (804, 804, 802)
    print(f"[800^3] captures {q['captures']} nucleation events {q['nn']} solidified {q['solidified']} "
          f"active-cell updates {q['frontier']}", flush=True)
    n = lambda a, v: int(np.count_nonzero(a == v))
    liquid0, liquid1 = n(ct0, LIQUID), n(ct1, LIQUID)
    assert liquid0 - liquid1 == q["captures"] + q["nn"]
    assert n(ct1, SOLID) - n(ct0, SOLID) == q["solidified"]
    assert n(ct1, ACTIVE) - n(ct0, ACTIVE) == q["captures"] + q["nn"] - q["solidified"]
    assert q["captures"] > 1_000_000 and q["solidified"] > 100_000, "the run should be well into the build"
    assert np.array_equal(ct0 == WALL, ct1 == WALL)
    assert np.all(ct0[ct1 == LIQUID] == LIQUID)
    left = (ct0 == LIQUID) & (ct1 != LIQUID)
    assert np.all(gid1[left] != 0)
    digest = (hashlib.sha1(ct1.tobytes()).hexdigest(), hashlib.sha1(gid1.tobytes()).hexdigest())
    del ct0, left
    print("[800^3] second run", flush=True)
    cycle2, q2, ct2, gid2 = _run(DECK)
    assert cycle2 == CYCLES
    for key in ("nn", "captures", "solidified", "frontier"):
        assert q2[key] == q[key], key
    assert (hashlib.sha1(ct2.tobytes()).hexdigest(), hashlib.sha1(gid2.tobytes()).hexdigest()) == digest
