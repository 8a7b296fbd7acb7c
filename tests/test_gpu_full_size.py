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


# The oracle cannot step the 800^3 deck from its start in test time, but it can continue the GPU's
# own state: at cycle 6000 (the middle of the run, a fully developed front of 3.4 M active cells)
# and at cycle 9200 (the front at the domain's top plane, the run winding down) the GPU's cell
# types, GrainIDs, half-diagonals and double octahedron centres are imported into the oracle
# (which derives the undercooling and the critical lengths as the reference holds them) and both
# paths step 12 more cycles. Every cell's CellType and GrainID, and the active cells'
# DiagonalLength and DOCenter, must then agree bit for bit.
import oracle as orc  # noqa: E402
from conftest import DATA  # noqa: E402

MORE = 12
# (first cycle, least active cells then, least captures in the window): the steady front in the
# middle of the run, and the front reaching the domain's top plane (it gets there around cycle
# 8500-9000; the run ends at cycle 11 000)
WINDOWS = {"steady_front_6000": (6000, 1_000_000, 500_000), "front_at_top_9200": (9200, 100_000, 50_000)}


@pytest.mark.parametrize("window", list(WINDOWS))
def test_800_front_continues_exactly_on_the_oracle(window):
    START, min_active, min_captures = WINDOWS[window]
    print(f"[800^3] GPU to cycle {START}", flush=True)
    with ea.Solver(DECK) as sol:
        sol.run(max_cycles=START)
        st = sol.front_state()
        print(f"[800^3] GPU {MORE} more steps", flush=True)
        sol.run(max_cycles=START + MORE)
        q = sol.query()
        ct_g = sol.field(ea.FIELD_CELLTYPE)
        gid_g = sol.field(ea.FIELD_GRAINID)
        dl_g = sol.field(ea.FIELD_DIAGONAL)
        doc_g = sol.field(ea.FIELD_DOCENTER).reshape(-1, 3)
    active0 = int(np.count_nonzero((st["ct_raw"] & 0x0F) == ACTIVE))
    assert active0 > min_active
    print(f"[800^3] oracle import of {active0} active cells", flush=True)
    o = orc.OracleSim.from_state(DECK, DATA, START, st["ct_raw"], st["gid"], st["dl"], st["centre"], st["nloc"],
                                 st["ntime"])
    del st
    print(f"[800^3] oracle {MORE} steps", flush=True)
    cycle, _ = o.run(max_cycles=START + MORE)
    assert cycle == START + MORE
    d = o.dims()
    ct_o = o.field(orc.CELLTYPE)
    assert np.array_equal(ct_g, ct_o), f"CellType differs at {np.count_nonzero(ct_g != ct_o)} cells"
    assert np.array_equal(gid_g, o.field(orc.GRAINID)), "GrainID differs"
    act = np.flatnonzero(ct_o == ACTIVE)
    assert np.array_equal(dl_g[act].view(np.uint32), o.field(orc.DIAGLEN)[act].view(np.uint32))
    assert np.array_equal(doc_g[act].view(np.uint32), o.field(orc.DOCENTER).reshape(-1, 3)[act].view(np.uint32))
    assert d["captures"] > min_captures, "the window should hold the front's captures"
    print(f"[800^3] {d['captures']} captures, {d['solidified']} solidifications, {d['frontier']} active-cell updates "
          f"identical", flush=True)
    o.close()
