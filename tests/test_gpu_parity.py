"""GPU parity: the HIP path, called through the C ABI, against the oracle on the same decks.

Integer fields (CellType, GrainID), cycle counts and nucleation counts must match exactly; the
octahedron state (DiagonalLength, DOCenter) is compared bit for bit as float32.
"""
import os
import shutil
from pathlib import Path

import numpy as np
import pytest

import exaca_amd as ea
import oracle as orc
from conftest import DATA, EXAMPLES, c_deck, s_deck

pytestmark = pytest.mark.gpu

ACTIVE = 2


def _compare_state(sol, osim, tag, check_octahedra=True):
    ct_g = sol.field(ea.FIELD_CELLTYPE)
    ct_o = osim.field(orc.CELLTYPE)
    bad = np.flatnonzero(ct_g != ct_o)
    assert bad.size == 0, f"{tag}: CellType differs at {bad.size} cells, first {bad[:5]} gpu {ct_g[bad[:5]]} oracle {ct_o[bad[:5]]}"
    gid_g = sol.field(ea.FIELD_GRAINID)
    gid_o = osim.field(orc.GRAINID)
    bad = np.flatnonzero(gid_g != gid_o)
    assert bad.size == 0, f"{tag}: GrainID differs at {bad.size} cells, first {bad[:5]}"
    if check_octahedra:
        # the oracle holds octahedra for its active window [zlow, zlow + nzA) only
        act = np.flatnonzero(ct_o == ACTIVE)
        d = osim.dims()
        lo = d["X"] * d["Y"] * d["zlow"]
        act = act[(act >= lo) & (act < lo + d["X"] * d["Y"] * d["nzA"])]
        dl_g = sol.field(ea.FIELD_DIAGONAL)[act]
        dl_o = osim.field(orc.DIAGLEN)[act - lo]
        assert np.array_equal(dl_g.view(np.uint32), dl_o.view(np.uint32)), f"{tag}: DiagonalLength differs"
        c_g = sol.field(ea.FIELD_DOCENTER).reshape(-1, 3)[act]
        c_o = osim.field(orc.DOCENTER).reshape(-1, 3)[act - lo]
        assert np.array_equal(c_g.view(np.uint32), c_o.view(np.uint32)), f"{tag}: DOCenter differs"


def test_capture_geometry_bitwise(tmp_path):
    """New octahedron centre, diagonal and all 26 critical lengths for random captures."""
    deck = c_deck(tmp_path / "g.txt", 6, density=0)
    guv = np.loadtxt(DATA / "Substrate" / "GrainOrientationVectors_Robert.csv", delimiter=",", skiprows=1,
                     dtype=np.float64).astype(np.float32)
    rng = np.random.default_rng(7)
    n = 4000
    rows = np.zeros((n, 7))
    rows[:, 0] = rng.integers(0, 10000, n)
    cell = rng.integers(1, 200, (n, 3))
    jitter = rng.uniform(-0.7, 0.7, (n, 3))
    cold = (cell + 0.5 + jitter).astype(np.float32).astype(np.float64)
    nb = np.array([[0, -1, 0], [1, -1, 0], [-1, -1, 0], [0, -1, 1], [0, -1, -1], [-1, -1, 1], [1, -1, 1],
                   [-1, -1, -1], [1, -1, -1], [0, 0, 1], [0, 0, -1], [1, 0, 1], [-1, 0, 1], [1, 0, -1],
                   [-1, 0, -1], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [1, 1, 0], [-1, 1, 0], [0, 1, 1], [0, 1, -1],
                   [1, 1, 1], [-1, 1, 1], [1, 1, -1], [-1, 1, -1]])
    rows[:, 1:4] = cold
    rows[:, 4:7] = cell + nb[rng.integers(0, 26, n)] + 0.5
    rows[:8, 1:4] = cell[:8] + 0.5  # cell-centred capturers (nuclei / substrate)
    with ea.Solver(deck) as sol:
        out = sol.geometry_probe(rows)
    for r in range(n):
        o = int(rows[r, 0])
        centre, dl, cdl = orc.capture_octahedron(guv.reshape(-1, 9)[o], rows[r, 1:4], rows[r, 4:7])
        assert np.array_equal(out[r, 0:3], centre), f"row {r}: centre {out[r, 0:3]} vs {centre}"
        assert np.float32(out[r, 3]) == np.float32(dl), f"row {r}: diagonal"
        assert np.array_equal(out[r, 4:30].astype(np.float32), cdl), f"row {r}: critical lengths"


def test_stepwise_small_domain(tmp_path):
    """Every one of the first 400 steps of a dense-nucleation 14^3 domain, state compared each step;
    the reference's nn after every step's nucleation (exaca_nucleation_nn) against the oracle's."""
    deck = c_deck(tmp_path / "s.txt", 14, density=3000, fraction=0.05)
    osim = orc.OracleSim(deck, DATA)
    with ea.Solver(deck) as sol:
        assert sol.query()["nuclei"] == osim.dims()["nuclei"] > 0
        nn_seen = set()
        for cycle in range(1, 401):
            nn = sol.nucleation_nn(cycle)
            sol.cell_capture(cycle)
            osim.step(cycle)
            assert nn == osim.dims()["nn"], f"cycle {cycle}: nn {nn} != {osim.dims()['nn']}"
            nn_seen.add(nn)
            if cycle % 20 == 0 or cycle < 30:
                _compare_state(sol, osim, f"cycle {cycle}")
        assert len(nn_seen) > 3, "the deck must nucleate over several steps"
        q = sol.query()
        d = osim.dims()
        assert q["nn"] == d["nn"]
        assert q["frontier"] == d["frontier"]
        assert q["captures"] == d["captures"]
        assert q["solidified"] == d["solidified"]


CASES = {
    "dirs_50_substrate_only": lambda p: EXAMPLES / "Inp_SmallDirSolidification.txt",
    "dirs_100_baseline": lambda p: EXAMPLES / "Inp_DirSolidification_100.txt",
    "dirs_40_dense_nuclei": lambda p: c_deck(p / "d.txt", 40, density=400, fraction=0.02),
    "dirs_ragged_37x23x31": lambda p: c_deck(p / "r.txt", 31, nx=37, ny=23, density=200, fraction=0.1),
    # every cell on an x or y edge (the bounds mask on each neighbour scan and capture)
    "dirs_narrow_2x3x40": lambda p: c_deck(p / "n.txt", 40, nx=2, ny=3, density=2000, fraction=0.5),
    "dirs_slab_48x2x24": lambda p: c_deck(p / "b.txt", 24, nx=48, ny=2, density=500, fraction=0.2),
    # G = 1e5: 0.01 K per step, so the growth increments reach their plateau only after 21 004
    # steps, past the 4096 entries staged in LDS (the kernels read the rest from global memory)
    "dirs_slow_cooling_24": lambda p: c_deck(p / "g.txt", 24, density=100, fraction=0.1, G=1e5),
    # no spread in the nucleation undercooling: every site of a plane is due at the same step, and
    # adjacent nuclei take each other out of shared neighbours' masks within one launch
    "dirs_simultaneous_nuclei_20": lambda p: c_deck(p / "z.txt", 20, density=20000, fraction=0.05, dTsigma=0.0),
    "spots_2x2": lambda p: s_deck(p / "s.txt", nsx=2, nsy=2, offset=30, radius=20, density=60),
    "spots_1x2_jumps": lambda p: EXAMPLES / "Inp_SmallSpotMelt.txt",
    # multilayer: three layers 6 cells apart, each pool overlapping the one below (the window slides
    # up at each transition, active cells restart as cell-centred octahedra, undercooling carries on)
    "spots_3_layers": lambda p: s_deck(p / "l.txt", nsx=2, nsy=2, offset=16, radius=10, density=300, spacing=8,
                                       layers=3, layer_height=6),
    "spots_2_layers_jumps": lambda p: s_deck(p / "j.txt", nsx=2, nsy=1, offset=30, radius=14, density=100,
                                             spacing=10, layers=2, layer_height=8, G=2e5),
}


@pytest.mark.parametrize("case", list(CASES))
def test_full_run_matches_oracle(case, tmp_path):
    deck = CASES[case](tmp_path)
    osim = orc.OracleSim(deck, DATA)
    ocycle, _ = osim.run()
    with ea.Solver(deck) as sol:
        gcycle, _ = sol.run()
        q = sol.query()
        assert gcycle == ocycle, f"final cycle gpu {gcycle} oracle {ocycle}"
        d = osim.dims()
        assert q["nn"] == d["nn"]
        assert q["captures"] == d["captures"]
        assert q["solidified"] == d["solidified"]
        assert q["frontier"] == d["frontier"]
        _compare_state(sol, osim, case)


def test_quiet_steps_leave_nucleation_out(tmp_path, monkeypatch):
    """A single-layer run leaves the nucleation kernel out of every step (and 50-step graph) with
    no nucleation time after the last step it ran; the capture pass advances the step's control
    words and passes the nucleation gate on. The run must equal one that launches the kernel at
    every step, counters and fields, and the oracle."""
    deck = c_deck(tmp_path / "q.txt", 60, density=200, fraction=0.02)
    runs = []
    for every in (False, True):
        if every:
            monkeypatch.setenv("EXACA_NUCLEATE_EVERY_STEP", "1")
        with ea.Solver(deck) as sol:
            cycle, _ = sol.run()
            q = sol.query()
            runs.append((cycle, {k: q[k] for k in ("nn", "captures", "solidified", "frontier")},
                         sol.field(ea.FIELD_CELLTYPE), sol.field(ea.FIELD_GRAINID), sol.field(ea.FIELD_DIAGONAL)))
    (c0, q0, ct0, g0, d0), (c1, q1, ct1, g1, d1) = runs
    assert q0["nn"] > 3, "the deck must nucleate"
    assert c0 == c1 and q0 == q1
    assert np.array_equal(ct0, ct1) and np.array_equal(g0, g1)
    assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32))
    osim = orc.OracleSim(deck, DATA)
    ocycle, _ = osim.run()
    d = osim.dims()
    assert c0 == ocycle and q0["nn"] == d["nn"] and q0["captures"] == d["captures"]


FEW_BLOCK_CASES = {
    "dirs_40_dense_nuclei": CASES["dirs_40_dense_nuclei"],
    "spots_2x2": CASES["spots_2x2"],
    # every bottom cell a substrate grain and dense nuclei: thousands of fronts at once
    "dirs_72_dense": lambda p: c_deck(p / "f.txt", 72, density=1000, fraction=1.0),
}


@pytest.mark.parametrize("case", list(FEW_BLOCK_CASES))
def test_few_workgroups_many_rounds(tmp_path, monkeypatch, case):
    """With 8 frontier workgroups every slice holds hundreds of due cells and claimants, so the
    capture pass, the winners pass (two claimants per thread and round, compacted in place) and the
    apply pass all run several rounds per workgroup, and the list rebuild runs several units per
    workgroup. A full run must still equal the oracle."""
    monkeypatch.setenv("EXACA_BLOCKS", "8")
    deck = FEW_BLOCK_CASES[case](tmp_path)
    osim = orc.OracleSim(deck, DATA)
    ocycle, _ = osim.run()
    with ea.Solver(deck) as sol:
        gcycle, _ = sol.run()
        q = sol.query()
        d = osim.dims()
        assert gcycle == ocycle
        for key in ("nn", "captures", "solidified", "frontier"):
            assert q[key] == d[key], (key, q[key], d[key])
        _compare_state(sol, osim, f"{case} on 8 workgroups")


@pytest.mark.parametrize("stage", [1, 3])
def test_candidate_overflow_path(tmp_path, monkeypatch, stage):
    """The apply pass stages a step's solidification candidates in LDS and sends the waves that do
    not fit to a global overflow list. With the stage narrowed to a few entries almost every
    candidate takes the overflow path: a full run must still equal the oracle (dense nucleation:
    many candidates per workgroup and step)."""
    monkeypatch.setenv("EXACA_CAND_STAGE", str(stage))
    deck = c_deck(tmp_path / "ov.txt", 40, density=400, fraction=0.02)
    osim = orc.OracleSim(deck, DATA)
    ocycle, _ = osim.run()
    with ea.Solver(deck) as sol:
        gcycle, _ = sol.run()
        q = sol.query()
        d = osim.dims()
        assert gcycle == ocycle
        for key in ("nn", "captures", "solidified", "frontier"):
            assert q[key] == d[key], (key, q[key], d[key])
        _compare_state(sol, osim, f"candidate stage {stage}")


@pytest.mark.parametrize("fixture", ["spot_array_6x6.json", "spot_melt_2_layers.json"])
def test_spot_decks_match_oracle_fixtures(fixture):
    """The spot-melt example decks: examples/Inp_SpotMeltArray_6x6.txt (the overlapping-spot
    workload: 36 hemispherical spots of radius 75 um on a 6 x 6 grid at 60 um pitch, melted in raster
    order, arrays 456 x 456 x 78, about 1e5 steps) and examples/Inp_SpotMelt.txt (3 x 2 spots, two
    layers). The serial oracle needs most of an hour and 13 minutes for them, so its results are
    committed fixtures (tests/golden/make_spot_golden.py): final cycle, step counters and digests of
    the final CellType and GrainID fields."""
    import hashlib
    import json

    gold = json.loads((Path(__file__).resolve().parent / "golden" / fixture).read_text())
    with ea.Solver(EXAMPLES / gold["deck"]) as sol:
        cycle, _ = sol.run()
        q = sol.query()
        ct = np.ascontiguousarray(sol.field(ea.FIELD_CELLTYPE), dtype=np.int32)
        gid = np.ascontiguousarray(sol.field(ea.FIELD_GRAINID), dtype=np.int32)
    assert [q["X"], q["Y"], q["nz"]] == gold["dims"]
    assert cycle == gold["final_cycle"]
    for key in ("nn", "frontier", "captures", "solidified"):
        assert q[key] == gold[key], (key, q[key], gold[key])
    assert hashlib.sha1(ct.tobytes()).hexdigest() == gold["celltype_sha1"]
    assert hashlib.sha1(gid.tobytes()).hexdigest() == gold["grainid_sha1"]


def test_drop_in_outputs_identical(tmp_path):
    """RunProgram_Reduced drop-in: vtk files byte-identical to the oracle's, log identical apart
    from the timing lines."""
    deck = c_deck(tmp_path / "io.txt", 30, density=300, fraction=0.05, printing="Y", base="io",
                  out=str(tmp_path / "gpu") + "/")
    (tmp_path / "gpu").mkdir()
    ea.run_program(deck)
    odeck = c_deck(tmp_path / "io_oracle.txt", 30, density=300, fraction=0.05, printing="Y", base="io",
                   out=str(tmp_path / "cpu") + "/")
    (tmp_path / "cpu").mkdir()
    osim = orc.OracleSim(odeck, DATA)
    osim.run()
    osim.write_outputs()
    for name in ["io.vtk", "io_Misorientations.vtk"]:
        a = (tmp_path / "gpu" / name).read_bytes()
        b = (tmp_path / "cpu" / name).read_bytes()
        assert a == b, f"{name} differs"
    la = (tmp_path / "gpu" / "io.log").read_text().splitlines()
    lb = (tmp_path / "cpu" / "io.log").read_text().splitlines()
    assert len(la) == len(lb)
    for x, y in zip(la, lb):
        if "time" in x or "took" in x or "input file" in x:
            continue
        assert x == y


def test_full_width_800_sample_matches_oracle(tmp_path):
    """The 800^3 benchmark deck cut to its bottom 40 planes (804 x 804 x 42 cells: the full
    benchmark width, and the same substrate, nuclei and times, since the random streams are drawn
    plane by plane), 300 steps on both paths: every cell's CellType and GrainID, the active cells'
    octahedra, and the step counters must agree."""
    lines = (EXAMPLES / "Inp_DirSolidification_800.txt").read_text().splitlines()
    lines = [("Domain size in z: 40" if ln.startswith("Domain size in z") else ln) for ln in lines]
    deck = tmp_path / "cut.txt"
    deck.write_text("\n".join(lines) + "\n")
    osim = orc.OracleSim(deck, DATA)
    ocycle, _ = osim.run(max_cycles=300)
    with ea.Solver(deck) as sol:
        gcycle, _ = sol.run(max_cycles=300)
        q = sol.query()
        d = osim.dims()
        assert (gcycle, ocycle) == (300, 300)
        assert (q["X"], q["Y"], q["nz"]) == (d["X"], d["Y"], d["nz"]) == (804, 804, 42)
        for key in ("nn", "captures", "solidified", "frontier"):
            assert q[key] == d[key], (key, q[key], d[key])
        assert d["captures"] > 100000, "the sample should exercise capture at scale"
        _compare_state(sol, osim, "800 cut to 40 planes")
    osim.close()
