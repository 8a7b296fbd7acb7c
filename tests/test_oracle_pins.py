"""Pins of the oracle (the CPU restatement used as the parity checker).

Each check is chosen so that a plausible restatement mistake (wrong stream, dropped term, wrong
sign or index, transposed operand) fails it: libstdc++'s own draws, closed forms the geometry
reduces to, the growth law's clamps, and invariants of whole runs.
"""
import json
import math
import shutil
from pathlib import Path

import numpy as np
import pytest

import oracle as orc
from conftest import DATA, EXAMPLES, REPO, c_deck, s_deck

GOLDEN = json.loads((REPO / "tests" / "golden" / "rng_golden.json").read_text())

NB = np.array([[0, -1, 0], [1, -1, 0], [-1, -1, 0], [0, -1, 1], [0, -1, -1], [-1, -1, 1], [1, -1, 1], [-1, -1, -1],
               [1, -1, -1], [0, 0, 1], [0, 0, -1], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1], [1, 0, 0],
               [-1, 0, 0], [0, 1, 0], [1, 1, 0], [-1, 1, 0], [0, 1, 1], [0, 1, -1], [1, 1, 1], [-1, 1, 1],
               [1, 1, -1], [-1, 1, -1]], dtype=np.float64)
IDENTITY = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float32)


def orientations():
    return np.loadtxt(DATA / "Substrate" / "GrainOrientationVectors_Robert.csv", delimiter=",", skiprows=1,
                      dtype=np.float64).astype(np.float32)


# ---------------------------------------------------------------- random streams (libstdc++)
@pytest.mark.parametrize("seed", [0, 1])
def test_mt19937_64_matches_libstdcxx(seed):
    want = np.array([int(x) for x in GOLDEN[f"mt64_raw_seed{seed}"]], dtype=np.uint64)
    assert np.array_equal(orc.mt64(seed, len(want)), want)


@pytest.mark.parametrize("seed", [0, 1])
def test_uniform_real_matches_libstdcxx(seed):
    want = np.array([float.fromhex(x) for x in GOLDEN[f"uniform_seed{seed}"]])
    got = orc.uniform(seed, len(want))
    assert np.array_equal(got, want)


def test_minstd_and_normal_match_libstdcxx():
    assert np.array_equal(orc.minstd(16), np.array(GOLDEN["minstd_raw"], dtype=np.uint32))
    want = np.array([float.fromhex(x) for x in GOLDEN["normal_5_0.5"]])
    assert np.array_equal(orc.normal(5.0, 0.5, len(want)), want)


# ---------------------------------------------------------------- octahedron geometry
def test_seed_lengths_identity_are_l1_norms():
    """A cell-centred octahedron aligned with the axes reaches neighbour (dx,dy,dz) at half-diagonal
    |dx|+|dy|+|dz| exactly: 1 for faces, 2 for edges, 3 for corners."""
    got = orc.seed_cdl(IDENTITY)
    assert np.array_equal(got, np.abs(NB).sum(axis=1).astype(np.float32))


def test_seed_lengths_rotated_are_rotated_l1_norms():
    """For an orthonormal frame d1, d2, d3 the face hit by the ray x0 has normal s1 d1 + s2 d2 + s3 d3,
    so the critical half-diagonal is sum_i |d_i . x0|."""
    g = orientations()
    for o in range(0, 10000, 97):
        d = g[o].reshape(3, 3).astype(np.float64)
        want = np.abs(NB @ d.T).sum(axis=1)
        assert np.allclose(orc.seed_cdl(g[o]), want, rtol=2e-6, atol=2e-6), o


def test_capture_face_neighbour_identity():
    """Axis-aligned seed at (0.5,0.5,0.5) capturing (1.5,0.5,0.5): the capturing corner is the new
    cell centre itself, so the new octahedron has half-diagonal sqrt(2)*sqrt(2)/2 = 1 and keeps
    the old centre."""
    c, dl, cdl = orc.capture_octahedron(IDENTITY, (0.5, 0.5, 0.5), (1.5, 0.5, 0.5))
    assert abs(dl - 1.0) < 1e-6
    assert np.allclose(c, (0.5, 0.5, 0.5), atol=1e-6)
    # critical lengths measured from the old centre: |1 + dx| + |dy| + |dz|
    want = np.abs(NB + np.array([1.0, 0, 0])).sum(axis=1)
    assert np.allclose(cdl, want, atol=1e-5)


def test_capture_corner_neighbour_identity():
    """Corner capture (1.5,1.5,1.5) from (0.5,0.5,0.5): the face triangle is (3.5,.5,.5),
    (.5,3.5,.5), (.5,.5,3.5), all sqrt(6) from the new centre; both edge projections exceed
    sqrt(3), so the new half-diagonal is sqrt(2)*sqrt(3) = sqrt(6) and the centre moves from the
    first corner back along +x by sqrt(6)."""
    c, dl, cdl = orc.capture_octahedron(IDENTITY, (0.5, 0.5, 0.5), (1.5, 1.5, 1.5))
    assert abs(dl - math.sqrt(6.0)) < 1e-6
    assert np.allclose(c, (3.5 - math.sqrt(6.0), 0.5, 0.5), atol=1e-6)


def test_capture_lengths_are_rotated_l1_from_new_centre():
    """The capture rule's max over the four face normals d1 +/- d2 +/- d3 equals sum_i |d_i . v|."""
    g = orientations()
    rng = np.random.default_rng(3)
    for _ in range(300):
        o = int(rng.integers(0, 10000))
        cold = (rng.integers(1, 50, 3) + 0.5 + rng.uniform(-0.6, 0.6, 3)).astype(np.float32).astype(np.float64)
        p = np.floor(cold) + NB[rng.integers(0, 26)] + 0.5
        c, dl, cdl = orc.capture_octahedron(g[o], cold, p)
        d = g[o].reshape(3, 3).astype(np.float64)
        want = np.abs((p + NB - c) @ d.T).sum(axis=1)
        assert np.allclose(cdl, want, rtol=1e-5, atol=1e-5)
        assert 0 < dl <= math.sqrt(6.0) + 1e-6  # L12, L13 <= sqrt(3)


# ---------------------------------------------------------------- growth law
def _baseline_poly():
    # Inconel625_Baseline normalised by deltat/deltax = 0.1 (G = R = 1e6, N = 10, dx = 1 um)
    return -1.0302e-7 * 0.1, 1.0533e-4 * 0.1, 2.2186e-3 * 0.1, 0.0


def test_growth_clamps_and_saturation():
    A, B, C, D = _baseline_poly()
    assert orc.growth_increment(A, B, C, D, 0.0) == 0.0
    assert orc.growth_increment(A, B, C, D, -3.0) == 0.0  # max(0, V) for a negative polynomial
    # the cap binds from about 57 K undercooling onwards; U is clamped at 210 K
    assert orc.growth_increment(A, B, C, D, 210.0) == 0.045
    assert orc.growth_increment(A, B, C, D, 500.0) == 0.045
    lo, hi = 0.0, 210.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if orc.growth_increment(A, B, C, D, mid) < 0.045:
            lo = mid
        else:
            hi = mid
    assert 56.0 < lo < 58.5
    # below the cap the increment increases with undercooling
    us = np.linspace(0.5, 55, 200)
    v = [orc.growth_increment(A, B, C, D, float(u)) for u in us]
    assert all(b > a for a, b in zip(v, v[1:]))


# ---------------------------------------------------------------- whole runs
def _identity_data_dir(tmp: Path) -> Path:
    dd = tmp / "data"
    (dd / "Materials").mkdir(parents=True)
    (dd / "Substrate").mkdir()
    shutil.copy(DATA / "Materials" / "Inconel625_Baseline", dd / "Materials")
    (dd / "Substrate" / "GrainOrientationVectors_Robert.csv").write_text(
        "10000\n" + "1,0,0,0,1,0,0,0,1\n" * 10000)
    return dd


def test_first_capture_time_of_axis_aligned_seeds(tmp_path):
    """Substrate seeds (liquidus time 0) grow from cycle 1 by the growth law; an axis-aligned
    octahedron first reaches its +z face neighbour (critical length 1) at the first cycle where
    the float-accumulated half-diagonal 0.01 + sum of increments reaches 1."""
    dd = _identity_data_dir(tmp_path)
    deck = c_deck(tmp_path / "id.txt", 12, density=0, fraction=0.03)
    sim = orc.OracleSim(deck, dd)
    p = sim.params()
    dl = np.float32(0.01)
    u = np.float32(0.0)
    t1 = None
    for cycle in range(1, 2000):
        u = np.float32(u + np.float32(p["dU"]))
        dl = np.float32(np.float64(dl) + orc.growth_increment(p["A"], p["B"], p["C"], p["D"], float(u)))
        if dl >= 1.0:
            t1 = cycle
            break
    assert t1 is not None
    d = sim.dims()
    X, Y = d["X"], d["Y"]
    ct0 = sim.field(orc.CELLTYPE).reshape(d["nz"], X, Y)
    gid0 = sim.field(orc.GRAINID).reshape(d["nz"], X, Y)
    seeds = np.argwhere(ct0 == 2)
    assert len(seeds) > 0
    for cycle in range(1, t1):
        sim.step(cycle)
    assert sim.dims()["captures"] == 0
    sim.step(t1)
    assert sim.dims()["captures"] >= len(seeds)
    gid = sim.field(orc.GRAINID).reshape(d["nz"], X, Y)
    ct = sim.field(orc.CELLTYPE).reshape(d["nz"], X, Y)
    for z, x, y in seeds:
        assert ct[z + 1, x, y] == 2 and gid[z + 1, x, y] == gid0[z, x, y]


def _grain_invariants(sim):
    d = sim.dims()
    shape = (d["nz"], d["X"], d["Y"])
    ct = sim.field(orc.CELLTYPE).reshape(shape)
    gid = sim.field(orc.GRAINID).reshape(shape)
    return d, ct, gid


@pytest.mark.parametrize("kind", ["C", "S"])
def test_completed_run_invariants(kind, tmp_path):
    """After a completed run: no liquid is left in the layer, every grain ID of a formerly liquid
    cell is found on one of its 26 neighbours or is its own nucleus ID, nucleated grains have
    negative IDs, and the nucleation count equals the number of nucleus IDs present at their sites."""
    if kind == "C":
        deck = c_deck(tmp_path / "c.txt", 24, density=800, fraction=0.05)
    else:
        deck = s_deck(tmp_path / "s.txt", nsx=2, nsy=1, offset=20, radius=12, density=800)
    sim = orc.OracleSim(deck, DATA)
    init_gid = sim.field(orc.GRAINID).copy()
    init_ct = sim.field(orc.CELLTYPE).copy()
    lid = sim.field(orc.LAYERID)
    nloc = sim.field(orc.NUCLEI_LOC)
    sim.run()
    d, ct, gid = _grain_invariants(sim)
    flat_ct = ct.ravel()
    assert not np.any((flat_ct == 5) & (lid == 0)), "liquid left in the layer"
    nz, X, Y = ct.shape
    was_liquid = (init_ct == 5).reshape(ct.shape)
    pad = np.pad(gid, 1, constant_values=0)
    for z, x, y in np.argwhere(was_liquid):
        g = gid[z, x, y]
        if g == init_gid.reshape(ct.shape)[z, x, y] and g < 0:
            continue  # a nucleus that nucleated keeps its own ID
        neigh = pad[z:z + 3, x:x + 3, y:y + 3].ravel()
        assert np.count_nonzero(neigh == g) >= 2, (z, x, y, g)  # itself plus a neighbour
    nucleated = sum(1 for q in nloc if gid.ravel()[q] == init_gid[q] and init_gid[q] < 0)
    assert nucleated == d["nn"]
    assert np.all(gid.ravel()[init_ct == 5] != 0)


def test_spot_melt_jump_and_termination():
    """The 1x2 spot deck has a >5000-step gap between its spots: the check at cycle 3000 finds no
    undercooled liquid and must move to the step before the earliest liquidus crossing of the
    liquid that is left; the run may only end on a check cycle with every spot cell solid."""
    sim = orc.OracleSim(EXAMPLES / "Inp_SmallSpotMelt.txt", DATA)
    cts = sim.field(orc.CRITTIMESTEP)
    lid = sim.field(orc.LAYERID)
    sim.run(max_cycles=2999)
    sim.step(3000)
    ct = sim.field(orc.CELLTYPE)
    liquid = (ct == 5) & (lid == 0)
    assert liquid.any() and np.all(cts[liquid] > 3000)
    nxt = int(cts[liquid].min())
    assert nxt - 3000 > 5000
    cycle, counts, xswitch = sim.check(3000)
    assert counts[1] == 0 and xswitch == 0 and cycle == nxt - 1
    cycle, _ = sim.run()
    ct = sim.field(orc.CELLTYPE)
    assert not np.any((ct == 5) & (lid == 0))
    assert cycle % 1000 == 0 and cycle >= cts[lid == 0].max()


@pytest.mark.parametrize("kind", ["dirs_40_dense", "spots_small"])
def test_counters_conserve_cells(kind, tmp_path):
    """Cells leave the liquid only by capture or nucleation and become solid only through the
    solidification counter, so the counters must account for every change of the CellType field
    (a double-counted capture or a solidification counted twice fails this), part way through a
    run and at its end."""
    deck = (c_deck(tmp_path / "d.txt", 40, density=400, fraction=0.02) if kind == "dirs_40_dense"
            else EXAMPLES / "Inp_SmallSpotMelt.txt")
    sim = orc.OracleSim(deck, DATA)
    ct0 = sim.field(orc.CELLTYPE).copy()
    for stop in (300, 0):
        sim.run(max_cycles=stop)
        ct = sim.field(orc.CELLTYPE)
        d = sim.dims()
        n = lambda a, v: int(np.count_nonzero(a == v))
        assert n(ct0, 5) - n(ct, 5) == d["captures"] + d["nn"], stop
        assert n(ct, 1) - n(ct0, 1) == d["solidified"], stop
        assert n(ct, 2) - n(ct0, 2) == d["captures"] + d["nn"] - d["solidified"], stop
        assert np.all(ct0[ct == 5] == 5) and np.array_equal(ct0 == 0, ct == 0), stop
    sim.close()


# ---------------------------------------------------------------- capture order (Serial semantics)
# Hand-built states on a 6^3 deck (arrays 10 x 10 x 8): every cell solid except the few placed,
# all past their liquidus time with a constant undercooling of 100 K (growth on the 0.045 cap),
# octahedra cell-centred with the identity orientation, so that a cell of half-diagonal d reaches
# its face neighbours at d >= 1 and nothing else below d = 2. The expected fields are derived by
# hand from the reference's pass 2 (src/CAupdate.cpp:245-726, Serial backend): the steering
# vector is visited in ascending index order (:217-240); each neighbour that is liquid at the
# visit sets LCount (:333-334) and, when reached, is captured at once by compare-exchange
# (:336-352), so a later cell no longer sees it liquid; a cell with LCount 0 turns solid
# (:722-725); a cell captured in this step is not in this step's steering vector (built in pass
# 1), so it does not grow until the next step.
WALL, SOLID, ACTIVE, LIQUID = 0, 1, 2, 5
GROWN = np.float32(np.float64(1.0) + 0.045)  # float(1 + 0.045): one step of growth on the cap


def _pin_state(tmp_path, cells, tag="pin"):
    """cells: {(x, y, z): (type, grain, half-diagonal)} on a state with everything else solid.
    Returns the oracle and a function giving the linear index of (x, y, z)."""
    deck = c_deck(tmp_path / f"{tag}.txt", 6, density=0, fraction=0.0)
    o = orc.OracleSim(deck, DATA)
    d = o.dims()
    X, Y, nz = d["X"], d["Y"], d["nz"]
    idx = lambda x, y, z: (z * X + x) * Y + y
    ct = o.field(orc.CELLTYPE)
    ct[ct != WALL] = SOLID
    gid = np.zeros_like(ct)
    dl = np.zeros(X * Y * d["nzA"], dtype=np.float32)
    doc = np.zeros(3 * X * Y * d["nzA"], dtype=np.float32)
    cdl = np.zeros(26 * X * Y * d["nzA"], dtype=np.float32)
    seed = np.abs(NB).sum(axis=1).astype(np.float32)  # identity orientation: L1 norms
    for (x, y, z), (t, g, h) in cells.items():
        p = idx(x, y, z)
        assert ct[p] == SOLID, "placed cells lie inside the walls"
        ct[p], gid[p], dl[p] = t, g, h
        doc[3 * p:3 * p + 3] = (x - 1 + 0.5, y - 1 + 0.5, z + 0.5)  # cell centre (offset -1 in x, y)
        cdl[26 * p:26 * p + 26] = seed
    o.set_orientation(0, IDENTITY)  # GrainIDs 1 and 10001 both use orientation 0
    o.set_field(orc.CELLTYPE, ct)
    o.set_field(orc.GRAINID, gid)
    o.set_field(orc.CRITTIMESTEP, np.full_like(ct, -100))
    o.set_field(orc.UNDERCOOLING, np.full(ct.size, 100.0, dtype=np.float32))
    o.set_field(orc.UCHANGE, np.zeros(ct.size, dtype=np.float32))
    o.set_field(orc.DIAGLEN, dl)
    o.set_field(orc.DOCENTER, doc)
    o.set_field(orc.CRITDIAG, cdl)
    return o, idx


def _contest(tmp_path, mutation=orc.MUT_NONE):
    # A (grain 1) and B (grain 10001) both reach liquid L, their common face neighbour, this step;
    # A's index is lower (x = 3 before x = 5, same y and z)
    o, idx = _pin_state(tmp_path, {(3, 4, 3): (ACTIVE, 1, 1.0), (4, 4, 3): (LIQUID, 0, 0.0),
                                   (5, 4, 3): (ACTIVE, 10001, 1.0)}, tag=f"contest{mutation}")
    o.set_mutation(mutation)
    a, l, b = idx(3, 4, 3), idx(4, 4, 3), idx(5, 4, 3)
    o.step(1)
    s1 = (o.field(orc.CELLTYPE)[[a, l, b]].tolist(), int(o.field(orc.GRAINID)[l]), o.field(orc.DIAGLEN)[[a, l]])
    o.step(2)
    s2 = (o.field(orc.CELLTYPE)[[a, l, b]].tolist(), o.field(orc.DIAGLEN)[l])
    return s1, s2


def test_contested_cell_goes_to_lowest_index_and_new_cell_does_not_grow(tmp_path):
    (ct1, gl, dl1), (ct2, dl2) = _contest(tmp_path)
    # step 1: A captures L (grain 1); A saw L liquid at its visit and stays active; B finds no
    # liquid neighbour and solidifies; L keeps the capture's half-diagonal (a face capture from a
    # centred octahedron gives exactly 1) and A has grown one step
    assert ct1 == [ACTIVE, ACTIVE, SOLID]
    assert gl == 1
    assert dl1[0] == GROWN and dl1[1] == np.float32(1.0)
    # step 2: L grows for the first time; nothing liquid is left, so A and L solidify
    assert ct2 == [SOLID, SOLID, SOLID]
    assert dl2 == GROWN


def test_capture_order_pins_fail_planted_mistakes(tmp_path):
    """Each planted mistake fails the pins above: the higher index winning the contested cell, a
    captured neighbour not counting as liquid for its capturer, a new cell growing at once."""
    (ct1, gl, _), _ = _contest(tmp_path, orc.MUT_DESCENDING)
    assert gl == 10001 and ct1 == [SOLID, ACTIVE, ACTIVE]
    (ct1, gl, _), _ = _contest(tmp_path, orc.MUT_LCOUNT)
    assert gl == 1 and ct1[0] == SOLID
    (_, _, dl1), _ = _contest(tmp_path, orc.MUT_GROW_NEW)
    assert dl1[1] == GROWN


@pytest.mark.parametrize("capturer_below", [True, False])
def test_last_liquid_neighbour_taken_lower_or_higher(tmp_path, capturer_below):
    """Cell C (half-diagonal 0.01: reaches nothing) has one liquid neighbour L, which a capturer
    K takes this step. K below C in index order: at C's visit L is already taken, so C turns solid
    this step. K above C: C still sees L liquid at its visit and stays active for one more step."""
    xc, xk = (5, 3) if capturer_below else (3, 5)
    o, idx = _pin_state(tmp_path, {(xc, 4, 3): (ACTIVE, 1, 0.01), (4, 4, 3): (LIQUID, 0, 0.0),
                                   (xk, 4, 3): (ACTIVE, 10001, 1.0)})
    c, l, k = idx(xc, 4, 3), idx(4, 4, 3), idx(xk, 4, 3)
    o.step(1)
    ct = o.field(orc.CELLTYPE)
    assert ct[l] == ACTIVE and o.field(orc.GRAINID)[l] == 10001
    assert ct[c] == (SOLID if capturer_below else ACTIVE)
    assert ct[k] == ACTIVE
    o.step(2)
    assert o.field(orc.CELLTYPE)[[c, l, k]].tolist() == [SOLID, SOLID, SOLID]
    # descending visits reverse the outcome: the pin fails the planted mistake
    o2, _ = _pin_state(tmp_path, {(xc, 4, 3): (ACTIVE, 1, 0.01), (4, 4, 3): (LIQUID, 0, 0.0),
                                  (xk, 4, 3): (ACTIVE, 10001, 1.0)}, tag="mutated")
    o2.set_mutation(orc.MUT_DESCENDING)
    o2.step(1)
    assert o2.field(orc.CELLTYPE)[c] == (ACTIVE if capturer_below else SOLID)


# ---------------------------------------------------------------- state import and the OpenMP pass
def _own_state(o):
    """The oracle's state at a cycle before any capture: every active cell holds a cell-centred
    octahedron (initial seed or nucleus), so its centre is the cell centre."""
    d = o.dims()
    X, Y, nz = d["X"], d["Y"], d["nz"]
    ct = o.field(orc.CELLTYPE).astype(np.uint8)
    k, i, j = np.meshgrid(np.arange(nz), np.arange(X), np.arange(Y), indexing="ij")
    centre = np.stack([i - 1 + 0.5, j - 1 + 0.5, k + 0.5], axis=-1).reshape(-1).astype(np.float64)
    raw = ct | np.where(ct == ACTIVE, 0x10, 0).astype(np.uint8)
    dl = np.zeros(ct.size, dtype=np.float32)
    dl[:X * Y * d["nzA"]] = o.field(orc.DIAGLEN)
    return raw, o.field(orc.GRAINID), dl, centre, o.field(orc.NUCLEI_LOC), o.field(orc.NUCLEI_TIME)


def test_state_import_continues_the_run(tmp_path):
    """A run imported at cycle 200 (undercooling accumulated over up to 210 steps, nuclei placed,
    no capture yet: the first comes at cycle 232) continues exactly as the run it was taken from."""
    deck = c_deck(tmp_path / "i.txt", 24, density=30000, fraction=0.05)
    a = orc.OracleSim(deck, DATA)
    a.run(max_cycles=200)
    assert a.dims()["captures"] == 0 and a.dims()["nn"] > 100
    b = orc.OracleSim.from_state(deck, DATA, 200, *_own_state(a))
    assert np.array_equal(b.field(orc.UNDERCOOLING), a.field(orc.UNDERCOOLING) * (
        np.isin(a.field(orc.CELLTYPE), [ACTIVE, LIQUID])))
    assert np.array_equal(b.field(orc.CRITDIAG), a.field(orc.CRITDIAG) * np.repeat(
        a.field(orc.CELLTYPE)[:b.field(orc.DIAGLEN).size] == ACTIVE, 26))
    for cycle in range(201, 600):
        a.step(cycle)
        b.step(cycle)
    assert a.dims()["captures"] > 1000
    for f in (orc.CELLTYPE, orc.GRAINID, orc.DIAGLEN, orc.DOCENTER):
        assert np.array_equal(a.field(f), b.field(f)), f


def test_openmp_pass_conserves_cells(tmp_path):
    """The OpenMP CellCapture (a timing baseline: which cell wins a contested capture depends on
    thread timing) still accounts for every cell: each solidification turns one active cell solid,
    each capture or nucleation event one liquid cell active, and the run ends with no liquid."""
    deck = c_deck(tmp_path / "p.txt", 30, density=300, fraction=0.05)
    o = orc.OracleSim(deck, DATA)
    o.set_threads(4)
    ct0 = o.field(orc.CELLTYPE)
    o.run()
    d = o.dims()
    ct = o.field(orc.CELLTYPE)
    assert int((ct == LIQUID).sum()) == 0
    assert int((ct == SOLID).sum()) - int((ct0 == SOLID).sum()) == d["solidified"]
    # nucleation events: at least the interior ones (nn), at most one per site
    events = int((ct == ACTIVE).sum()) - int((ct0 == ACTIVE).sum()) - d["captures"] + d["solidified"]
    assert d["nn"] <= events <= d["nuclei"]
    assert int((ct0 == LIQUID).sum()) == d["captures"] + events
