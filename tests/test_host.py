"""Host logic on the CPU: deck parsing, the one-rank initial state, and the C ABI surface.

The product's initial state is built with libstdc++'s <random> directly; the oracle's with its own
C restatement of the same streams. Identical fields pin both.
"""
import re

import numpy as np
import pytest

import exaca_amd as ea
import oracle as orc
from conftest import DATA, EXAMPLES, REPO, c_deck, s_deck

FIELDS = [(ea.FIELD_CELLTYPE, orc.CELLTYPE), (ea.FIELD_GRAINID, orc.GRAINID),
          (ea.FIELD_CRITTIMESTEP, orc.CRITTIMESTEP), (ea.FIELD_LAYERID, orc.LAYERID),
          (ea.FIELD_NUCLEI_LOC, orc.NUCLEI_LOC), (ea.FIELD_NUCLEI_TIME, orc.NUCLEI_TIME)]

DECKS = {
    "small_dirs": lambda p: EXAMPLES / "Inp_SmallDirSolidification.txt",
    "dirs_200": lambda p: EXAMPLES / "Inp_DirSolidification.txt",
    "dirs_100_baseline": lambda p: EXAMPLES / "Inp_DirSolidification_100.txt",
    "small_spots": lambda p: EXAMPLES / "Inp_SmallSpotMelt.txt",
    "ragged": lambda p: c_deck(p / "r.txt", 17, nx=29, ny=11, density=2000, fraction=0.2),
    "spots_3x2": lambda p: s_deck(p / "s.txt", nsx=3, nsy=2, offset=18, radius=11, density=500, spacing=9),
    "empty_substrate": lambda p: c_deck(p / "e.txt", 9, density=0, fraction=0.0),
    # every third draw a nucleus site, and every draw one (probability above 1): the product's
    # block generator and integer cut must reproduce <random>'s decisions draw for draw
    "dense_nuclei": lambda p: c_deck(p / "n.txt", 21, nx=19, ny=23, density=300000, fraction=0.1),
    "all_nuclei": lambda p: c_deck(p / "a.txt", 8, density=2000000, fraction=0.1),
    # multilayer spot melts: later layers overwrite CritTimeStep and LayerID where their pools overlap
    "spot_melt_2_layers": lambda p: EXAMPLES / "Inp_SpotMelt.txt",
    "spots_3_layers": lambda p: s_deck(p / "l.txt", nsx=2, nsy=2, offset=16, radius=10, density=300, spacing=8,
                                       layers=3, layer_height=6),
}


@pytest.mark.parametrize("name", list(DECKS))
def test_initial_state_matches_oracle(name, tmp_path):
    deck = DECKS[name](tmp_path)
    sim = orc.OracleSim(deck, DATA)
    od = sim.dims()
    for ef, of in FIELDS:
        dims, got = ea.host_state(deck, ef)
        want = sim.field(of)
        assert got.shape == want.shape and np.array_equal(got, want), (name, ef)
    for k in ["X", "Y", "nz", "nzA", "nx", "ny", "nuclei"]:
        assert dims[k] == od[k]


def test_deck_values_of_reference_examples():
    d = ea.host_state(EXAMPLES / "Inp_SmallDirSolidification.txt")
    assert (d["nx"], d["ny"], d["nz"], d["X"], d["Y"], d["nzA"]) == (52, 52, 52, 54, 54, 52)
    d = ea.host_state(EXAMPLES / "Inp_SmallSpotMelt.txt")
    # nx = 4 + 2r + offset (nspots_x - 1), ny = 4 + 2r + offset (nspots_y - 1), nz = r + 3
    assert (d["nx"], d["ny"], d["nz"], d["nzA"]) == (54, 79, 28, 27)
    d = ea.host_state(EXAMPLES / "Inp_SpotMeltArray_6x6.txt")
    assert (d["nx"], d["ny"], d["nz"]) == (454, 454, 78)


def test_substrate_and_nuclei_statistics():
    """Substrate fraction and bulk nucleation density follow the deck (binomial 4-sigma bands);
    nucleation undercooling (time after the liquidus times dU) follows N(5 K, 0.5 K)."""
    deck = EXAMPLES / "Inp_DirSolidification.txt"  # 200^3, fraction 0.08, 10e12 m^-3, G/R/N = 5e5/3e5/25
    d, ct = ea.host_state(deck, ea.FIELD_CELLTYPE)
    _, cts = ea.host_state(deck, ea.FIELD_CRITTIMESTEP)
    _, t = ea.host_state(deck, ea.FIELD_NUCLEI_TIME)
    _, loc = ea.host_state(deck, ea.FIELD_NUCLEI_LOC)
    n_sites = (d["X"] - 2) * (d["Y"] - 2)
    active = int(np.count_nonzero(ct == 2))
    assert abs(active - 0.08 * n_sites) < 4 * np.sqrt(n_sites * 0.08 * 0.92)
    n_liquid_cells = (d["X"] - 2) * (d["Y"] - 2) * (d["nz"] - 3)
    assert abs(d["nuclei"] - 1e-5 * n_liquid_cells) < 4 * np.sqrt(1e-5 * n_liquid_cells)
    du = 3e5 * (1e-6 / (25 * 0.6))
    und = (t - cts[loc]) * du
    assert abs(und.mean() - 5.0) < 0.25 and abs(und.std() - 0.5) < 0.2
    assert np.all(np.diff(loc) > 0)  # numbered in index order


@pytest.mark.parametrize("edit,msg", [
    (lambda s: s.replace("Problem type: C", "Problem type C"), "must be separated"),
    (lambda s: s.replace("Cooling rate (K/s)", "Heating rate (K/s)"), "Cooling rate"),
    (lambda s: s.replace("misorientation values (Y or N)?: N", "misorientation values (Y or N)?: maybe"), "\"Y\" or \"N\""),
    (lambda s: s.replace("Problem type: C", "Problem type: R"), "problem type R"),
    (lambda s: s.replace("Material: Inconel625_Baseline", "Material: Unobtainium"), "Materials"),
])
def test_deck_errors(edit, msg, tmp_path):
    deck = c_deck(tmp_path / "bad.txt", 8)
    deck.write_text(edit(deck.read_text()))
    with pytest.raises(ea.ExacaError) as e:
        ea.host_state(deck)
    assert msg in str(e.value)
    with pytest.raises(RuntimeError):
        orc.OracleSim(deck, DATA)


def test_multilayer_deck_values():
    """Inp_SpotMelt.txt: 2 layers 20 cells apart; nz = r + 3 + (layers - 1) * offset, and the first
    layer's active window is r + 2 planes."""
    d = ea.host_state(EXAMPLES / "Inp_SpotMelt.txt")
    assert (d["nz"], d["nzA"]) == (75 + 3 + 20, 77)


def test_oracle_multilayer_run(tmp_path):
    """The oracle through three layers: each transition moves the window to just below the lowest
    active cell and up to the next layer's top (src/CAinitialize.cpp:2772-2832), the cycle count
    restarts, no nucleus or capture falls outside a window, and each layer ends with none of its
    cells liquid."""
    deck = s_deck(tmp_path / "l.txt", nsx=2, nsy=2, offset=16, radius=10, density=300, spacing=8, layers=3,
                  layer_height=6)
    o = orc.OracleSim(deck, DATA)
    lid = o.field(orc.LAYERID)
    d = o.dims()
    assert d["layers"] == 3 and (d["zlow"], d["zhigh"]) == (0, 11)
    XY = d["X"] * d["Y"]
    tops = [int(np.flatnonzero(lid == k).max() // XY) for k in range(3)]
    assert tops == [11, 17, 23]
    for layer in range(3):
        o.run(max_cycles=-1)  # this layer to its end
        d = o.dims()
        assert d["layer"] == layer and d["outside"] == 0
        ct = o.field(orc.CELLTYPE)
        assert not np.any((ct == 5) & (lid == layer)), "liquid left in the finished layer"
        if layer < 2:
            lowest = int((np.flatnonzero(ct == 2) // XY).min())
            o.next_layer()
            d = o.dims()
            assert (d["zlow"], d["zhigh"], d["cycle"]) == (lowest - 1, tops[layer + 1], 0)
    assert not np.any(o.field(orc.CELLTYPE) == 5)
    with pytest.raises(RuntimeError, match="no layer after"):
        o.next_layer()


def test_library_exports_every_declared_symbol():
    header = (REPO / "include" / "exaca_ca.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(exaca_\w+)\(", header, re.M))
    assert len(declared) >= 12
    lib = ea.library()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert set(ea._SIGNATURES) == declared


def test_device_entry_points_fail_cleanly_without_gpu(tmp_path):
    """Without a GPU the library reports a device error through the status code (no crash)."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    deck = c_deck(tmp_path / "n.txt", 6, density=0)
    with pytest.raises(ea.ExacaError) as e:
        ea.Solver(deck)
    assert e.value.status == 2
