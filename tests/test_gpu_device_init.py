"""GPU: the initial state of directional-solidification decks built on the device (the host draws
the random streams only; launch_init_directional builds every cell) against the full host
initialisation (exaca_host_state, which the CPU tests pin against the oracle): CellType, GrainID,
CritTimeStep, LayerID, the Melted flags and the nucleation sites must be identical, and every
initially active cell must hold a cell-centred octahedron of half-diagonal 0.01
(src/CAinitialize.cpp:1025-1055, 1460-1529, 1834-2248).
"""
import numpy as np
import pytest

import exaca_amd as ea
from conftest import EXAMPLES, c_deck

pytestmark = pytest.mark.gpu

DECKS = {
    "small_dirs": lambda p: EXAMPLES / "Inp_SmallDirSolidification.txt",
    "dirs_100_baseline": lambda p: EXAMPLES / "Inp_DirSolidification_100.txt",
    "ragged": lambda p: c_deck(p / "r.txt", 17, nx=29, ny=11, density=2000, fraction=0.2),
    "empty_substrate": lambda p: c_deck(p / "e.txt", 9, density=0, fraction=0.0),
    "dense_nuclei": lambda p: c_deck(p / "n.txt", 21, nx=19, ny=23, density=300000, fraction=0.1),
    "all_nuclei": lambda p: c_deck(p / "a.txt", 8, density=2000000, fraction=0.1),
    "narrow_2x2": lambda p: c_deck(p / "w.txt", 12, nx=2, ny=2, density=50000, fraction=0.5),
    "dirs_800": lambda p: EXAMPLES / "Inp_DirSolidification_800.txt",
}
FIELDS = [ea.FIELD_CELLTYPE, ea.FIELD_GRAINID, ea.FIELD_CRITTIMESTEP, ea.FIELD_LAYERID, ea.FIELD_MELTED]


@pytest.mark.parametrize("name", list(DECKS))
def test_device_init_equals_host_init(tmp_path, name):
    deck = DECKS[name](tmp_path)
    print(f"[device init] {name}", flush=True)
    with ea.Solver(deck) as sol:
        q = sol.query()
        for f in FIELDS:
            dims, want = ea.host_state(deck, f)
            got = sol.field(f)
            bad = np.flatnonzero(got != want)
            assert bad.size == 0, (name, f, bad.size, bad[:5], got[bad[:5]], want[bad[:5]])
            del want, got
        for f in (ea.FIELD_NUCLEI_LOC, ea.FIELD_NUCLEI_TIME):
            _, want = ea.host_state(deck, f)
            assert np.array_equal(sol.field(f), want), (name, f)
        assert q["nuclei"] == dims["nuclei"]
        ct = sol.field(ea.FIELD_CELLTYPE)
        act = np.flatnonzero(ct == 2)
        assert act.size == dims["active"]
        dl = sol.field(ea.FIELD_DIAGONAL)[act]
        assert np.all(dl == np.float32(0.01))
        X, Y = q["X"], q["Y"]
        centre = sol.field(ea.FIELD_DOCENTER).reshape(-1, 3)[act]
        z, rem = act // (X * Y), act % (X * Y)
        want_c = np.stack([rem // Y - 1 + 0.5, rem % Y - 1 + 0.5, z + 0.5], axis=1).astype(np.float32)
        assert np.array_equal(centre, want_c)
