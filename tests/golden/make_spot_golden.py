"""Writes the oracle's full runs of the spot-melt example decks as fixtures for the GPU tests:
  spot_array_6x6.json      examples/Inp_SpotMeltArray_6x6.txt (6 x 6 spots of radius 75 um at 60 um
                           pitch, arrays 456 x 456 x 78, about 1e5 steps: most of an hour serially)
  spot_melt_2_layers.json  examples/Inp_SpotMelt.txt (3 x 2 spots, two layers 20 cells apart:
                           about 13 minutes serially)
far longer than a GPU test may take. Each holds the final cycle (of the last layer), the step
counters and SHA-1 digests of the final CellType and GrainID fields (int32, the reference's linear
order).

Run from the repository root with the oracle built (make -C oracle):
    python tests/golden/make_spot_golden.py [deck fixture]   (default: the 6 x 6 array)
"""
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO / "oracle"))
import oracle as orc  # noqa: E402

DATA = REPO / "exaca_amd" / "data"


def main() -> None:
    deck_name, fixture = (sys.argv[1], sys.argv[2]) if len(sys.argv) > 2 else ("Inp_SpotMeltArray_6x6.txt",
                                                                             "spot_array_6x6.json")
    DECK = REPO / "examples" / deck_name
    OUT = Path(__file__).resolve().parent / fixture
    t0 = time.time()
    sim = orc.OracleSim(DECK, DATA)
    cycle, cap_s = sim.run()
    d = sim.dims()
    ct = np.ascontiguousarray(sim.field(orc.CELLTYPE), dtype=np.int32)
    gid = np.ascontiguousarray(sim.field(orc.GRAINID), dtype=np.int32)
    rec = {
        "deck": DECK.name,
        "dims": [d["X"], d["Y"], d["nz"]],
        "final_cycle": cycle,
        "nn": d["nn"],
        "frontier": d["frontier"],
        "captures": d["captures"],
        "solidified": d["solidified"],
        "layers": d["layers"],
        "outside_window": d["outside"],
        "celltype_sha1": hashlib.sha1(ct.tobytes()).hexdigest(),
        "grainid_sha1": hashlib.sha1(gid.tobytes()).hexdigest(),
        "celltype_counts": {str(int(k)): int(v) for k, v in zip(*np.unique(ct, return_counts=True))},
        "oracle_capture_s": cap_s,
        "oracle_wall_s": time.time() - t0,
    }
    OUT.write_text(json.dumps(rec, indent=1) + "\n")
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
