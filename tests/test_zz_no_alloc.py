"""GPU: exaca_grain_stats_noalloc, the entry point that allocates nothing, against
exaca_grain_stats on the same solved domain, through the C ABI.

Two calls in a row (the second over a smaller region: the static scratch keeps the first call's
contents, so whatever must start at zero has to be cleared by the call), each compared with the
allocating entry point: integer results and histograms exactly, the means bit for bit (both run the
same kernels and the same host reduction). A domain whose GrainID range exceeds the static scratch
is refused before any kernel runs.
"""
import numpy as np
import pytest

import exaca_amd as ea
from conftest import c_deck

pytestmark = pytest.mark.gpu


def _same(a, b, tag):
    for key in ("cells", "melted", "grains", "nucleated_cells", "melted_top", "histogram"):
        assert a[key] == b[key], (tag, key, a[key], b[key])
    for key in ("mean_misorientation", "mean_misorientation_top", "mean_grain_volume", "nucleated_fraction"):
        assert np.float64(a[key]).tobytes() == np.float64(b[key]).tobytes(), (tag, key, a[key], b[key])


def test_noalloc_matches_allocating_entry_point(tmp_path):
    print("[no-alloc] 40^3 dense-nucleation deck: two regions, large then small", flush=True)
    deck = c_deck(tmp_path / "d.txt", 40, density=400, fraction=0.02)
    with ea.Solver(deck) as sol:
        sol.run()
        q = sol.query()
        nx, ny, nz = q["nx"], q["ny"], q["nz"]
        big = (0, nx - 1, 0, ny - 1, 0, nz - 1)
        small = (7, 9, 5, 30, 10, 12)
        first = sol.grain_stats(big, noalloc=True)
        second = sol.grain_stats(small, noalloc=True)
        _same(first, sol.grain_stats(big), "whole view")
        _same(second, sol.grain_stats(small), "small box after the whole view")
        assert second["cells"] < first["cells"] and second["grains"] < first["grains"]
        # and again in the other order, after the allocating calls
        _same(sol.grain_stats(big, noalloc=True), first, "whole view, second time")


def test_noalloc_refuses_a_grain_range_beyond_its_scratch(tmp_path):
    """Every cell a nucleation site (probability above 1): more than 2^20 GrainIDs."""
    print("[no-alloc] 104^3 deck with every cell a nucleus: refused on the host", flush=True)
    deck = c_deck(tmp_path / "a.txt", 104, density=2000000, fraction=0.1)
    with ea.Solver(deck) as sol:
        assert sol.query()["nuclei"] > (1 << 20)
        with pytest.raises(ea.ExacaError, match="static scratch") as e:
            sol.grain_stats(noalloc=True)
        assert e.value.status == 1
        # the allocating entry point still serves it
        assert sol.grain_stats()["cells"] > 0
