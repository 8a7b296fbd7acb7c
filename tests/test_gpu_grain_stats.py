"""Grain statistics on the GPU (exaca_grain_stats) against the same quantities computed with numpy
from the downloaded GrainID field, the host Melted flags and the orientation file.

The reference computes them in its analysis executable (analysis/src/GAprint.cpp:17-108) and in
the misorientation vtk writer (src/CAprint.cpp:419-432). Integer results must match exactly; the
misorientation means within 1e-9 relative (the device path sums per-orientation counts in long
double, the reference sums cell by cell in long double: only the summation order differs).
"""
import math

import numpy as np
import pytest

import exaca_amd as ea
from conftest import DATA, c_deck

pytestmark = pytest.mark.gpu


def _orientation_angles():
    rows = (DATA / "Substrate" / "GrainOrientationVectors_Robert.csv").read_text().splitlines()
    n = int(rows[0])
    guv = np.array([[float(v) for v in r.split(",")] for r in rows[1:n + 1]], dtype=np.float32)
    ang = np.full(n, 62.7)
    for ll in range(3):
        a = np.abs((180 / math.pi) * np.arccos(guv[:, 3 * ll + 2].astype(np.float64)))
        ang = np.minimum(ang, a)
    return ang


def _whole_domain(field, q):
    """The reference's whole-domain view (k, i, j): this domain's interior for z = 1 .. nz-2, zero
    on the bottom and top planes."""
    v = field.reshape(q["nz"], q["X"], q["Y"])[:, 1:q["nx"] + 1, 1:q["ny"] + 1].astype(np.int64)
    v[0] = 0
    v[-1] = 0
    return v


def _expected(gid, melted, region, ang, deltax=1e-6):
    x0, x1, y0, y1, z0, z1 = region
    g = gid[z0:z1 + 1, x0:x1 + 1, y0:y1 + 1]
    m = melted[z0:z1 + 1, x0:x1 + 1, y0:y1 + 1] != 0
    vol = g.size
    # a melted wall cell (GrainID 0) counts in orientation 0 (the reference's writers never visit one)
    o = np.where(g == 0, 0, (np.abs(g) - 1) % 10000)
    af = ang.astype(np.float32).astype(np.float64)
    nm = int(m.sum())
    top = m[-1]
    grains = np.unique(g).size
    hist = np.bincount(np.floor(ang[o[m]] + 0.5).astype(np.int64), minlength=64)
    return {
        "cells": vol, "melted": nm, "grains": grains, "nucleated_cells": int((g < 0).sum()),
        "melted_top": int(top.sum()),
        "mean_misorientation": af[o[m]].sum() / nm,
        "mean_misorientation_top": af[o[-1][top]].sum() / max(int(top.sum()), 1),
        "mean_grain_volume": deltax * deltax * 1e12 * (vol / grains),
        "nucleated_fraction": (g < 0).sum() / vol,
        "histogram": hist,
    }


def test_grain_statistics_match_numpy(tmp_path):
    deck = c_deck(tmp_path / "d.txt", 40, density=400, fraction=0.02)
    ang = _orientation_angles()
    _, melted_flat = ea.host_state(deck, ea.FIELD_MELTED)
    with ea.Solver(deck) as sol:
        sol.run()
        q = sol.query()
        gid = _whole_domain(sol.field(ea.FIELD_GRAINID), q)
        melted = _whole_domain(melted_flat, q)
        nx, ny, nz = q["nx"], q["ny"], q["nz"]
        regions = [
            (1, nx - 2, 1, ny - 2, 1, nz - 2),      # everything inside the walls (the default)
            (3, nx - 5, 2, ny - 9, 6, nz - 8),      # an interior box
            (0, nx - 1, 0, ny - 1, 0, nz - 1),      # the whole view, walls and zero planes included
            (7, 7, 5, 30, 10, 10),                  # one row
        ]
        for region in regions:
            got = sol.grain_stats(region)
            exp = _expected(gid, melted, region, ang)
            for key in ("cells", "melted", "grains", "nucleated_cells", "melted_top"):
                assert got[key] == exp[key], (region, key, got[key], exp[key])
            assert np.array_equal(np.array(got["histogram"]), exp["histogram"]), region
            assert sum(got["histogram"]) == got["melted"]
            for key in ("mean_misorientation", "mean_misorientation_top", "mean_grain_volume", "nucleated_fraction"):
                assert got[key] == pytest.approx(exp[key], rel=1e-9, abs=1e-12), (region, key)
        default = sol.grain_stats()
        assert default["cells"] == (nx - 2) * (ny - 2) * (nz - 2)
        # a finished type-C run: every interior cell melted, nuclei present in the dense-nucleation deck
        assert default["melted"] == default["cells"]
        assert default["nucleated_cells"] > 0
        with pytest.raises(ea.ExacaError):
            sol.grain_stats((0, nx, 0, ny - 1, 0, nz - 1))
