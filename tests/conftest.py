"""Shared pytest configuration: the `gpu` marker, paths and a deck writer for small cases."""
import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
DATA = REPO / "exaca_amd" / "data"
EXAMPLES = REPO / "examples"
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def c_deck(path: Path, n: int, *, nx=None, ny=None, density=10.0, fraction=0.04, G=1e6, R=1e6, N=10,
           material="Inconel625_Baseline", base="case", out="./", printing="N", dTN=5, dTsigma=0.5) -> Path:
    """Writes a directional-solidification (type C) deck in the reference's format."""
    nx = n if nx is None else nx
    ny = n if ny is None else ny
    path.write_text(f"""small directional solidification test case
*****
Problem type: C
Decomposition strategy (1 for 1D domain decomposition, 2 for simple 2D decomposition, 3 for complex 2D decomposition): 1
Material: {material}
Cell size (microns: 1
Heterogeneous nucleation density (normalized by 1*10^12 per m^3): {density}
Mean nucleation undercooling (in K): {dTN}
Standard deviation of nucleation undercooling: {dTsigma}
Path to output: {out}
Output file base name: {base}
File of grain orientations: GrainOrientationVectors_Robert.csv
Thermal gradient (K/m): {G}
Cooling rate (K/s): {R}
Domain size in x: {nx}
Domain size in y: {ny}
Domain size in z: {n}
"N", where the time step = deltax/(Steady State Velocity * N): {N}
Fraction of surface sites active: {fraction}
***Output data printing options: (Y or N) which data should be printed***
Print Paraview vtk file of grain misorientation values (Y or N)?: {printing}
Print Paraview vtk file of all ExaCA data for post-processing (Y or N)?: {printing}
Debug check (reduced)- Print data for main Kokkos views (CellType, LayerID, CritTimeStep) following initialization for debug check (Y or N)?: N
Debug check (extensive) - Print data for all main data structures (CellType, GrainID, CritTimeStep, UndercoolingCurrent, UndercoolingChange, Melted, LayerID) following initialization for debug check (Y or N)?: N
""")
    return path


def s_deck(path: Path, *, nsx=2, nsy=2, offset=30, radius=20, density=10.0, G=1e6, R=1e6, N=10, spacing=25,
           layers=1, layer_height=20, material="Inconel625_Baseline", base="spots", out="./", printing="N") -> Path:
    """Writes a spot-melt array (type S) deck in the reference's format."""
    path.write_text(f"""small spot melt array test case
*****
Problem type: S
Decomposition strategy (1 for 1D domain decomposition, 2 for simple 2D decomposition, 3 for complex 2D decomposition): 1
Material: {material}
Cell size (microns): 1
Heterogeneous nucleation density (normalized by 1*10^12 per m^3): {density}
Mean nucleation undercooling (in K): 5
Standard deviation of nucleation undercooling: 0.5
Path to output: {out}
Output file base name: {base}
File of grain orientations: GrainOrientationVectors_Robert.csv
Thermal gradient (K/m): {G}
Cooling rate (K/s): {R}
Number of spots in x: {nsx}
Number of spots in y: {nsy}
Offset between spot centers (in microns): {offset}
Radii of spots (in microns): {radius}
Number of layers: {layers}
Offset between layers (in cells): {layer_height}
"N", where the time step = deltax/(Steady State Velocity * N): {N}
Substrate grain spacing (in microns): {spacing}
***Output data printing options: (Y or N) which data should be printed***
Print Paraview vtk file of grain misorientation values (Y or N)?: {printing}
Print Paraview vtk file of all ExaCA data for post-processing (Y or N)?: {printing}
Debug check (reduced)- Print data for main Kokkos views (CellType, LayerID, CritTimeStep) following initialization for debug check (Y or N)?: N
Debug check (extensive) - Print data for all main data structures (CellType, GrainID, CritTimeStep, UndercoolingCurrent, UndercoolingChange, Melted, LayerID) following initialization for debug check (Y or N)?: N
""")
    return path


@pytest.fixture
def data_dir():
    return DATA
