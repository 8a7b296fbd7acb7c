"""Input decks in the reference's format (src/CAinitialize.cpp:232-576) and the random process-map
requests of the ensemble workloads (directional solidification with sizes, thermal gradient,
cooling rate and nucleation density drawn log-uniformly)."""
from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np


def directional_deck(path: Path, nx: int, ny: int, nz: int, *, G: float, R: float, N: int = 10,
                     density: float = 10.0, fraction: float = 0.04, material: str = "Inconel625_Baseline",
                     base: str = "case", out: str = "./", printing: bool = False) -> Path:
    """Writes a type-C deck. density is in the deck's unit (1e12 m^-3)."""
    yn = "Y" if printing else "N"
    Path(path).write_text(f"""directional solidification request
*****
Problem type: C
Decomposition strategy (1 for 1D domain decomposition, 2 for simple 2D decomposition, 3 for complex 2D decomposition): 1
Material: {material}
Cell size (microns: 1
Heterogeneous nucleation density (normalized by 1*10^12 per m^3): {density!r}
Mean nucleation undercooling (in K): 5
Standard deviation of nucleation undercooling: 0.5
Path to output: {out}
Output file base name: {base}
File of grain orientations: GrainOrientationVectors_Robert.csv
Thermal gradient (K/m): {G!r}
Cooling rate (K/s): {R!r}
Domain size in x: {nx}
Domain size in y: {ny}
Domain size in z: {nz}
"N", where the time step = deltax/(Steady State Velocity * N): {N}
Fraction of surface sites active: {fraction}
***Output data printing options: (Y or N) which data should be printed***
Print Paraview vtk file of grain misorientation values (Y or N)?: {yn}
Print Paraview vtk file of all ExaCA data for post-processing (Y or N)?: {yn}
Debug check (reduced)- Print data for main Kokkos views (CellType, LayerID, CritTimeStep) following initialization for debug check (Y or N)?: N
Debug check (extensive) - Print data for all main data structures (CellType, GrainID, CritTimeStep, UndercoolingCurrent, UndercoolingChange, Melted, LayerID) following initialization for debug check (Y or N)?: N
""")
    return Path(path)


@dataclass
class Request:
    index: int
    deck: Path
    seed: int
    nx: int
    ny: int
    nz: int
    G: float
    R: float
    density: float  # m^-3
    arrival: float = 0.0  # seconds after the stream starts


def _loguniform(rng, lo, hi, size=None):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size))


def process_map_requests(n: int, directory: Path, *, rng_seed: int = 2024, cubes: bool = False,
                         side=(32, 256), rate: float | None = None) -> list[Request]:
    """n requests: sides log-uniform in `side` (independent per axis unless cubes), G in
    [1e5, 1e7] K/m, R in [1e4, 1e7] K/s and nucleation density in [1e11, 1e14] m^-3, all
    log-uniform; a per-request random-stream seed; Poisson arrivals at `rate` per second."""
    rng = np.random.default_rng(rng_seed)
    out = []
    t = 0.0
    for i in range(n):
        if cubes:
            s = int(round(_loguniform(rng, *side)))
            nx = ny = nz = s
        else:
            nx, ny, nz = (int(round(v)) for v in _loguniform(rng, side[0], side[1], 3))
        G = float(_loguniform(rng, 1e5, 1e7))
        R = float(_loguniform(rng, 1e4, 1e7))
        dens = float(_loguniform(rng, 1e11, 1e14))
        if rate:
            t += float(rng.exponential(1.0 / rate))
        deck = directional_deck(Path(directory) / f"req{i:04d}.txt", nx, ny, nz, G=G, R=R, density=dens / 1e12,
                                base=f"req{i:04d}")
        out.append(Request(i, deck, 1000 + i, nx, ny, nz, G, R, dens, t if rate else 0.0))
    return out
