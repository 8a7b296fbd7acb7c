// init.hpp -- one-rank initial state of a CA domain, built on the host the way the reference's
// serial init loops build it (src/CAinitialize.cpp:1025-2248).
#pragma once
#include <cstdint>
#include <memory>
#include <new>
#include <utility>
#include <vector>

#include "deck.hpp"

namespace exaca {

// Allocator whose value-less construct leaves the element uninitialised, so that resizing a
// whole-domain field does not zero it on one thread (the init fills the fields plane by plane on
// threads instead).
template <typename T>
struct NoInitAlloc : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <typename U>
    NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    template <typename U>
    void construct(U *p) noexcept {
        ::new (static_cast<void *>(p)) U;
    }
    template <typename U, typename... A>
    void construct(U *p, A &&...a) {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};
template <typename T>
using Field = std::vector<T, NoInitAlloc<T>>;

struct HostDomain {
    int nx = 0, ny = 0, nz = 0;   // global sizes including the wall planes
    int X = 0, Y = 0;             // MyXSlices, MyYSlices
    int xoff = -1, yoff = -1;     // MyXOffset, MyYOffset
    int nzA = 0, zlow = 0, zhigh = 0;
    long long ncell = 0;          // X*Y*nz
    Field<uint8_t> ct;            // reference cell types (no seed bit)
    Field<int32_t> gid, cts;
    Field<int8_t> lid;
    Field<uint8_t> melted;
    float dU = 0;                 // UndercoolingChange shared by every cell that ever accumulates
    std::vector<int32_t> nloc, ntime;  // nucleation sites and steps (NucleiLocation/NucleationTimes)
    std::vector<float> guv;       // GrainUnitVector, 10000 x 9
    std::vector<uint32_t> active; // initially active cells, ascending index
    // Directional-solidification domains initialised on the device (initialize_sparse): the
    // whole-domain fields above stay empty; the device builds them from the deck, the per-plane
    // CritTimeStep and these lists (active = the substrate's active cells, numbered 1, 2, ... in
    // index order; nloc = the nucleation sites, numbered -1, -2, ... in index order).
    bool sparse = false;
    std::vector<int32_t> plane_cts;  // CritTimeStep of each plane
    // normalised interfacial response (A, B, C scaled by deltat/deltax; D unscaled, runCA.cpp:198-200)
    double A = 0, B = 0, C = 0, D = 0;
    int substrate_active = 0, baseplate_grains = 0;
};

// Runs TempInit_{DirSolidification,SpotMelt}, OrientationInit, the substrate init for the deck's
// problem type, ActiveCellWallInit (type S), GrainInit and NucleiInit for rank 0 of 1.
HostDomain initialize(const Deck &deck, bool verbose);

// The same initial state of a directional-solidification deck without its whole-domain fields:
// only the random draws run on the host (the substrate's active cells, GrainInit's nucleation
// sites, NucleiInit's undercoolings), in the reference's streams and orders; the device builds every
// cell from them (launch_init_directional). Decks of other types, and domains narrower than two
// interior cells in x or y (where a liquid cell could lack any non-wall neighbour and GrainInit's
// last pass would turn it solid), return a full host state instead (sparse == false).
HostDomain initialize_sparse(const Deck &deck, bool verbose);

// The 26 critical diagonal lengths of a cell-centred octahedron of one orientation
// (src/CAupdate.cpp:112-194), evaluated with the host libm exactly as the reference does.
void seed_critical_lengths(const float *guv9, float out26[26]);

// Growth increment per step after k executed steps below the liquidus, k = 0 .. n-1; the last
// entry is the saturated value (undercooling clamped at 210 K or no longer increasing).
std::vector<double> growth_table(double A, double B, double C, double D, float dU, int max_entries);
// Smallest k from which every entry of the growth table equals the last one.
int growth_plateau(const std::vector<double> &grow);

// Orientation file reader (src/CAinitialize.cpp:1408-1442).
std::vector<float> read_orientations(const std::string &path);

}  // namespace exaca
