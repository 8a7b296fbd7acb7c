// ca_common.hpp -- types shared by the host runtime and the CDNA4 kernels of the CA time step.
//
// Cell-state layout on the GPU (structure of arrays, one entry per cell of the whole local
// domain, reference linear order p = z*X*Y + x*Y + y with y fastest, src/CAupdate.cpp:220-225):
//   ct   uint8   cell type (values of src/CAtypes.hpp:11) | SEED_BIT for cell-centred octahedra
//   gid  int32   GrainID
//   cts  int32   CritTimeStep (liquidus crossing step)
//   lid  int8    LayerID
//   dl   float   octahedron half-diagonal (DiagonalLength)
//   cx/cy/cz double  octahedron centre as computed at capture (DOCenter is its float rounding)
//   win  uint64  capture claim key of the current step (generation stamp << 32 | capturer)
// The reference's 26 critical diagonal lengths per cell (104 B) are not stored: they are
// recomputed bit-exactly from the double centre (captured cells) or read from a per-orientation
// table (cell-centred octahedra: nuclei and the initial substrate).
#pragma once
#include <cstdint>

namespace exaca {

enum CellType : uint8_t {
    kWall = 0,
    kSolid = 1,
    kActive = 2,
    kTemporaryUpdate = 3,
    kTemporaryInit = 4,
    kLiquid = 5,
    kGhost = 6,
};
// Marks an active cell whose octahedron is centred on the cell (set by nucleation and by the
// initial grain structure); its critical diagonal lengths come from the orientation table.
constexpr uint8_t kSeedBit = 0x10;
constexpr uint8_t kTypeMask = 0x0F;
constexpr int kNumOrientations = 10000;
constexpr uint32_t kWaitMask = 0x80000000u;  // frontier entry not yet past its liquidus time
constexpr int kMaxJumps = 4096;

// The 26 neighbour offsets in the reference order: 0-8 at dy = -1, 9-16 in the y plane, 17-25 at
// dy = +1 (src/CAinitialize.cpp:595-674). The order fixes the seed-table column of each neighbour.
constexpr int kNbX[26] = {0, 1, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, 0, 1, -1, 0, 0, 1, -1, 1, -1};
constexpr int kNbY[26] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1};
constexpr int kNbZ[26] = {0, 0, 0, 1, -1, 1, 1, -1, -1, 1, -1, 1, 1, -1, -1, 0, 0, 0, 0, 0, 1, -1, 1, 1, -1, -1};

// Device-resident step control block, double-buffered by step parity so that the kernels of a
// step read slot p while the first kernel prepares slot 1-p for the next step.
struct StepCtl {
    int32_t cycle[2];     // time step number seen by the kernels of a step of parity p
    uint32_t serial[2];   // executed-step counter (claim generation stamp)
    uint32_t count[2];    // frontier list lengths
    int32_t njumps;       // cycle jumps taken so far (spot-melt problems)
    int32_t error;        // nonzero: frontier list overflow or invalid grain
    int32_t done;         // set by the device when a domain with autostop finishes
    int32_t final_cycle;  // the check cycle at which it finished
    int32_t jump_from[kMaxJumps];  // cycle J before the jump; cycles J+1 .. T-1 were skipped
    int32_t jump_to[kMaxJumps];    // T, the first cycle executed after the jump
};

// Per-run counters kept on the device.
struct StepStats {
    unsigned long long frontier;        // sum over steps of cells past their liquidus time (numSteer)
    unsigned long long captures;        // liquid cells captured
    unsigned long long solidified;      // active cells turned solid
    unsigned long long nucleated;       // nucleation events counted as the reference's nn
    unsigned long long waiting;         // frontier entries still above their liquidus time
    unsigned long long activated;       // all nucleation events (interior or not)
};

// Termination-check reduction (src/CAupdate.cpp:738-758 plus the min CritTimeStep of :792-808).
struct CheckResult {
    unsigned long long counts[5];  // superheated, undercooled, active, solid, future-layer liquid
    unsigned long long next_cts;   // min CritTimeStep (as unsigned long) of liquid cells in the layer
};

// Everything a kernel needs to address one CA domain.
struct DomainView {
    int X, Y, nz, nzA;      // MyXSlices, MyYSlices, nz, nzActive (ZBound_Low = 0 on one GPU)
    int xoff, yoff;         // MyXOffset, MyYOffset (-1 for one rank)
    long long ncell;        // X*Y*nz
    uint8_t *ct;
    int32_t *gid;
    int32_t *cts;
    int8_t *lid;
    float *dl;
    double *cx, *cy, *cz;
    unsigned long long *win;
    uint32_t *list[2];
    uint32_t *mask;
    uint32_t *slice_out;    // next-frontier entries produced by each workgroup slice
    uint32_t list_cap;
    int32_t nnuclei;
    const int32_t *nloc;
    const int32_t *ntime;
    const float *guv;       // 10000 x 9 unit-vector components (GrainUnitVector)
    const float *seed_cdl;  // 10000 x 26 critical diagonal lengths of cell-centred octahedra
    const double *reach;    // per orientation: c with critical length^2 >= c |x - centre|^2
    const double *grow;     // growth increment after k executed steps below the liquidus
    int32_t ngrow;          // table length; the last entry holds for every larger k
    StepCtl *ctl;
    StepStats *stats;
    uint32_t *bits;         // one bit per cell: in the frontier (active)
    uint32_t *chunk_sum;    // frontier cells per 2048-word chunk of bits, then their exclusive prefix
    uint32_t nchunks;
    int32_t index_order;    // rebuild each next frontier from bits, in cell-index order
    int32_t autostop;       // end the domain on the device at the first check with no liquid left
    long long liquid0;      // liquid cells of the layer at the start
    int32_t *status;        // where the device reports the final cycle of an autostopped domain
};

}  // namespace exaca
