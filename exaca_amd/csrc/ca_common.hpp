// ca_common.hpp -- types shared by the host runtime and the CDNA4 kernels of the CA time step.
//
// Cell-state layout on the GPU: one 32-byte record per cell of the whole local domain, in the
// reference linear order p = z*X*Y + x*Y + y with y fastest (src/CAupdate.cpp:220-225). Active
// cells are sparse in that order (a front of a few percent of each plane over a band of planes),
// so everything a step reads or atomically updates for a cell sits in one aligned 32-byte sector:
//   lm   uint64  active cell: liquid neighbours in bits 0-25 (bit 31 set: not yet known, for a
//                nucleus or an initial seed); bits 32-63: serial + 1 of the last step in which a
//                cell of index >= this one captured one of its liquid neighbours (keep tag).
//                Liquid cell: low word = liquid marker (bits 26-31 set), high word = claim word,
//                the lowest index of a cell that claimed it in the current step (0xFFFFFFFF:
//                none). A claimed cell is captured in the same step and never turns liquid again,
//                so a claim word is read only in the step that wrote it.
//   dl   float   octahedron half-diagonal (DiagonalLength) after the growth of steps up to `win`
//   thr  float   lower bound on the critical lengths of the still-liquid neighbours (0: unknown)
//   cts  int32   CritTimeStep (liquidus crossing step)
//   gid  int32   GrainID
//   win  uint32  active cell: the growth index: dl includes the growth of every executed step up
//                to that many steps past the liquidus time (not yet set while kFresh or unknown)
//   lid  int8    LayerID
//   ct   uint8   cell type (values of src/CAtypes.hpp:11) | kSeedBit for cell-centred octahedra
//   flags        kCounted, kFresh
// The octahedron is a second record per cell, one aligned 128-byte line: the centre in double (as
// computed at capture; DOCenter is its float rounding) and, for a captured cell, the reference's
// 26 critical diagonal lengths, computed once by the capture. A cell-centred octahedron (nuclei,
// the initial substrate, layer seeds) reads its lengths from a per-orientation table instead. A
// frontier cell's evaluation reads the one line of lengths; a capture reads the claimant's centre.
#pragma once
#include <cstddef>
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
constexpr unsigned long long kLmUnknown = 1ull << 31;
constexpr unsigned long long kLmNow = (1ull << 26) - 1;
constexpr uint16_t kCounted = 1;
constexpr uint16_t kFresh = 2;  // mask built by the cell's capture; not yet looked at as an active cell
// Steps ahead a cell can be scheduled; a cell with nothing to do for longer is scheduled at the
// horizon and checks again.
constexpr uint32_t kAlarmSlots = 64;
// The apply pass: one workgroup per kApplyGroup capture-pass workgroups, each with a region of
// kCandStage solidification candidates.
constexpr uint32_t kApplyGroup = 2;
constexpr uint32_t kCandStage = 2048;
constexpr uint32_t apply_blocks(uint32_t claim_blocks) { return (claim_blocks + kApplyGroup - 1) / kApplyGroup; }
// Words of one calendar slot: one bit per cell, rounded up to 32-byte rows (the list rebuild reads
// a slot in 32-byte pieces).
constexpr size_t alarm_stride(long long ncell) {
    return ((static_cast<size_t>(ncell) + 31) / 32 + 7) & ~static_cast<size_t>(7);
}
constexpr int kMaxJumps = 4096;
constexpr int kMaxLayers = 256;

// The 26 neighbour offsets in the reference order: 0-8 at dy = -1, 9-16 in the y plane, 17-25 at
// dy = +1 (src/CAinitialize.cpp:595-674). The order fixes the seed-table column of each neighbour.
constexpr int kNbX[26] = {0, 1, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, 0, 1, -1, 0, 0, 1, -1, 1, -1};
constexpr int kNbY[26] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1};
constexpr int kNbZ[26] = {0, 0, 0, 1, -1, 1, 1, -1, -1, 1, -1, 1, 1, -1, -1, 0, 0, 0, 0, 0, 1, -1, 1, 1, -1, -1};
// The same offsets packed two bits apiece (offset + 1 at bits 2l, 2l+1), so that device code
// decodes a per-lane neighbour number with shifts; an indexed table would be placed in scratch.
constexpr unsigned long long pack_offsets(const int (&t)[26]) {
    unsigned long long v = 0;
    for (int l = 0; l < 26; l++) v |= static_cast<unsigned long long>(t[l] + 1) << (2 * l);
    return v;
}
constexpr unsigned long long kPackX = pack_offsets(kNbX);
constexpr unsigned long long kPackY = pack_offsets(kNbY);
constexpr unsigned long long kPackZ = pack_offsets(kNbZ);
constexpr unsigned kClaimNone = 0xFFFFFFFFu;
// Neighbour number of the opposite offset: cell q is neighbour l of r iff r is neighbour
// kOpposite[l] of q.
constexpr int opposite_of(int l) {
    for (int k = 0; k < 26; k++)
        if (kNbX[k] == -kNbX[l] && kNbY[k] == -kNbY[l] && kNbZ[k] == -kNbZ[l]) return k;
    return -1;
}
struct OppositeTable {
    int v[26];
};
constexpr OppositeTable make_opposites() {
    OppositeTable t{};
    for (int l = 0; l < 26; l++) t.v[l] = opposite_of(l);
    return t;
}
constexpr OppositeTable kOpposite = make_opposites();

// Device-resident step control block, double-buffered by step parity so that the kernels of a
// step read slot p while the first kernel prepares slot 1-p for the next step.
struct StepCtl {
    int32_t cycle[2];     // time step number seen by the kernels of a step of parity p
    uint32_t serial[2];   // steps executed before the step of parity p (its calendar slot)
    uint32_t count[2];    // lengths of the steps' lists of cells to process
    uint32_t elo, ehi;    // cell-index range [elo, ehi] holding every active cell (elo > ehi: none)
    uint32_t ncand;       // cells whose last liquid neighbour was captured in the current step
    uint32_t rebuild_seq; // advanced before every list rebuild (its flags' tag)
    int32_t njumps;       // cycle jumps taken so far (spot-melt problems)
    int32_t error;        // nonzero: list overflow or invalid grain
    int32_t done;         // set by the device when a domain with autostop finishes
    int32_t final_cycle;  // the check cycle at which it finished
    int32_t last_cycle;   // the last executed step
    int32_t next_nuc[2];  // least NucleationTime of the sites not yet decided, for the step of parity p
    unsigned long long nactive;     // active cells at the end of the last step
    unsigned long long nonwaiting;  // of those, past their liquidus time (the reference's numSteer)
    int32_t jump_from[kMaxJumps];  // cycle J before the jump; cycles J+1 .. T-1 were skipped
    int32_t jump_to[kMaxJumps];    // T, the first cycle executed after the jump
    int32_t layer;                 // the layer being solidified; every layer's cycles start at 1
    int32_t layer_end[kMaxLayers];   // last cycle of each finished layer
    int32_t layer_jump0[kMaxLayers]; // index of each layer's first cycle jump
    int32_t layer_zlo[kMaxLayers];   // each layer's active window [ZBound_Low, ZBound_High]: only
    int32_t layer_zhi[kMaxLayers];   // cells inside it accumulate undercooling in that layer
};

// Per-run counters kept on the device.
struct StepStats {
    unsigned long long frontier;        // sum over steps of cells past their liquidus time (numSteer)
    unsigned long long captures;        // liquid cells captured
    unsigned long long solidified;      // active cells turned solid
    unsigned long long nucleated;       // nucleation events counted as the reference's nn
    unsigned long long waiting;         // frontier entries still above their liquidus time
    unsigned long long activated;       // all nucleation events (interior or not)
    unsigned long long evaluated;       // frontier cells whose critical lengths were evaluated
    unsigned long long masks;           // liquid-neighbour masks built from the neighbour types
};

// Termination-check reduction (src/CAupdate.cpp:738-758 plus the min CritTimeStep of :792-808).
struct CheckResult {
    unsigned long long counts[5];  // superheated, undercooled, active, solid, future-layer liquid
    unsigned long long next_cts;   // min CritTimeStep (as unsigned long) of liquid cells in the layer
    unsigned long long amin, amax; // least and greatest index of an active cell (amin > amax: none)
};

struct alignas(32) Cell {
    unsigned long long lm;
    float dl;
    float thr;
    int32_t cts;
    int32_t gid;
    uint32_t win;
    int8_t lid;
    uint8_t ct;
    uint16_t flags;  // kCounted: past its liquidus time and counted in StepCtl::nonwaiting; kFresh
};
static_assert(sizeof(Cell) == 32, "one sector per cell");
struct alignas(128) Octa {
    float cdl[26];  // CritDiagonalLength in the neighbour order kNbX/kNbY/kNbZ (captured cells)
    double x, y, z; // centre
};
static_assert(sizeof(Octa) == 128, "one line per cell");
// Row stride (floats) of the per-orientation table of cell-centred critical lengths: rows are
// aligned like the lengths of an Octa record.
constexpr int kCdlStride = 32;

// Per-cell fields a caller can read back (values of EXACA_FIELD_* in the C interface).
enum FieldKind : int {
    kFieldCellType = 0,
    kFieldGrainId = 1,
    kFieldDiagonal = 2,
    kFieldCentre = 3,
    kFieldCritTimeStep = 4,
    kFieldLayerId = 5,
    kFieldCellTypeRaw = 6,
    kFieldCentreF64 = 10,
};

// Counters of the current step, one slot per workgroup (part[row * nparts + workgroup]): every
// workgroup overwrites its own slot, and the step's last kernel sums the rows. Same-address atomics
// from thousands of workgroups would serialise.
enum PartRow : int {
    kPartClaimSol,      // solidified in the capture pass
    kPartClaimEntered,  // first step past the liquidus time
    kPartClaimEval,     // critical-length evaluations
    kPartClaimMask,     // liquid masks built
    kPartNucSol,        // solidified at nucleation (mask emptied by a nucleus)
    kPartNucEntered,
    kPartNucAct,        // nucleation events
    kPartNucNn,         // nucleation events counted as the reference's nn
    kPartCap,           // captures
    kPartSettleSol,     // solidified in the settle pass
    kPartRows
};

// Everything a kernel needs to address one CA domain.
struct DomainView {
    int X, Y, nz, nzA;      // MyXSlices, MyYSlices, nz, and the top of the active window: cells of
                            // z < nzA (ZBound_High + 1, so nzActive while ZBound_Low = 0)
    int zlo;                // ZBound_Low, the window's bottom plane (0 in the first layer)
    int xoff, yoff;         // MyXOffset, MyYOffset (-1 for one rank)
    long long ncell;        // X*Y*nz
    Cell *cells;
    Octa *octa;             // octahedron records (centre, critical lengths)
    uint32_t *list[2];      // frontier of a step of parity p (list[1-p] also holds the step's
                            // solidification candidates until the next frontier is written)
    uint32_t *claim_cell;   // cells with claims in the current step, stored in their own frontier
    uint32_t *claim_bits;   //   slice, and the neighbours each claimed
    uint32_t *slice_claims; // entries of each workgroup's slice of the claim lists
    uint32_t *cand;         // kCandStage solidification candidates per apply workgroup
    uint32_t *cand_count;   // candidates in each region
    uint32_t cand_cap;      // candidates staged per apply workgroup (<= kCandStage; the rest go to
                            // the overflow list)
    uint32_t list_cap;
    int32_t nnuclei;
    const int32_t *nloc;
    int32_t *ntime;         // NucleationTimes; a site examined past its time is set to INT32_MAX
    const float *guv;       // 10000 x 9 unit-vector components (GrainUnitVector)
    const float *seed_cdl;  // 10000 rows of kCdlStride: the 26 critical diagonal lengths of a
                            // cell-centred octahedron of each orientation
    const double *grow;     // growth increment after k executed steps below the liquidus
    int32_t ngrow;          // table length; the last entry holds for every larger k
    int32_t ksat;           // from this k on every increment is the last entry (the 0.045 cap or the
                            // 210 K saturation)
    StepCtl *ctl;
    StepStats *stats;
    unsigned long long *part;  // kPartRows x nparts per-workgroup counters of the current step
    uint32_t nparts;
    uint32_t *alarms;       // kAlarmSlots bitmaps of one bit per cell: process the cell at the step
                            // whose serial is the bitmap's slot (mod kAlarmSlots)
    unsigned long long *chunk_flag;  // per unit of the list rebuild: tag | state | count or prefix
    uint32_t rebuild_cap;   // list-rebuild workgroups the device holds at once (0: no limit)
    uint32_t nchunks;
    int32_t autostop;       // end the domain on the device at the first check with no liquid left
    long long liquid0;      // liquid cells of the layer at the start
    int32_t *status;        // where the device reports the final cycle of an autostopped domain
    int32_t *status_err;    // where it reports the error flags of a domain stopped by an error
};

}  // namespace exaca
