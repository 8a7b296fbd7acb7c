// ca_kernels.hpp -- host-side launchers for the CDNA4 kernels of one CA time step.
#pragma once
#include <hip/hip_runtime.h>

#include "ca_common.hpp"

namespace exaca {

constexpr int kBlock = 256;  // four wave64s per workgroup

// Nucleation (src/CAupdate.cpp:67-200). A step may leave it out when no site's time falls in the
// steps since nucleation last ran (it would examine nothing); its capture pass is then launched
// with nucleated = false.
void launch_nucleation(const DomainView &d, int parity, int layer, hipStream_t s);
// Its workgroups (each writes its own slot of the nucleation counter rows).
int nucleation_blocks(const DomainView &d);
// CellCapture pass A: liquidus-time gate, growth, neighbour scan, claims (src/CAupdate.cpp:217-352).
// Also prepares the control slot of the next step.
void launch_claim(const DomainView &d, int parity, int layer, int blocks, hipStream_t s);
// CellCapture pass B: claims decided, winners build new octahedra, cells left without a liquid
// neighbour by lower-indexed captures solidify, and the next frontier is written
// (src/CAupdate.cpp:353-725).
// nucleated: whether this step's nucleation kernel ran (its counter rows are booked).
void launch_apply(const DomainView &d, int parity, int layer, int blocks, bool nucleated, hipStream_t s);
// Every active cell due at the next step, and that step's list rebuilt: after a cycle jump, whose
// skipped steps the calendar did not count. parity_next is the parity of the next step.
void launch_rewake(const DomainView &d, int parity_next, int blocks, hipStream_t s);
// The next layer's window: out[0] = least z of an active cell, out[1] = greatest z of a cell of
// next_layer (out initialised to INT32_MAX, -1 by the caller).
void launch_layer_bounds(const DomainView &d, int next_layer, int *out, hipStream_t s);
// LayerSetup on the resident state for the window [d.zlo, d.nzA), then the first step's list.
void launch_layer_setup(const DomainView &d, int parity_next, int blocks, hipStream_t s);
// One step of every domain in a device-resident table of `slots` views (grid row = slot).
void launch_step_n(const DomainView *views, int slots, int parity, int layer, int nuc_blocks, int blocks,
                   hipStream_t s);
// IntermediateOutputAndCheck's reduction (src/CAupdate.cpp:738-758, 792-808).
void launch_check(const DomainView &d, int cycle, int layer, CheckResult *out, hipStream_t s);
// Octahedra of the initially active cells (src/CAinitialize.cpp:1936-2043).
void launch_seed_init(const DomainView &d, const uint32_t *cells, uint32_t n, hipStream_t s);
// One step with an event pair around each kernel: nucleation, claim, winners, apply, settle, and
// the frontier rebuild. ev holds 2 * kTimedKernels events.
constexpr int kTimedKernels = 6;
void launch_step_timed(const DomainView &d, int parity, int layer, int blocks, hipStream_t s, hipEvent_t *ev);
// Initial cell records from compact device arrays (CellType with the seed bit, LayerID, GrainID,
// CritTimeStep), and one field back out of the records (FieldKind layout).
void launch_pack(const DomainView &d, const uint8_t *ct, const int8_t *lid, const int32_t *gid, const int32_t *cts,
                 hipStream_t s);
void launch_extract(const DomainView &d, int field, void *out, hipStream_t s);
// The initial state of a directional-solidification domain built on the device from its per-plane
// CritTimeStep, the substrate's active cells and the nucleation sites (both in index order); the
// Melted flags go to `melted` (one byte per cell).
void launch_init_directional(const DomainView &d, const int32_t *plane_cts, int nx, int ny, uint8_t *melted,
                             const uint32_t *substrate, uint32_t nsub, const int32_t *sites, uint32_t nsites,
                             hipStream_t s);
// Bytes per cell of a field as launch_extract writes it (0: not a per-cell field).
inline size_t field_bytes(int field) {
    switch (field) {
    case kFieldCellTypeRaw: return 1;
    case kFieldCentre: return 12;
    case kFieldCentreF64: return 24;
    case kFieldCellType:
    case kFieldGrainId:
    case kFieldDiagonal:
    case kFieldCritTimeStep:
    case kFieldLayerId: return 4;
    default: return 0;
    }
}
// Grain statistics of the whole-domain region [k0, k1] x [i0, i1] x [j0, j1] (inclusive; the
// reference's vtk / analysis layout): flags gflag[GrainID - gmin] (ngid entries), counts[0..2] +=
// melted cells, cells with a negative GrainID, cells whose GrainID falls outside the flag range;
// obins / otop += melted cells per orientation, in the region / in its top plane k1.
struct GrainRegion {
    int k0, k1, i0, i1, j0, j1;
};
void launch_grain_stats(const DomainView &d, const uint8_t *melted, const GrainRegion &r, int gmin, uint32_t ngid,
                        uint32_t *gflag, unsigned long long *counts, unsigned int *obins, unsigned int *otop,
                        hipStream_t s);
// The list rebuild: workgroups the device holds at once (its occupancy on cus CUs), and the units
// of a domain of ncell cells. An ensemble domain has at most kEnsembleUnits units, so that one
// round of workgroups covers it.
uint32_t rebuild_resident(int cus);
uint32_t rebuild_units(long long ncell);
constexpr uint32_t kEnsembleUnits = 128;
// Unique GrainIDs of a grain-statistics pass: the nonzero entries of gflag[0..n), added to *out.
void launch_count_flags(const uint32_t *gflag, uint32_t n, unsigned long long *out, hipStream_t s);
// p[0..n) = v, as a kernel (no memset node in a captured graph).
void launch_fill(uint32_t *p, size_t n, uint32_t v, hipStream_t s);
// Plain copy of n 16-byte words (HBM calibration).
void launch_copy(const uint4 *src, uint4 *dst, size_t n, hipStream_t s);
// Device-side self test of the capture geometry against host values: one octahedron per input row.
void launch_geometry_probe(const float *guv, const double *in, double *out, int n, hipStream_t s);

}  // namespace exaca
