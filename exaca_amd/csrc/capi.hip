// capi.hip -- the C ABI (include/exaca_ca.h): device state of one CA domain, the per-step entry
// points, the device-resident layer loop replayed from HIP graphs, and the RunProgram_Reduced
// drop-in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdlib>
#include <limits>
#include <memory>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/exaca_ca.h"
#include "ca_common.hpp"
#include "ca_kernels.hpp"
#include "deck.hpp"
#include "init.hpp"
#include "output.hpp"
#include "runtime_util.hpp"

namespace exaca {
thread_local std::string g_last_error;
}

namespace exaca {
// The control block of a domain before its first step: the initially active cells are the first
// step's list, and the range of active cells is theirs.
StepCtl initial_ctl(const HostDomain &hd) {
    StepCtl c;
    std::memset(&c, 0, sizeof c);
    c.cycle[0] = 1;
    c.count[0] = static_cast<uint32_t>(hd.active.size());
    c.elo = 0xFFFFFFFFu;
    c.ehi = 0;
    for (uint32_t q : hd.active) {
        c.elo = std::min(c.elo, q);
        c.ehi = std::max(c.ehi, q);
    }
    c.nactive = hd.active.size();
    c.layer_zlo[0] = hd.zlow;
    c.layer_zhi[0] = hd.zhigh;
    c.next_nuc[0] = 0x7FFFFFFF;
    c.next_nuc[1] = 0x7FFFFFFF;
    for (int32_t t : hd.ntime) c.next_nuc[0] = std::min(c.next_nuc[0], t);
    return c;
}

// Initial cell records of a domain: the compact host arrays are staged in device buffers that are
// free until the first step (the claim lists and the second list: at least 12 bytes per cell) and
// packed on the device. The initially active cells are the first step's list (list[0]), with their
// cell-centred octahedra; the calendar starts empty. Synchronises the stream (the host sources
// are temporaries).
void upload_cells(const DomainView &d, const HostDomain &hd, uint8_t *melted, hipStream_t s) {
    const size_t n = static_cast<size_t>(hd.ncell);
    if (hd.sparse) {
        // built on the device from the per-plane CritTimeStep and the draws' lists, staged in
        // buffers that are free until the first step; the active list is the first step's list
        int32_t *plane_cts = reinterpret_cast<int32_t *>(d.claim_bits);
        int32_t *sites = reinterpret_cast<int32_t *>(d.claim_cell);
        HIP_OK(hipMemcpyAsync(plane_cts, hd.plane_cts.data(), 4 * hd.plane_cts.size(), hipMemcpyHostToDevice, s));
        if (!hd.nloc.empty())
            HIP_OK(hipMemcpyAsync(sites, hd.nloc.data(), 4 * hd.nloc.size(), hipMemcpyHostToDevice, s));
        if (!hd.active.empty())
            HIP_OK(hipMemcpyAsync(d.list[0], hd.active.data(), 4 * hd.active.size(), hipMemcpyHostToDevice, s));
        launch_init_directional(d, plane_cts, hd.nx, hd.ny, melted, d.list[0], static_cast<uint32_t>(hd.active.size()),
                                sites, static_cast<uint32_t>(hd.nloc.size()), s);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemsetAsync(d.alarms, 0, 4 * kAlarmSlots * alarm_stride(d.ncell), s));
        HIP_OK(hipMemsetAsync(d.chunk_flag, 0, 8ull * d.nchunks, s));
        HIP_OK(hipStreamSynchronize(s));
        return;
    }
    if (melted) HIP_OK(hipMemcpyAsync(melted, hd.melted.data(), n, hipMemcpyHostToDevice, s));
    std::vector<uint8_t> ct(hd.ct.begin(), hd.ct.end());
    for (uint32_t q : hd.active) ct[q] |= kSeedBit;
    uint8_t *ct_d = reinterpret_cast<uint8_t *>(d.claim_cell);
    int8_t *lid_d = reinterpret_cast<int8_t *>(ct_d + n);
    int32_t *gid_d = reinterpret_cast<int32_t *>(d.list[1]);
    int32_t *cts_d = reinterpret_cast<int32_t *>(d.claim_bits);
    HIP_OK(hipMemcpyAsync(ct_d, ct.data(), n, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(lid_d, hd.lid.data(), n, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(gid_d, hd.gid.data(), 4 * n, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(cts_d, hd.cts.data(), 4 * n, hipMemcpyHostToDevice, s));
    launch_pack(d, ct_d, lid_d, gid_d, cts_d, s);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemsetAsync(d.alarms, 0, 4 * kAlarmSlots * alarm_stride(d.ncell), s));
    HIP_OK(hipMemsetAsync(d.chunk_flag, 0, 8ull * d.nchunks, s));
    if (!hd.active.empty()) {
        HIP_OK(hipMemcpyAsync(d.list[0], hd.active.data(), 4 * hd.active.size(), hipMemcpyHostToDevice, s));
        launch_seed_init(d, d.list[0], static_cast<uint32_t>(hd.active.size()), s);
        HIP_OK(hipGetLastError());
    }
    HIP_OK(hipStreamSynchronize(s));
}

// One per-cell field of a domain into host memory (the layout launch_extract writes).
void download_field(const DomainView &d, int field, void *dst, size_t dst_bytes, hipStream_t s) {
    const size_t per = field_bytes(field);
    if (per == 0) throw std::invalid_argument("unknown field " + std::to_string(field));
    const size_t bytes = per * static_cast<size_t>(d.ncell);
    if (dst_bytes < bytes) throw std::invalid_argument("destination too small");
    void *tmp = nullptr;
    HIP_OK(hipMalloc(&tmp, bytes));
    try {
        launch_extract(d, field, tmp, s);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(dst, tmp, bytes, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    } catch (...) {
        (void)hipFree(tmp);
        throw;
    }
    HIP_OK(hipFree(tmp));
}
}  // namespace exaca

namespace {

using exaca::DeviceError;
using exaca::StateError;
using exaca::guarded;
using exaca::g_last_error;

constexpr int kGraphSteps = 50;  // even: a graph starts and ends on parity 0

}  // namespace

struct exaca_state {
    exaca::Deck deck;
    exaca::HostDomain host;
    exaca::DomainView dv{};
    hipStream_t stream = nullptr;
    exaca::StepCtl *ctl = nullptr;
    exaca::StepStats *stats = nullptr;
    exaca::CheckResult *check = nullptr;
    std::vector<void *> buffers;
    size_t bytes = 0;
    int parity = 0;        // parity of the next step
    int cycle = 0;         // last executed step
    int pending = -1;      // cycle whose nucleation ran but whose capture has not
    int blocks = 2048;     // workgroups of the frontier kernels
    int njumps = 0;
    hipGraphExec_t graph = nullptr;       // kGraphSteps steps
    hipGraphExec_t graph_quiet = nullptr; // the same without the nucleation kernels
    int graph_layer = -1;
    // Nucleation times in ascending order, and the last step whose nucleation kernel ran: a step
    // may leave the kernel out when no time falls after that step and up to it (every site of a
    // time at or before a step where the kernel ran is decided there). Single-layer decks only: a
    // site of a later layer waits past its time.
    std::vector<int32_t> nuc_times;
    int nuc_through = 0;
    bool nuc_skip = false;
    double init_seconds = 0;
    int device = 0;
    int layer = 0;              // the layer being solidified
    uint8_t *melted = nullptr;  // the reference's Melted flag per cell (grain statistics)
    int gmin = 0;               // least initial GrainID and the size of the GrainID range
    uint32_t ngid = 0;          //   (0: not yet found)

    template <typename T>
    T *alloc(size_t n) {
        void *p = nullptr;
        HIP_OK(hipMalloc(&p, n * sizeof(T) > 0 ? n * sizeof(T) : 1));
        buffers.push_back(p);
        bytes += n * sizeof(T);
        return static_cast<T *>(p);
    }
    ~exaca_state() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (graph_quiet) (void)hipGraphExecDestroy(graph_quiet);
        for (void *p : buffers) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void build_device_state(exaca_state *h) {
    using namespace exaca;
    const HostDomain &hd = h->host;
    HIP_OK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    h->device = dev;
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, dev));
    // Frontier kernels: up to 4 workgroups per CU (the capture pass is then resident in one round;
    // measured at 800^3: 2, 3, 4, 6, 8 per CU give 110, 101, 96, 102, 118 us), fewer for small
    // domains (the frontier never exceeds the cell count, and a slice is at least one workgroup of
    // cells). (At most 2048: the per-workgroup step counters are summed by one workgroup, 8 slots
    // a thread.)
    const long long cap = std::min<long long>(2048, ((prop.multiProcessorCount * 4 + 7) / 8) * 8);
    const long long want = ((hd.ncell / 4096 + 7) / 8) * 8;
    h->blocks = static_cast<int>(std::max<long long>(8, std::min(cap, want)));
    if (const int forced = exaca::frontier_blocks_override()) h->blocks = forced;

    DomainView &d = h->dv;
    d.X = hd.X;
    d.Y = hd.Y;
    d.nz = hd.nz;
    d.nzA = hd.nzA;
    d.xoff = hd.xoff;
    d.yoff = hd.yoff;
    d.ncell = hd.ncell;
    const size_t n = static_cast<size_t>(hd.ncell);
    d.cells = h->alloc<Cell>(n);
    d.octa = h->alloc<Octa>(n);
    d.list_cap = static_cast<uint32_t>(n);
    d.list[0] = h->alloc<uint32_t>(n);
    d.list[1] = h->alloc<uint32_t>(n);
    d.claim_cell = h->alloc<uint32_t>(n);
    d.claim_bits = h->alloc<uint32_t>(n);
    d.slice_claims = h->alloc<uint32_t>(h->blocks);
    d.cand = h->alloc<uint32_t>(static_cast<size_t>(apply_blocks(h->blocks)) * kCandStage);
    d.cand_count = h->alloc<uint32_t>(apply_blocks(h->blocks));
    d.cand_cap = cand_stage_cap();
    const size_t nwords = (n + 31) / 32;
    d.nchunks = static_cast<uint32_t>((nwords + 2047) / 2048);
    d.alarms = h->alloc<uint32_t>(kAlarmSlots * alarm_stride(d.ncell));
    d.chunk_flag = h->alloc<unsigned long long>(d.nchunks);
    d.rebuild_cap = exaca::rebuild_resident(prop.multiProcessorCount);
    d.nnuclei = static_cast<int32_t>(hd.nloc.size());
    int32_t *nloc = h->alloc<int32_t>(hd.nloc.size());
    int32_t *ntime = h->alloc<int32_t>(hd.ntime.size());
    d.nloc = nloc;
    d.ntime = ntime;
    float *guv = h->alloc<float>(hd.guv.size());
    d.guv = guv;
    std::vector<float> seed(static_cast<size_t>(kCdlStride) * kNumOrientations, 0.0f);
    for (int o = 0; o < kNumOrientations; o++) seed_critical_lengths(&hd.guv[9 * o], &seed[kCdlStride * o]);
    float *seed_d = h->alloc<float>(seed.size());
    d.seed_cdl = seed_d;
    std::vector<double> grow = growth_table(hd.A, hd.B, hd.C, hd.D, hd.dU, 1 << 25);
    if (grow.size() >= (1u << 25)) throw std::runtime_error("undercooling never saturates within 2^25 steps");
    double *grow_d = h->alloc<double>(grow.size());
    d.grow = grow_d;
    d.ngrow = static_cast<int32_t>(grow.size());
    d.ksat = growth_plateau(grow);
    h->ctl = h->alloc<StepCtl>(1);
    h->stats = h->alloc<StepStats>(1);
    h->check = h->alloc<CheckResult>(1);
    d.ctl = h->ctl;
    d.stats = h->stats;
    d.nparts = 2048;  // the capture grid (<= 2048) and the nucleation grid (<= 1024)
    d.part = h->alloc<unsigned long long>(static_cast<size_t>(kPartRows) * d.nparts);

    // Initial state; per-cell undercooling is not stored (it is a function of the steps executed
    // below the liquidus time).
    hipStream_t s = h->stream;
    if (!hd.nloc.empty()) {
        HIP_OK(hipMemcpyAsync(nloc, hd.nloc.data(), hd.nloc.size() * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(ntime, hd.ntime.data(), hd.ntime.size() * 4, hipMemcpyHostToDevice, s));
    }
    HIP_OK(hipMemcpyAsync(guv, hd.guv.data(), hd.guv.size() * 4, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(seed_d, seed.data(), seed.size() * 4, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(grow_d, grow.data(), grow.size() * 8, hipMemcpyHostToDevice, s));
    h->melted = h->alloc<uint8_t>(n);
    upload_cells(d, hd, h->melted, s);
    StepCtl c = initial_ctl(hd);
    HIP_OK(hipMemcpyAsync(h->ctl, &c, sizeof c, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemsetAsync(h->stats, 0, sizeof(StepStats), s));
    HIP_OK(hipMemsetAsync(d.part, 0, 8ull * kPartRows * d.nparts, s));
    HIP_OK(hipStreamSynchronize(s));
}

// Sets the cycle the next step will see. A forward move past the next step records a jump so
// that the lazily derived undercooling skips the steps that were not executed.
void set_next_cycle(exaca_state *h, int next) {
    if (next == h->cycle + 1) return;
    if (next <= h->cycle) throw StateError("time steps must increase (last " + std::to_string(h->cycle) + ")");
    if (h->njumps >= exaca::kMaxJumps) throw StateError("too many cycle jumps");
    const int from = h->cycle, to = next, slot = h->njumps;
    exaca::StepCtl *c = h->ctl;
    HIP_OK(hipMemcpyAsync(&c->jump_from[slot], &from, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(&c->jump_to[slot], &to, 4, hipMemcpyHostToDevice, h->stream));
    h->njumps++;
    HIP_OK(hipMemcpyAsync(&c->njumps, &h->njumps, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(&c->cycle[h->parity], &next, 4, hipMemcpyHostToDevice, h->stream));
    // The calendar counts executed steps and waits for liquidus times in cycles: every active cell
    // is due at the first step after the jump, and is scheduled afresh from there.
    exaca::launch_rewake(h->dv, h->parity, h->blocks, h->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(h->stream));  // host values above are stack temporaries
    h->cycle = next - 1;
}

// The transition to the next layer (src/runCA.cpp:343-418): DomainShiftAndResize's new window
// (src/CAinitialize.cpp:2772-2832), LayerSetup's cell-centred octahedra for the window's active
// cells (:2835-2982), and the cycle count restarting at 0. The finished layer's last cycle and cycle
// jumps stay on record: the undercooling of a cell carried into the new layer keeps counting its
// earlier layers' steps.
void next_layer(exaca_state *h, int layer_done, int verbose) {
    using namespace exaca;
    if (h->pending >= 0) throw StateError("layer change between nucleation and capture");
    if (layer_done != h->layer) throw StateError("layer " + std::to_string(layer_done) + " is not the current layer");
    if (layer_done + 1 >= h->deck.num_layers) throw StateError("no layer after the last one");
    if (layer_done + 1 >= kMaxLayers) throw StateError("too many layers");
    HIP_OK(hipStreamSynchronize(h->stream));
    int bounds[2] = {0x7FFFFFFF, -1};
    int *bounds_d = reinterpret_cast<int *>(h->check);  // the check's result block is free between steps
    HIP_OK(hipMemcpyAsync(bounds_d, bounds, sizeof bounds, hipMemcpyHostToDevice, h->stream));
    launch_layer_bounds(h->dv, layer_done + 1, bounds_d, h->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(bounds, bounds_d, sizeof bounds, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    const int zlo = bounds[0] - 1, zhigh = bounds[1];
    if (bounds[0] == 0x7FFFFFFF || zhigh < 0 || zlo < 0 || zhigh < zlo || zhigh >= h->dv.nz)
        throw StateError("no active window for layer " + std::to_string(layer_done + 1));
    if (verbose) {
        std::cout << "New domain bottom at Z = " << zlo << ", a shift of " << zlo - h->dv.zlo << " cells" << std::endl;
        std::cout << "New domain top at Z = " << zhigh << std::endl;
        std::cout << "New active domain height is " << zhigh - zlo + 1 << std::endl;
    }
    h->dv.zlo = zlo;
    h->dv.nzA = zhigh + 1;
    for (hipGraphExec_t *ge : {&h->graph, &h->graph_quiet})  // they hold the old window in their arguments
        if (*ge) {
            HIP_OK(hipGraphExecDestroy(*ge));
            *ge = nullptr;
        }
    // the layer's record, the new layer's first cycle, and fresh counts of the active cells
    StepCtl *c = h->ctl;
    const int32_t end = h->cycle, j0 = h->njumps, layer = layer_done + 1, one = 1, zero = 0, open = 0;
    const uint32_t range[2] = {0xFFFFFFFFu, 0};
    const unsigned long long none[2] = {0, 0};
    HIP_OK(hipMemcpyAsync(&c->layer_end[layer_done], &end, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(&c->layer_jump0[layer], &j0, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(&c->layer, &layer, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(&c->layer_zlo[layer], &zlo, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(&c->layer_zhi[layer], &zhigh, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(&c->cycle[h->parity], &one, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(&c->last_cycle, &zero, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(&c->next_nuc[h->parity], &open, 4, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(&c->elo, range, sizeof range, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipMemcpyAsync(&c->nactive, none, sizeof none, hipMemcpyHostToDevice, h->stream));
    // the calendar starts empty: every active cell of the window is due at the first step
    launch_fill(h->dv.alarms, static_cast<size_t>(kAlarmSlots) * alarm_stride(h->dv.ncell), 0u, h->stream);
    launch_layer_setup(h->dv, h->parity, h->blocks, h->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(h->stream));  // host values above are stack temporaries
    h->layer = layer;
    h->cycle = 0;
}

void step_nucleation(exaca_state *h, int layer) {
    exaca::launch_nucleation(h->dv, h->parity, layer, h->stream);
}
void step_capture(exaca_state *h, int layer, bool nucleated = true) {
    exaca::launch_claim(h->dv, h->parity, layer, h->blocks, h->stream);
    exaca::launch_apply(h->dv, h->parity, layer, h->blocks, nucleated, h->stream);
}
// Whether steps h->cycle + 1 .. last need the nucleation kernel.
bool nucleation_due(const exaca_state *h, int last) {
    if (!h->nuc_skip) return true;
    const auto it = std::upper_bound(h->nuc_times.begin(), h->nuc_times.end(), h->nuc_through);
    return it != h->nuc_times.end() && *it <= last;
}

void ensure_graph(exaca_state *h, int layer) {
    if (h->graph && h->graph_layer == layer) return;
    for (hipGraphExec_t *ge : {&h->graph, &h->graph_quiet})
        if (*ge) {
            HIP_OK(hipGraphExecDestroy(*ge));
            *ge = nullptr;
        }
    for (bool nucleate : {true, false}) {
        hipGraph_t g;
        HIP_OK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < kGraphSteps; k++) {
            const int p = k & 1;
            if (nucleate) exaca::launch_nucleation(h->dv, p, layer, h->stream);
            exaca::launch_claim(h->dv, p, layer, h->blocks, h->stream);
            exaca::launch_apply(h->dv, p, layer, h->blocks, nucleate, h->stream);
        }
        HIP_OK(hipStreamEndCapture(h->stream, &g));
        HIP_OK(hipGraphInstantiate(nucleate ? &h->graph : &h->graph_quiet, g, nullptr, nullptr, 0));
        HIP_OK(hipGraphDestroy(g));
    }
    h->graph_layer = layer;
}

void run_steps(exaca_state *h, int layer, int n, bool use_graph) {
    while (n > 0) {
        if (use_graph && h->parity == 0 && n >= kGraphSteps) {
            ensure_graph(h, layer);
            const int last = h->cycle + kGraphSteps;
            const bool nucleate = nucleation_due(h, last);
            HIP_OK(hipGraphLaunch(nucleate ? h->graph : h->graph_quiet, h->stream));
            if (nucleate) h->nuc_through = last;
            h->cycle = last;
            n -= kGraphSteps;
        }
        else {
            const bool nucleate = nucleation_due(h, h->cycle + 1);
            if (nucleate) {
                step_nucleation(h, layer);
                h->nuc_through = h->cycle + 1;
            }
            step_capture(h, layer, nucleate);
            h->parity ^= 1;
            h->cycle += 1;
            n -= 1;
        }
    }
    HIP_OK(hipGetLastError());
}

void device_check(exaca_state *h, int *cycle, int layer, uint64_t counts[5], int *xswitch, int verbose) {
    if (h->pending >= 0) throw StateError("check called between nucleation and capture");
    if (*cycle != h->cycle) throw StateError("check cycle differs from the last executed step");
    exaca::CheckResult r;
    std::memset(&r, 0, sizeof r);
    r.next_cts = ~0ull;
    r.amin = ~0ull;
    r.amax = 0;
    HIP_OK(hipMemcpyAsync(h->check, &r, sizeof r, hipMemcpyHostToDevice, h->stream));
    exaca::launch_check(h->dv, *cycle, layer, h->check, h->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(&r, h->check, sizeof r, hipMemcpyDeviceToHost, h->stream));
    exaca::StepStats st;
    int32_t err = 0;
    HIP_OK(hipMemcpyAsync(&st, h->stats, sizeof st, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipMemcpyAsync(&err, &h->ctl->error, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    if (err) throw DeviceError(exaca::device_error_text(err));
    if (counts)
        for (int k = 0; k < 5; k++) counts[k] = r.counts[k];
    // The range of active cells only widens between checks; narrow it to the cells active now.
    uint32_t range[2] = {0xFFFFFFFFu, 0};
    if (r.amin <= r.amax) {
        range[0] = static_cast<uint32_t>(r.amin);
        range[1] = static_cast<uint32_t>(r.amax);
    }
    HIP_OK(hipMemcpyAsync(&h->ctl->elo, range, sizeof range, hipMemcpyHostToDevice, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    if (verbose)
        std::cout << "cycle = " << *cycle << " Superheated liquid cells = " << r.counts[0]
                  << " Undercooled liquid cells = " << r.counts[1] << " Number of nucleation events this layer "
                  << st.nucleated << " Remaining liquid cells in future layers of this simulation = " << r.counts[4]
                  << std::endl;
    if (r.counts[0] + r.counts[1] == 0) *xswitch = 1;
    // Spot-melt problems skip ahead when the next liquidus crossing is more than 5000 steps away
    // (src/CAupdate.cpp:787-817; unsigned long arithmetic as there).
    if (*xswitch == 0 && h->deck.sim_type == "S" && r.counts[1] == 0) {
        const unsigned long next = static_cast<unsigned long>(r.next_cts);
        if ((next - static_cast<unsigned long>(static_cast<long>(*cycle))) > 5000) {
            set_next_cycle(h, static_cast<int>(next));
            *cycle = h->cycle;
            if (verbose) std::cout << "Jumping to cycle " << *cycle + 1 << std::endl;
        }
    }
}

void run_layer(exaca_state *h, int layer, int max_cycles, int verbose, int *out_cycle, double *gpu_ms) {
    if (h->pending >= 0) throw StateError("run_layer called between nucleation and capture");
    if (layer != h->layer) throw StateError("layer " + std::to_string(layer) + " is not the current layer " +
                                            std::to_string(h->layer) + " (exaca_next_layer moves on)");
    const bool graphs = std::getenv("EXACA_NO_GRAPHS") == nullptr;
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, h->stream));
    int xswitch = 0;
    do {
        int target = (h->cycle / 1000 + 1) * 1000;
        if (max_cycles > 0 && target > max_cycles) target = max_cycles;
        run_steps(h, layer, target - h->cycle, graphs);
        if (h->cycle % 1000 == 0) {
            int c = h->cycle;
            device_check(h, &c, layer, nullptr, &xswitch, verbose);
        }
        if (max_cycles > 0 && h->cycle >= max_cycles) break;
    } while (xswitch == 0);
    HIP_OK(hipEventRecord(e1, h->stream));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int32_t err = 0;
    HIP_OK(hipMemcpy(&err, &h->ctl->error, 4, hipMemcpyDeviceToHost));
    if (err) throw DeviceError(exaca::device_error_text(err));
    if (out_cycle) *out_cycle = h->cycle;
    if (gpu_ms) *gpu_ms = ms;
}

// The writers read LayerID and Melted from the host state; a domain initialised on the device has
// them only there (LayerID is 0 everywhere in a directional-solidification domain).
void host_fields_for_output(exaca_state *h) {
    exaca::HostDomain &hd = h->host;
    if (!hd.sparse) return;
    const size_t n = static_cast<size_t>(hd.ncell);
    hd.lid.assign(n, 0);
    hd.melted.resize(n);
    HIP_OK(hipMemcpy(hd.melted.data(), h->melted, n, hipMemcpyDeviceToHost));
}

exaca_state *create_state(const std::string &input, const std::string &dd, bool verbose, uint64_t seed = 0) {
    auto t0 = std::chrono::steady_clock::now();
    auto h = std::make_unique<exaca_state>();
    h->deck = exaca::read_deck(input, dd);
    h->deck.seed = seed;
    h->host = exaca::initialize_sparse(h->deck, verbose);
    build_device_state(h.get());
    h->nuc_times = h->host.ntime;
    std::sort(h->nuc_times.begin(), h->nuc_times.end());
    h->nuc_through = std::numeric_limits<int>::min();
    h->nuc_skip = h->deck.num_layers == 1 && std::getenv("EXACA_NUCLEATE_EVERY_STEP") == nullptr;
    h->init_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return h.release();
}

// Steps one at a time with HIP events around every kernel (no graphs): per-kernel device time of
// nucleation, claim, apply, settle and the frontier rebuild (bitmap count / scan / write).
void profile(exaca_state *h, int layer, int nsteps, double *ms, int64_t counts[5], int64_t *launches) {
    using exaca::kTimedKernels;
    exaca::StepStats before, after;
    HIP_OK(hipStreamSynchronize(h->stream));
    HIP_OK(hipMemcpy(&before, h->stats, sizeof before, hipMemcpyDeviceToHost));
    const int per_batch = 128;
    std::vector<hipEvent_t> ev(2 * kTimedKernels * per_batch);
    for (auto &e : ev) HIP_OK(hipEventCreate(&e));
    double acc[kTimedKernels] = {};
    int done = 0, xswitch = 0;
    while (done < nsteps && !xswitch) {
        const int batch = std::min(per_batch, nsteps - done);
        int k = 0;
        for (; k < batch; k++) {
            exaca::launch_step_timed(h->dv, h->parity, layer, h->blocks, h->stream, &ev[2 * kTimedKernels * k]);
            h->parity ^= 1;
            h->cycle += 1;
            h->nuc_through = h->cycle;
            if (h->cycle % 1000 == 0) {
                k++;
                break;
            }
        }
        HIP_OK(hipStreamSynchronize(h->stream));
        for (int q = 0; q < k; q++)
            for (int r = 0; r < kTimedKernels; r++) {
                float t = 0;
                HIP_OK(hipEventElapsedTime(&t, ev[2 * kTimedKernels * q + 2 * r], ev[2 * kTimedKernels * q + 2 * r + 1]));
                acc[r] += t;
            }
        done += k;
        if (h->cycle % 1000 == 0) {
            int c = h->cycle;
            device_check(h, &c, layer, nullptr, &xswitch, 0);
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    HIP_OK(hipMemcpy(&after, h->stats, sizeof after, hipMemcpyDeviceToHost));
    for (int r = 0; r < kTimedKernels; r++) ms[r] = acc[r];
    counts[0] = static_cast<int64_t>(after.frontier - before.frontier);
    counts[1] = static_cast<int64_t>(after.captures - before.captures);
    counts[2] = static_cast<int64_t>(after.solidified - before.solidified);
    counts[3] = done;
    counts[4] = done;
    *launches = done;
}

// ---- Grain statistics (analysis/src/GAprint.cpp:17-108): one device pass over a region into
// caller-provided buffers, then the host reduction of its counters and orientation bins. ----
struct StatsBuffers {
    uint32_t *gflag;          // one flag per GrainID of the initial range
    unsigned long long *cnt;  // melted cells, nucleated cells, cells outside the range, unique GrainIDs
    unsigned int *bins;       // melted cells per orientation: the region, then its top plane
};

exaca::GrainRegion checked_region(const exaca_state *h, const int32_t region[6]) {
    const exaca::HostDomain &hd = h->host;
    const int x0 = region[0], x1 = region[1], y0 = region[2], y1 = region[3], z0 = region[4], z1 = region[5];
    if (x0 < 0 || x1 >= hd.nx || x0 > x1 || y0 < 0 || y1 >= hd.ny || y0 > y1 || z0 < 0 || z1 >= hd.nz || z0 > z1)
        throw std::invalid_argument("region outside the domain");
    return exaca::GrainRegion{z0, z1, x0, x1, y0, y1};
}

// Every GrainID of a run is an initial one (captures copy them) or 0: the flag range is the
// initial field's, found once per handle.
void grain_id_range(exaca_state *h) {
    if (h->ngid) return;
    int gmin = 0, gmax = 0;
    if (h->host.sparse) {  // substrate grains 1 .. n, nuclei -1 .. -P
        gmin = -static_cast<int>(h->host.nloc.size());
        gmax = static_cast<int>(h->host.active.size());
    }
    for (int32_t g : h->host.gid) {
        gmin = g < gmin ? g : gmin;
        gmax = g > gmax ? g : gmax;
    }
    h->gmin = gmin;
    h->ngid = static_cast<uint32_t>(static_cast<long long>(gmax) - gmin + 1);
}

void grain_stats_pass(exaca_state *h, const exaca::GrainRegion &r, const StatsBuffers &b) {
    // cleared by a fill kernel, not a memset node (a caller may capture these launches)
    exaca::launch_fill(b.gflag, h->ngid, 0u, h->stream);
    exaca::launch_fill(reinterpret_cast<uint32_t *>(b.cnt), 8, 0u, h->stream);
    exaca::launch_fill(b.bins, 2ull * exaca::kNumOrientations, 0u, h->stream);
    exaca::launch_grain_stats(h->dv, h->melted, r, h->gmin, h->ngid, b.gflag, b.cnt, b.bins,
                              b.bins + exaca::kNumOrientations, h->stream);
    exaca::launch_count_flags(b.gflag, h->ngid, b.cnt + 3, h->stream);
    HIP_OK(hipGetLastError());
}

// The printed quantities from the device results c[4] and ob[2 * kNumOrientations] (no heap use).
void grain_stats_finish(const exaca_state *h, const int32_t region[6], const unsigned long long c[4],
                        const unsigned int *ob, int64_t counts[5], double values[4], int64_t hist[64]) {
    const exaca::HostDomain &hd = h->host;
    if (c[2]) throw DeviceError("a GrainID outside the initial range");
    const long long grains = static_cast<long long>(c[3]);
    // per orientation: the least angle of its three axes to +Z, capped at 62.7 degrees
    // (analysis/src/GAprint.cpp:38-47: stored as float; src/CAprint.cpp:423-430: rounded)
    long double sum = 0, sum_top = 0;
    long long top = 0;
    int64_t hb[64] = {};
    for (int o = 0; o < exaca::kNumOrientations; o++) {
        double amin = 62.7;
        for (int ll = 0; ll < 3; ll++) {
            const double a = std::abs((180 / M_PI) * std::acos(static_cast<double>(hd.guv[9 * o + 3 * ll + 2])));
            if (a < amin) amin = a;
        }
        const float af = static_cast<float>(amin);
        sum += static_cast<long double>(ob[o]) * af;
        sum_top += static_cast<long double>(ob[exaca::kNumOrientations + o]) * af;
        top += ob[exaca::kNumOrientations + o];
        hb[static_cast<int>(std::round(amin))] += ob[o];
    }
    const long long vol = static_cast<long long>(region[1] - region[0] + 1) * (region[3] - region[2] + 1) *
                          (region[5] - region[4] + 1);
    const long long nmelt = static_cast<long long>(c[0]);
    if (counts) {
        counts[0] = vol;
        counts[1] = nmelt;
        counts[2] = grains;
        counts[3] = static_cast<int64_t>(c[1]);
        counts[4] = top;
    }
    if (values) {
        values[0] = nmelt ? static_cast<double>(sum / nmelt) : 0.0;
        values[1] = top ? static_cast<double>(sum_top / top) : 0.0;
        values[2] = h->deck.deltax * h->deck.deltax * std::pow(10.0, 12) * (static_cast<double>(vol) / grains);
        values[3] = static_cast<double>(c[1]) / static_cast<double>(vol);
    }
    if (hist) std::copy(hb, hb + 64, hist);
}

// ---- The static scratch of exaca_grain_stats_noalloc: counters, orientation bins and GrainID
// flags in one file-scope device array (each GPU holds its own copy until the process ends). ----
constexpr size_t kScratchCountsOff = 0;
constexpr size_t kScratchBinsOff = 256;
constexpr size_t kScratchFlagsOff = kScratchBinsOff + ((8ull * exaca::kNumOrientations + 255) / 256) * 256;
constexpr uint32_t kScratchMaxGrainIds = 1u << 20;
constexpr size_t kScratchBytes = kScratchFlagsOff + 4ull * kScratchMaxGrainIds;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Device address of `scratch` on the current device, looked up once per device (host code must
// never pass the symbol itself to a kernel: on the host it names a stand-in, not device memory).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};
void *scratch_address() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw DeviceError("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// Makes the handle's device current for the duration of a call.
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev) {
        HIP_OK(hipGetDevice(&prev));
        if (prev != dev) HIP_OK(hipSetDevice(dev));
        else prev = -1;
    }
    ~DeviceScope() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace

extern "C" {

const char *exaca_last_error(void) { return g_last_error.c_str(); }

int exaca_create(const char *input_file, const char *data_dir, int verbose, exaca_state **out) {
    return guarded([&] {
        if (!input_file || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        *out = create_state(input_file, data_dir ? data_dir : exaca::data_dir(), verbose != 0);
    });
}

int exaca_create_seeded(const char *input_file, const char *data_dir, uint64_t seed, int verbose, exaca_state **out) {
    return guarded([&] {
        if (!input_file || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        *out = create_state(input_file, data_dir ? data_dir : exaca::data_dir(), verbose != 0, seed);
    });
}

void nucleation_call(exaca_state *h, int cycle, int layer) {
    if (!h) throw std::invalid_argument("null handle");
    if (h->pending >= 0) throw StateError("nucleation called twice without capture");
    if (layer != h->layer) throw StateError("layer " + std::to_string(layer) + " is not the current layer");
    set_next_cycle(h, cycle);
    step_nucleation(h, layer);
    HIP_OK(hipGetLastError());
    h->pending = cycle;
    h->nuc_through = cycle;
}

int exaca_nucleation(exaca_state *h, int cycle, int layer) {
    return guarded([&] { nucleation_call(h, cycle, layer); });
}

int exaca_nucleation_nn(exaca_state *h, int cycle, int layer, int32_t *nn) {
    return guarded([&] {
        if (!nn) throw std::invalid_argument("null nn");
        nucleation_call(h, cycle, layer);
        // the count booked by the steps before, plus this step's per-workgroup counts (the capture
        // pass books them at the end of the step)
        const int nb = exaca::nucleation_blocks(h->dv);
        std::vector<unsigned long long> row(static_cast<size_t>(nb));
        exaca::StepStats st;
        HIP_OK(hipMemcpyAsync(row.data(), h->dv.part + static_cast<size_t>(exaca::kPartNucNn) * h->dv.nparts,
                              row.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipMemcpyAsync(&st, h->stats, sizeof st, hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        unsigned long long total = st.nucleated;
        for (unsigned long long v : row) total += v;
        *nn = static_cast<int32_t>(total);
    });
}

int exaca_cell_capture(exaca_state *h, int cycle, int layer) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null handle");
        if (h->pending != cycle) throw StateError("capture requires nucleation of the same cycle first");
        step_capture(h, layer);
        HIP_OK(hipGetLastError());
        h->parity ^= 1;
        h->cycle = cycle;
        h->pending = -1;
    });
}

int exaca_check(exaca_state *h, int *cycle, int layer, uint64_t counts[5], int *xswitch, int verbose) {
    return guarded([&] {
        if (!h || !cycle || !xswitch) throw std::invalid_argument("null argument");
        device_check(h, cycle, layer, counts, xswitch, verbose);
    });
}

int exaca_run_layer(exaca_state *h, int layer, int max_cycles, int verbose, int *cycle, double *gpu_ms) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null handle");
        run_layer(h, layer, max_cycles, verbose, cycle, gpu_ms);
    });
}

int exaca_download(exaca_state *h, int field, void *dst, size_t dst_bytes) {
    return guarded([&] {
        if (!h || !dst) throw std::invalid_argument("null argument");
        if (field == EXACA_FIELD_MELTED) {  // int32 per cell, from the device's flags
            const size_t n = static_cast<size_t>(h->host.ncell);
            if (dst_bytes < 4 * n) throw std::invalid_argument("destination too small");
            std::vector<uint8_t> m(n);
            HIP_OK(hipStreamSynchronize(h->stream));
            HIP_OK(hipMemcpy(m.data(), h->melted, n, hipMemcpyDeviceToHost));
            int32_t *o = static_cast<int32_t *>(dst);
            for (size_t q = 0; q < n; q++) o[q] = m[q];
            return;
        }
        if (field == EXACA_FIELD_NUCLEI_LOC || field == EXACA_FIELD_NUCLEI_TIME) {
            // the sites as initialised (the device marks decided sites in its own copy)
            const std::vector<int32_t> &v = field == EXACA_FIELD_NUCLEI_LOC ? h->host.nloc : h->host.ntime;
            if (dst_bytes < 4 * v.size()) throw std::invalid_argument("destination too small");
            if (!v.empty()) std::memcpy(dst, v.data(), 4 * v.size());
            return;
        }
        exaca::download_field(h->dv, field, dst, dst_bytes, h->stream);
    });
}

int exaca_grain_stats(exaca_state *h, const int32_t region[6], int64_t counts[5], double values[4], int64_t hist[64]) {
    return guarded([&] {
        if (!h || !region) throw std::invalid_argument("null argument");
        if (h->pending >= 0) throw StateError("grain_stats called between nucleation and capture");
        const exaca::GrainRegion r = checked_region(h, region);
        grain_id_range(h);
        // device buffers of this call: GrainID flags, the counters and the orientation bins
        struct Free {
            std::vector<void *> p;
            ~Free() {
                for (void *q : p) (void)hipFree(q);
            }
        } tmp;
        StatsBuffers b{};
        HIP_OK(hipMalloc(&b.gflag, 4ull * h->ngid));
        tmp.p.push_back(b.gflag);
        HIP_OK(hipMalloc(&b.cnt, 8 * 4));
        tmp.p.push_back(b.cnt);
        HIP_OK(hipMalloc(&b.bins, 4ull * 2 * exaca::kNumOrientations));
        tmp.p.push_back(b.bins);
        grain_stats_pass(h, r, b);
        std::vector<unsigned int> ob(2 * exaca::kNumOrientations);
        unsigned long long c[4];
        HIP_OK(hipMemcpyAsync(ob.data(), b.bins, 4ull * ob.size(), hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipMemcpyAsync(c, b.cnt, sizeof c, hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        grain_stats_finish(h, region, c, ob.data(), counts, values, hist);
    });
}

int exaca_grain_stats_noalloc(exaca_state *h, const int32_t region[6], int64_t counts[5], double values[4],
                              int64_t hist[64]) {
    return guarded([&] {
        if (!h || !region) throw std::invalid_argument("null argument");
        if (h->pending >= 0) throw StateError("grain_stats called between nucleation and capture");
        const exaca::GrainRegion r = checked_region(h, region);
        grain_id_range(h);
        if (h->ngid > kScratchMaxGrainIds)
            throw std::invalid_argument("exaca_grain_stats_noalloc: the domain's GrainID range holds " +
                                        std::to_string(h->ngid) + " values, more than the " +
                                        std::to_string(kScratchMaxGrainIds) +
                                        " its static scratch covers; use exaca_grain_stats");
        DeviceScope on(h->device);
        char *base = static_cast<char *>(scratch_address());
        StatsBuffers b{};
        b.cnt = reinterpret_cast<unsigned long long *>(base + kScratchCountsOff);
        b.bins = reinterpret_cast<unsigned int *>(base + kScratchBinsOff);
        b.gflag = reinterpret_cast<uint32_t *>(base + kScratchFlagsOff);
        grain_stats_pass(h, r, b);
        // host copies into per-thread static arrays: this entry point allocates nothing
        static thread_local unsigned int ob[2 * exaca::kNumOrientations];
        static thread_local unsigned long long c[4];
        HIP_OK(hipMemcpyAsync(ob, b.bins, sizeof ob, hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipMemcpyAsync(c, b.cnt, sizeof c, hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        grain_stats_finish(h, region, c, ob, counts, values, hist);
    });
}

int exaca_next_layer(exaca_state *h, int layer, int verbose) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null handle");
        next_layer(h, layer, verbose);
    });
}

int exaca_query(const exaca_state *h, int64_t out[20]) {
    return guarded([&] {
        if (!h || !out) throw std::invalid_argument("null argument");
        exaca::StepStats st;
        exaca::StepCtl c;
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipMemcpy(&st, h->stats, sizeof st, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(&c, h->ctl, sizeof c, hipMemcpyDeviceToHost));
        const exaca::HostDomain &hd = h->host;
        out[0] = hd.X;
        out[1] = hd.Y;
        out[2] = hd.nz;
        out[3] = hd.nzA;
        out[4] = hd.nx;
        out[5] = hd.ny;
        out[6] = static_cast<int64_t>(hd.nloc.size());
        out[7] = h->cycle;
        out[8] = static_cast<int64_t>(st.nucleated);
        out[9] = static_cast<int64_t>(st.frontier);
        out[10] = static_cast<int64_t>(st.captures);
        out[11] = static_cast<int64_t>(st.solidified);
        out[12] = static_cast<int64_t>(st.waiting);
        out[13] = c.count[h->parity];
        out[14] = h->dv.list_cap;
        out[15] = static_cast<int64_t>(h->bytes);
        out[16] = static_cast<int64_t>(st.evaluated);
        out[17] = static_cast<int64_t>(st.masks);
        out[18] = h->deck.num_layers;
        out[19] = h->layer;
    });
}

int exaca_destroy(exaca_state *h) {
    return guarded([&] { delete h; });
}

int exaca_profile_steps(exaca_state *h, int layer, int nsteps, double ms[3], int64_t counts[5]) {
    return guarded([&] {
        if (!h || !ms || !counts || nsteps < 0) throw std::invalid_argument("bad argument");
        if (h->pending >= 0) throw StateError("profile_steps called between nucleation and capture");
        double all[exaca::kTimedKernels];
        int64_t launches = 0;
        profile(h, layer, nsteps, all, counts, &launches);
        ms[0] = all[0];
        ms[1] = all[1];
        ms[2] = 0;
        for (int k = 2; k < exaca::kTimedKernels; k++) ms[2] += all[k];
    });
}

int exaca_profile_kernels(exaca_state *h, int layer, int nsteps, double ms[6], int64_t counts[5]) {
    return guarded([&] {
        if (!h || !ms || !counts || nsteps < 0) throw std::invalid_argument("bad argument");
        if (h->pending >= 0) throw StateError("profile_kernels called between nucleation and capture");
        int64_t launches = 0;
        profile(h, layer, nsteps, ms, counts, &launches);
    });
}

int exaca_copy_probe(size_t bytes, int reps, double *ms) {
    return guarded([&] {
        if (!ms || reps <= 0 || bytes < 16) throw std::invalid_argument("bad argument");
        const size_t n = bytes / 16;
        uint4 *a = nullptr, *b = nullptr;
        HIP_OK(hipMalloc(&a, n * 16));
        HIP_OK(hipMalloc(&b, n * 16));
        HIP_OK(hipMemset(a, 1, n * 16));
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        exaca::launch_copy(a, b, n, nullptr);
        HIP_OK(hipEventRecord(e0, nullptr));
        for (int r = 0; r < reps; r++) exaca::launch_copy(a, b, n, nullptr);
        HIP_OK(hipEventRecord(e1, nullptr));
        HIP_OK(hipEventSynchronize(e1));
        float t = 0;
        HIP_OK(hipEventElapsedTime(&t, e0, e1));
        *ms = t / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(a);
        (void)hipFree(b);
    });
}

int exaca_geometry_probe(exaca_state *h, const double *in, double *out, int n) {
    return guarded([&] {
        if (!h || !in || !out || n <= 0) throw std::invalid_argument("bad argument");
        for (int r = 0; r < n; r++) {
            const double o = in[7 * r];
            if (!(o >= 0 && o < exaca::kNumOrientations)) throw std::invalid_argument("orientation out of range");
        }
        double *din = nullptr, *dout = nullptr;
        HIP_OK(hipMalloc(&din, sizeof(double) * 7 * n));
        HIP_OK(hipMalloc(&dout, sizeof(double) * 30 * n));
        HIP_OK(hipMemcpy(din, in, sizeof(double) * 7 * n, hipMemcpyHostToDevice));
        exaca::launch_geometry_probe(h->dv.guv, din, dout, n, h->stream);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipMemcpy(out, dout, sizeof(double) * 30 * n, hipMemcpyDeviceToHost));
        (void)hipFree(din);
        (void)hipFree(dout);
    });
}

int exaca_host_state(const char *input_file, const char *data_dir, int field, void *dst, size_t dst_bytes,
                     int64_t dims[8]) {
    return guarded([&] {
        if (!input_file || !dims) throw std::invalid_argument("null argument");
        const exaca::Deck deck = exaca::read_deck(input_file, data_dir ? data_dir : exaca::data_dir());
        const exaca::HostDomain hd = exaca::initialize(deck, false);
        dims[0] = hd.X;
        dims[1] = hd.Y;
        dims[2] = hd.nz;
        dims[3] = hd.nzA;
        dims[4] = hd.nx;
        dims[5] = hd.ny;
        dims[6] = static_cast<int64_t>(hd.nloc.size());
        dims[7] = static_cast<int64_t>(hd.active.size());
        if (!dst) return;
        const size_t n = static_cast<size_t>(hd.ncell);
        auto put = [&](const int32_t *src, size_t count) {
            if (dst_bytes < 4 * count) throw std::invalid_argument("destination too small");
            std::memcpy(dst, src, 4 * count);
        };
        std::vector<int32_t> t;
        switch (field) {
        case EXACA_FIELD_CELLTYPE:
            t.assign(hd.ct.begin(), hd.ct.end());
            put(t.data(), n);
            break;
        case EXACA_FIELD_GRAINID: put(hd.gid.data(), n); break;
        case EXACA_FIELD_CRITTIMESTEP: put(hd.cts.data(), n); break;
        case EXACA_FIELD_LAYERID:
            t.assign(hd.lid.begin(), hd.lid.end());
            put(t.data(), n);
            break;
        case EXACA_FIELD_NUCLEI_LOC: put(hd.nloc.data(), hd.nloc.size()); break;
        case EXACA_FIELD_NUCLEI_TIME: put(hd.ntime.data(), hd.ntime.size()); break;
        case EXACA_FIELD_MELTED:
            t.assign(hd.melted.begin(), hd.melted.end());
            put(t.data(), n);
            break;
        default: throw std::invalid_argument("field not available from the host state");
        }
    });
}

int exaca_run_program(int id, int np, const char *input_file) {
    return guarded([&] {
        if (!input_file) throw std::invalid_argument("Error: Must provide path to input file on the command line.");
        if (np != 1) throw std::invalid_argument("one GPU holds the whole domain: np must be 1");
        (void)id;
        auto t0 = std::chrono::steady_clock::now();
        std::unique_ptr<exaca_state> h(create_state(input_file, exaca::data_dir(), true));
        std::cout << "Data initialized: Time spent: " << h->init_seconds << " s" << std::endl;
        int cycle = 0;
        double ms = 0;
        auto t1 = std::chrono::steady_clock::now();
        const int layers = h->deck.num_layers;
        for (int layer = 0; layer < layers; layer++) {
            auto tl = std::chrono::steady_clock::now();
            double lms = 0;
            run_layer(h.get(), layer, 0, 1, &cycle, &lms);
            ms += lms;
            const double took = std::chrono::duration<double>(std::chrono::steady_clock::now() - tl).count();
            if (layer + 1 < layers) {
                next_layer(h.get(), layer, 1);
                std::cout << "Time for layer number " << layer << " was " << took << " s, starting layer "
                          << layer + 1 << std::endl;
            }
        }
        auto t2 = std::chrono::steady_clock::now();
        exaca::RunTimes times;
        times.init = std::chrono::duration<double>(t1 - t0).count();
        times.run = std::chrono::duration<double>(t2 - t1).count();
        times.capture = ms / 1000.0;
        std::cout << "Time for final layer was " << times.run << " s" << std::endl;
        const exaca::Deck &d = h->deck;
        if (d.print_misorientation || d.print_full) {
            std::cout << "Collecting data on rank 0 and printing to files" << std::endl;
            std::vector<int32_t> gid(static_cast<size_t>(h->dv.ncell));
            exaca::download_field(h->dv, exaca::kFieldGrainId, gid.data(), gid.size() * 4, h->stream);
            host_fields_for_output(h.get());
            exaca::write_vtk_files(d, h->host, gid);
        }
        else {
            std::cout << "No output files to be printed, exiting program" << std::endl;
        }
        times.out = std::chrono::duration<double>(std::chrono::steady_clock::now() - t2).count();
        exaca::write_log(d, h->host, cycle, times);
        std::cout << "===================================================================================" << std::endl;
        std::cout << "Having run with = 1 processors" << std::endl;
        std::cout << "Output written at cycle = " << cycle << std::endl;
        std::cout << "Total time = " << times.init + times.run + times.out << std::endl;
        std::cout << "Time spent initializing data = " << times.init << " s" << std::endl;
        std::cout << "Time spent performing CA calculations = " << times.run << " s" << std::endl;
        std::cout << "Time spent collecting and printing output data = " << times.out << " s\n" << std::endl;
    });
}

}  // extern "C"
