// ensemble.hip -- many CA domains resident on one GPU, stepped together.
//
// Each domain occupies one slot: its state lives in a sub-range of a single device arena, and its
// DomainView sits in a device-resident slot table. One step of every slot is four launches whose
// grid row y serves slot y (launch_step_n); 50-step chunks of those launches are captured once in a
// HIP graph and replayed. Every domain keeps its own cycle, control block and frontier, and stops
// itself on the device at the first 1000-step check with no liquid left, so each one evolves
// exactly as it would alone on the plain path. The host admits new domains into free slots
// between chunks (continuous batching) and polls the per-slot status words after each chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "../../include/exaca_ca.h"
#include "ca_common.hpp"
#include "ca_kernels.hpp"
#include "deck.hpp"
#include "init.hpp"
#include "runtime_util.hpp"

namespace exaca {
void upload_cells(const DomainView &d, const HostDomain &hd, uint8_t *melted, hipStream_t s);
StepCtl initial_ctl(const HostDomain &hd);
void download_field(const DomainView &d, int field, void *dst, size_t dst_bytes, hipStream_t s);
}  // namespace exaca

namespace {

using exaca::DeviceError;
using exaca::StateError;
using exaca::guarded;

constexpr int kChunk = 50;         // steps per graph replay (even: parity 0 at chunk boundaries)
constexpr int kMaxSlices = 256;    // largest workgroups-per-slot candidate
constexpr size_t kAlign = 256;

size_t up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over the arena's byte range.
struct Arena {
    char *base = nullptr;
    size_t bytes = 0;
    std::map<size_t, size_t> free;  // offset -> length
    bool take(size_t n, size_t &off) {
        for (auto it = free.begin(); it != free.end(); ++it)
            if (it->second >= n) {
                off = it->first;
                const size_t rest = it->second - n;
                free.erase(it);
                if (rest) free[off + n] = rest;
                return true;
            }
        return false;
    }
    void give(size_t off, size_t n) {
        auto it = free.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free.erase(nx);
        }
        if (it != free.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free.erase(it);
            }
        }
    }
    size_t available() const {
        size_t s = 0;
        for (auto &kv : free) s += kv.second;
        return s;
    }
};

// Device bytes one domain needs (state, frontier lists, tables, control).
size_t domain_bytes(const exaca::HostDomain &h, size_t ngrow) {
    const size_t n = static_cast<size_t>(h.ncell);
    size_t b = 0;
    b += up(sizeof(exaca::Cell) * n) + up(sizeof(exaca::Octa) * n);  // cell records
    b += 4 * up(4 * n) + up(4 * kMaxSlices);                            // frontier, claim lists
    b += up(4ull * exaca::apply_blocks(kMaxSlices) * exaca::kCandStage) + up(4 * exaca::apply_blocks(kMaxSlices));
    b += 2 * up(4 * std::max<size_t>(h.nloc.size(), 1)) + up(8 * ngrow);
    b += up(sizeof(exaca::StepCtl)) + up(sizeof(exaca::StepStats)) + up(8ull * exaca::kPartRows * kMaxSlices);
    const size_t nwords = (n + 31) / 32;
    b += up(4 * exaca::kAlarmSlots * exaca::alarm_stride(h.ncell)) + up(8 * ((nwords + 2047) / 2048));  // calendar, chunk flags
    return b;
}

}  // namespace

// A domain initialised on the host, ready to be admitted.
struct exaca_prepared {
    exaca::Deck deck;
    exaca::HostDomain host;
    std::vector<double> grow;
    long long liquid0 = 0;
    size_t bytes = 0;
};

struct exaca_ensemble {
    hipStream_t stream = nullptr;
    int slots = 0;
    int blocks = 64;      // workgroups per slot (frontier kernels)
    int nuc_blocks = 4;   // workgroups per slot (nucleation)
    Arena arena;
    exaca::DomainView *views_d = nullptr;
    std::vector<exaca::DomainView> views;
    int32_t *status_d = nullptr;  // per slot: final cycle (slots words), then error flags (slots words)
    std::vector<int32_t> status;  // both, as last polled
    std::vector<size_t> off, len;
    std::vector<char> used;
    std::vector<exaca::StepStats *> stats_d;
    // tables shared by every slot (one orientation file per ensemble)
    std::string orientation_file;
    float *guv = nullptr, *seed_cdl = nullptr;
    std::map<int, hipGraphExec_t> graphs;  // per workgroups-per-slot setting
    long long steps = 0;

    ~exaca_ensemble() {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        if (arena.base) (void)hipFree(arena.base);
        if (views_d) (void)hipFree(views_d);
        if (status_d) (void)hipFree(status_d);
        if (guv) (void)hipFree(guv);
        if (seed_cdl) (void)hipFree(seed_cdl);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void ensure_tables(exaca_ensemble *e, const exaca::HostDomain &h, const std::string &orientation_file) {
    using namespace exaca;
    if (e->guv) {
        if (orientation_file != e->orientation_file)
            throw std::invalid_argument("all domains of an ensemble share one orientation file");
        return;
    }
    e->orientation_file = orientation_file;
    std::vector<float> seed(static_cast<size_t>(kCdlStride) * kNumOrientations, 0.0f);
    for (int o = 0; o < kNumOrientations; o++) seed_critical_lengths(&h.guv[9 * o], &seed[kCdlStride * o]);
    HIP_OK(hipMalloc(&e->guv, h.guv.size() * 4));
    HIP_OK(hipMalloc(&e->seed_cdl, seed.size() * 4));
    HIP_OK(hipMemcpy(e->guv, h.guv.data(), h.guv.size() * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(e->seed_cdl, seed.data(), seed.size() * 4, hipMemcpyHostToDevice));
}

hipGraphExec_t graph_for(exaca_ensemble *e, int blocks) {
    auto it = e->graphs.find(blocks);
    if (it != e->graphs.end()) return it->second;
    hipGraph_t g;
    HIP_OK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < kChunk; k++)
        exaca::launch_step_n(e->views_d, e->slots, k & 1, 0, e->nuc_blocks, blocks, e->stream);
    HIP_OK(hipStreamEndCapture(e->stream, &g));
    hipGraphExec_t x;
    HIP_OK(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    HIP_OK(hipGraphDestroy(g));
    e->graphs[blocks] = x;
    return x;
}

void admit(exaca_ensemble *e, const exaca_prepared *pr, int *slot_out) {
    using namespace exaca;
    ensure_tables(e, pr->host, pr->deck.orientation_file);
    int slot = -1;
    for (int s = 0; s < e->slots; s++)
        if (!e->used[s]) {
            slot = s;
            break;
        }
    if (slot < 0) throw StateError("no free slot");
    size_t off = 0;
    if (!e->arena.take(pr->bytes, off)) throw StateError("device memory budget exhausted");
    const HostDomain &hd = pr->host;
    const size_t n = static_cast<size_t>(hd.ncell);
    char *p = e->arena.base + off;
    auto carve = [&](size_t bytes) {
        char *q = p;
        p += up(bytes);
        return q;
    };
    DomainView d{};
    d.X = hd.X;
    d.Y = hd.Y;
    d.nz = hd.nz;
    d.nzA = hd.nzA;
    d.xoff = hd.xoff;
    d.yoff = hd.yoff;
    d.ncell = hd.ncell;
    d.cells = reinterpret_cast<Cell *>(carve(sizeof(Cell) * n));
    d.octa = reinterpret_cast<Octa *>(carve(sizeof(Octa) * n));
    d.list[0] = reinterpret_cast<uint32_t *>(carve(4 * n));
    d.list[1] = reinterpret_cast<uint32_t *>(carve(4 * n));
    d.claim_cell = reinterpret_cast<uint32_t *>(carve(4 * n));
    d.claim_bits = reinterpret_cast<uint32_t *>(carve(4 * n));
    d.slice_claims = reinterpret_cast<uint32_t *>(carve(4 * kMaxSlices));
    d.cand = reinterpret_cast<uint32_t *>(carve(4ull * exaca::apply_blocks(kMaxSlices) * exaca::kCandStage));
    d.cand_count = reinterpret_cast<uint32_t *>(carve(4 * exaca::apply_blocks(kMaxSlices)));
    d.cand_cap = exaca::cand_stage_cap();
    d.list_cap = static_cast<uint32_t>(n);
    int32_t *nloc = reinterpret_cast<int32_t *>(carve(4 * std::max<size_t>(hd.nloc.size(), 1)));
    int32_t *ntime = reinterpret_cast<int32_t *>(carve(4 * std::max<size_t>(hd.nloc.size(), 1)));
    double *grow = reinterpret_cast<double *>(carve(8 * pr->grow.size()));
    const size_t nwords = (n + 31) / 32;
    d.nchunks = static_cast<uint32_t>((nwords + 2047) / 2048);
    d.alarms = reinterpret_cast<uint32_t *>(carve(4 * kAlarmSlots * alarm_stride(d.ncell)));
    d.chunk_flag = reinterpret_cast<unsigned long long *>(carve(8ull * d.nchunks));
    d.ctl = reinterpret_cast<StepCtl *>(carve(sizeof(StepCtl)));
    d.stats = reinterpret_cast<StepStats *>(carve(sizeof(StepStats)));
    d.nparts = kMaxSlices;
    d.part = reinterpret_cast<unsigned long long *>(carve(8ull * kPartRows * kMaxSlices));
    d.nnuclei = static_cast<int32_t>(hd.nloc.size());
    d.nloc = nloc;
    d.ntime = ntime;
    d.guv = e->guv;
    d.seed_cdl = e->seed_cdl;
    d.grow = grow;
    d.ngrow = static_cast<int32_t>(pr->grow.size());
    d.ksat = exaca::growth_plateau(pr->grow);
    d.autostop = 1;
    d.liquid0 = pr->liquid0;
    d.status = e->status_d + slot;
    d.status_err = e->status_d + e->slots + slot;

    hipStream_t s = e->stream;
    if (!hd.nloc.empty()) {
        HIP_OK(hipMemcpyAsync(nloc, hd.nloc.data(), 4 * hd.nloc.size(), hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(ntime, hd.ntime.data(), 4 * hd.ntime.size(), hipMemcpyHostToDevice, s));
    }
    HIP_OK(hipMemcpyAsync(grow, pr->grow.data(), 8 * pr->grow.size(), hipMemcpyHostToDevice, s));
    exaca::upload_cells(d, hd, nullptr, s);
    // Chunks end on parity 0, so a new domain always starts in slot 0 of its control block.
    StepCtl c = exaca::initial_ctl(hd);
    HIP_OK(hipMemcpyAsync(d.ctl, &c, sizeof c, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemsetAsync(d.stats, 0, sizeof(StepStats), s));
    HIP_OK(hipMemsetAsync(d.part, 0, 8ull * kPartRows * kMaxSlices, s));
    const int32_t zero = 0;
    HIP_OK(hipMemcpyAsync(e->status_d + slot, &zero, 4, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(e->status_d + e->slots + slot, &zero, 4, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(e->views_d + slot, &d, sizeof d, hipMemcpyHostToDevice, s));
    HIP_OK(hipStreamSynchronize(s));  // the host-side sources above are temporaries
    e->views[slot] = d;
    e->off[slot] = off;
    e->len[slot] = pr->bytes;
    e->used[slot] = 1;
    e->status[slot] = 0;
    e->status[e->slots + slot] = 0;
    *slot_out = slot;
}

void release(exaca_ensemble *e, int slot) {
    if (slot < 0 || slot >= e->slots || !e->used[slot]) throw std::invalid_argument("slot not in use");
    exaca::DomainView empty{};
    HIP_OK(hipMemcpyAsync(e->views_d + slot, &empty, sizeof empty, hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    e->arena.give(e->off[slot], e->len[slot]);
    e->views[slot] = empty;
    e->used[slot] = 0;
    e->status[slot] = -1;
    e->status[e->slots + slot] = 0;
}

void advance(exaca_ensemble *e, int steps) {
    const int chunks = (steps + kChunk - 1) / kChunk;
    hipGraphExec_t g = graph_for(e, e->blocks);
    for (int c = 0; c < chunks; c++) HIP_OK(hipGraphLaunch(g, e->stream));
    e->steps += static_cast<long long>(chunks) * kChunk;
}

void poll(exaca_ensemble *e) {
    HIP_OK(hipMemcpyAsync(e->status.data(), e->status_d, 8 * e->slots, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
    for (int s = 0; s < e->slots; s++)
        if (!e->used[s]) {
            e->status[s] = -1;
            e->status[e->slots + s] = 0;
        }
}

unsigned long long frontier_total(exaca_ensemble *e) {
    unsigned long long t = 0;
    for (int s = 0; s < e->slots; s++) {
        if (!e->used[s]) continue;
        exaca::StepStats st;
        HIP_OK(hipMemcpy(&st, e->views[s].stats, sizeof st, hipMemcpyDeviceToHost));
        t += st.frontier;
    }
    return t;
}

}  // namespace

extern "C" {

int exaca_prepare(const char *input_file, const char *data_dir, uint64_t seed, exaca_prepared **out) {
    return guarded([&] {
        if (!input_file || !out) throw std::invalid_argument("null argument");
        *out = nullptr;
        auto pr = std::make_unique<exaca_prepared>();
        pr->deck = exaca::read_deck(input_file, data_dir ? data_dir : exaca::data_dir());
        if (pr->deck.sim_type != "C")
            throw std::invalid_argument("ensembles hold directional-solidification (type C) domains");
        pr->deck.seed = seed;
        pr->host = exaca::initialize_sparse(pr->deck, false);
        const exaca::HostDomain &h = pr->host;
        pr->grow = exaca::growth_table(h.A, h.B, h.C, h.D, h.dU, 1 << 25);
        if (pr->grow.size() >= (1u << 25)) throw std::runtime_error("undercooling never saturates within 2^25 steps");
        if (exaca::rebuild_units(h.ncell) > exaca::kEnsembleUnits)
            throw std::invalid_argument("domain too large for an ensemble slot (at most " +
                                        std::to_string(exaca::kEnsembleUnits) +
                                        " list-rebuild units, about 33 M cells); run it on the plain path");
        if (h.sparse)  // every cell inside the walls is liquid but the substrate's active cells
            pr->liquid0 = static_cast<long long>(h.nx) * h.ny * (h.nz - 2) - static_cast<long long>(h.active.size());
        else
            for (long long q = 0; q < h.ncell; q++) pr->liquid0 += (h.ct[q] == exaca::kLiquid && h.lid[q] == 0);
        pr->bytes = domain_bytes(h, pr->grow.size());
        *out = pr.release();
    });
}

int exaca_prepared_info(const exaca_prepared *pr, int64_t out[4]) {
    return guarded([&] {
        if (!pr || !out) throw std::invalid_argument("null argument");
        out[0] = pr->host.ncell;
        out[1] = static_cast<int64_t>(pr->bytes);
        out[2] = static_cast<int64_t>(pr->host.nloc.size());
        out[3] = pr->liquid0;
    });
}

int exaca_prepared_free(exaca_prepared *pr) {
    return guarded([&] { delete pr; });
}

int exaca_ensemble_create(int slots, size_t arena_bytes, exaca_ensemble **out) {
    return guarded([&] {
        if (!out || slots <= 0 || slots > 65535 || arena_bytes < kAlign) throw std::invalid_argument("bad argument");
        *out = nullptr;
        auto e = std::make_unique<exaca_ensemble>();
        e->slots = slots;
        HIP_OK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
        HIP_OK(hipMalloc(&e->arena.base, arena_bytes));
        e->arena.bytes = arena_bytes;
        e->arena.free[0] = arena_bytes;
        HIP_OK(hipMalloc(&e->views_d, sizeof(exaca::DomainView) * slots));
        HIP_OK(hipMemset(e->views_d, 0, sizeof(exaca::DomainView) * slots));
        HIP_OK(hipMalloc(&e->status_d, 8 * slots));
        HIP_OK(hipMemset(e->status_d, 0, 8 * slots));
        e->views.assign(slots, exaca::DomainView{});
        e->status.assign(2 * static_cast<size_t>(slots), -1);
        std::fill(e->status.begin() + slots, e->status.end(), 0);
        e->off.assign(slots, 0);
        e->len.assign(slots, 0);
        e->used.assign(slots, 0);
        *out = e.release();
    });
}

int exaca_ensemble_admit(exaca_ensemble *e, const exaca_prepared *pr, int *slot) {
    return guarded([&] {
        if (!e || !pr || !slot) throw std::invalid_argument("null argument");
        admit(e, pr, slot);
    });
}

int exaca_ensemble_release(exaca_ensemble *e, int slot) {
    return guarded([&] {
        if (!e) throw std::invalid_argument("null handle");
        release(e, slot);
    });
}

int exaca_ensemble_advance(exaca_ensemble *e, int steps) {
    return guarded([&] {
        if (!e || steps < 0) throw std::invalid_argument("bad argument");
        advance(e, steps);
    });
}

int exaca_ensemble_poll(exaca_ensemble *e, int32_t *final_cycles, int32_t *errors) {
    return guarded([&] {
        if (!e || !final_cycles) throw std::invalid_argument("null argument");
        poll(e);
        std::copy(e->status.begin(), e->status.begin() + e->slots, final_cycles);
        if (errors) std::copy(e->status.begin() + e->slots, e->status.end(), errors);
    });
}

int exaca_ensemble_set_blocks(exaca_ensemble *e, int blocks) {
    return guarded([&] {
        if (!e || blocks < 8 || blocks > kMaxSlices || blocks % 8) throw std::invalid_argument("bad workgroup count");
        e->blocks = blocks;
    });
}

int exaca_ensemble_autotune(exaca_ensemble *e, int chunks_per_candidate, int *chosen) {
    return guarded([&] {
        if (!e || !chosen || chunks_per_candidate <= 0) throw std::invalid_argument("bad argument");
        // Online tuning on the live ensemble: the workgroup count per slot changes only the
        // schedule, never the result, so candidates are tried on consecutive chunks and scored by
        // device time per active-cell update.
        const int candidates[] = {16, 32, 64, 128, 256};
        double best = 1e300;
        int best_b = e->blocks;
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        for (int b : candidates) {
            hipGraphExec_t g = graph_for(e, b);
            HIP_OK(hipStreamSynchronize(e->stream));
            const unsigned long long f0 = frontier_total(e);
            HIP_OK(hipEventRecord(e0, e->stream));
            for (int c = 0; c < chunks_per_candidate; c++) HIP_OK(hipGraphLaunch(g, e->stream));
            HIP_OK(hipEventRecord(e1, e->stream));
            HIP_OK(hipEventSynchronize(e1));
            e->steps += static_cast<long long>(chunks_per_candidate) * kChunk;
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            const unsigned long long f1 = frontier_total(e);
            const double score = ms / static_cast<double>(std::max<unsigned long long>(f1 - f0, 1));
            if (score < best) {
                best = score;
                best_b = b;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        e->blocks = best_b;
        *chosen = best_b;
    });
}

int exaca_ensemble_query(exaca_ensemble *e, int slot, int64_t out[8]) {
    return guarded([&] {
        if (!e || !out || slot < 0 || slot >= e->slots || !e->used[slot]) throw std::invalid_argument("bad slot");
        exaca::StepStats st;
        exaca::StepCtl c;
        HIP_OK(hipStreamSynchronize(e->stream));
        HIP_OK(hipMemcpy(&st, e->views[slot].stats, sizeof st, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(&c, e->views[slot].ctl, sizeof c, hipMemcpyDeviceToHost));
        out[0] = c.done ? c.final_cycle : 0;
        out[1] = static_cast<int64_t>(st.nucleated);
        out[2] = static_cast<int64_t>(st.frontier);
        out[3] = static_cast<int64_t>(st.captures);
        out[4] = static_cast<int64_t>(st.solidified);
        out[5] = c.error;
        out[6] = static_cast<int64_t>(e->arena.available());
        out[7] = e->steps;
    });
}

int exaca_ensemble_download(exaca_ensemble *e, int slot, int field, void *dst, size_t dst_bytes) {
    return guarded([&] {
        if (!e || !dst || slot < 0 || slot >= e->slots || !e->used[slot]) throw std::invalid_argument("bad slot");
        exaca::download_field(e->views[slot], field, dst, dst_bytes, e->stream);
    });
}

int exaca_ensemble_destroy(exaca_ensemble *e) {
    return guarded([&] { delete e; });
}

}  // extern "C"
