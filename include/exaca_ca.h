/*
 * exaca_ca.h -- C ABI of the MI355X-native ExaCA cellular-automaton time step.
 *
 * Plain C types only: an opaque handle, integers, pointers and sizes. Every entry point returns
 * 0 on success and a nonzero status on failure (the message is available from
 * exaca_last_error()); no exception crosses this boundary. One host thread per handle; each
 * handle owns one HIP stream on the current device.
 *
 * Each declaration names the reference interface it replaces (paths relative to the ExaCA
 * repository root).
 */
#ifndef EXACA_CA_H
#define EXACA_CA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct exaca_state exaca_state;

enum exaca_status {
    EXACA_OK = 0,
    EXACA_ERR_INPUT = 1,   /* bad deck / file / argument (the reference throws std::runtime_error) */
    EXACA_ERR_DEVICE = 2,  /* HIP runtime failure or device-side error flag */
    EXACA_ERR_STATE = 3,   /* call out of order (e.g. capture without nucleation for that cycle) */
};

/* Field selectors for exaca_download. */
enum exaca_field {
    EXACA_FIELD_CELLTYPE = 0,     /* int32, reference TypeNames values (src/CAtypes.hpp:11) */
    EXACA_FIELD_GRAINID = 1,      /* int32 */
    EXACA_FIELD_DIAGONAL = 2,     /* float, DiagonalLength */
    EXACA_FIELD_DOCENTER = 3,     /* float[3] per cell, DOCenter (x, y, z interleaved) */
    EXACA_FIELD_CRITTIMESTEP = 4, /* int32 */
    EXACA_FIELD_LAYERID = 5,      /* int32 */
    EXACA_FIELD_CELLTYPE_RAW = 6, /* uint8, type | 0x10 for cell-centred octahedra */
    EXACA_FIELD_NUCLEI_LOC = 7,   /* int32 per nucleation site (NucleiLocation), host state only */
    EXACA_FIELD_NUCLEI_TIME = 8,  /* int32 per nucleation site (NucleationTimes), host state only */
    EXACA_FIELD_MELTED = 9,       /* int32 per cell, the reference's Melted flag */
    EXACA_FIELD_CENTRE_F64 = 10,  /* double[3] per cell: the octahedron centre as computed at capture
                                     (DOCenter is its float rounding); with CELLTYPE_RAW, GRAINID and
                                     DIAGONAL it is the whole front state of a step */
};

/* Replaces the initialisation half of RunProgram_Reduced (src/runCA.cpp:37-256): reads the deck
 * (InputReadFromFile, src/CAinitialize.cpp:232), builds the np = 1 initial state (TempInit_*,
 * OrientationInit, SubstrateInit_*, ActiveCellWallInit, GrainInit, NucleiInit) on the host and
 * uploads it. data_dir holds Materials/ and Substrate/ (NULL: the library's data directory). */
int exaca_create(const char *input_file, const char *data_dir, int verbose, exaca_state **out);

/* exaca_create with a random-stream seed for the initial state (no reference counterpart: the
 * reference's seeds are fixed). seed 0 is the reference's own streams; ensemble instances use
 * their own seeds so that each can be re-run alone on the plain path. */
int exaca_create_seeded(const char *input_file, const char *data_dir, uint64_t seed, int verbose,
                        exaca_state **out);

/* Nucleation (src/CAupdate.hpp:15 / src/CAupdate.cpp:67) for time step `cycle` of `layer`.
 * The nucleation count (the reference's nn) accumulates on the device; see exaca_query. */
int exaca_nucleation(exaca_state *h, int cycle, int layer);

/* Nucleation with the reference's `int &nn` out-parameter (src/CAupdate.hpp:15): *nn receives the
 * run's nucleation count including this step's. Reads it back from the device, so it waits for
 * the step's nucleation kernel, as the reference's own reduction of the count does
 * (src/CAupdate.cpp:75-199). */
int exaca_nucleation_nn(exaca_state *h, int cycle, int layer, int32_t *nn);

/* CellCapture (src/CAupdate.hpp:21 / src/CAupdate.cpp:204) for the same cycle. Asynchronous:
 * unlike the reference (deep_copy of numSteer, src/CAupdate.cpp:243) nothing is read back. */
int exaca_cell_capture(exaca_state *h, int cycle, int layer);

/* IntermediateOutputAndCheck (src/CAupdate.hpp:31 / src/CAupdate.cpp:732): reduces the five
 * cell counts, sets *xswitch when no liquid is left in the layer, and for spot-melt problems may
 * move *cycle forward to the step before the next liquidus crossing. counts may be NULL. */
int exaca_check(exaca_state *h, int *cycle, int layer, uint64_t counts[5], int *xswitch, int verbose);

/* The layer's do-while loop of RunProgram_Reduced (src/runCA.cpp:290-342) on the device:
 * Nucleation + CellCapture every step (replayed from captured HIP graphs), the check every 1000
 * steps. Stops at completion or after max_cycles steps (<= 0: no limit). *cycle receives the
 * final cycle; *gpu_ms (may be NULL) the device time of the loop. */
int exaca_run_layer(exaca_state *h, int layer, int max_cycles, int verbose, int *cycle, double *gpu_ms);

/* Device -> host copy of one whole-domain field (kind: enum exaca_field) into dst of
 * dst_bytes bytes. Replaces the deep_copy at src/runCA.cpp:423-424. NUCLEI_LOC and NUCLEI_TIME
 * give the handle's nucleation sites as initialised (int32 per site). */
int exaca_download(exaca_state *h, int field, void *dst, size_t dst_bytes);

/* Sizes and counters. out[0..19]: X, Y, nz, nzActive, nx, ny, nuclei, last cycle,
 * nucleation events (nn), sum of frontier sizes, captures, solidifications, waiting-entry steps,
 * frontier list length, frontier list capacity, device bytes allocated, critical-length
 * evaluations, liquid-neighbour masks built, number of layers, current layer. */
int exaca_query(const exaca_state *h, int64_t out[20]);

/* The transition after layer `layer` (src/runCA.cpp:343-418): DomainShiftAndResize and LayerSetup
 * (src/CAinitialize.cpp:2772-2982) on the resident state, and the cycle count restarting at 0.
 * Call it when exaca_run_layer (or the check) has ended the layer; then run layer + 1. */
int exaca_next_layer(exaca_state *h, int layer, int verbose);

/* Grain statistics of a region, computed on the device from the current GrainID field: the
 * quantities of the reference's analysis executable (analysis/src/GAprint.cpp:17-108,
 * PrintMisorientationData and the head of PrintSizeData) and its misorientation vtk
 * (src/CAprint.cpp:419-432). region[6] = XMin, XMax, YMin, YMax, ZMin, ZMax, inclusive, in the
 * whole-domain coordinates of the reference's vtk files (0 .. nx-1, 0 .. ny-1, 0 .. nz-1). A melted
 * cell of GrainID 0 (a wall of a type-C domain) counts in orientation 0.
 * counts[0..3]: cells in the region, melted cells, unique GrainIDs, cells of nucleated grains
 * (GrainID < 0); counts[4]: melted cells in the top plane ZMax.
 * values[0..3]: mean misorientation to +Z of the melted cells (degrees), the same in the top plane,
 * mean grain volume (deltax^2 * 1e12 * region cells / grains, the reference's printed value),
 * volume fraction of nucleated grains. hist[0..63]: melted cells per misorientation rounded to
 * whole degrees (the misorientation vtk's values). Any output pointer may be NULL. */
int exaca_grain_stats(exaca_state *h, const int32_t region[6], int64_t counts[5], double values[4], int64_t hist[64]);

/* exaca_grain_stats without any device or host allocation (no counterpart in the reference: for
 * callers that cannot allocate, such as code under HIP graph capture or a serving loop that
 * forbids it). Same arguments and results. Its counters, orientation bins and GrainID flags live
 * in one static device array of 4.2 MB per GPU, held until the process ends; domains whose
 * initial GrainID range spans more than 2^20 values are rejected with EXACA_ERR_INPUT. Every call
 * on a GPU shares that GPU's array: calls that overlap on one GPU (several host threads, or
 * handles on streams run concurrently) race and must be serialised by the caller. */
int exaca_grain_stats_noalloc(exaca_state *h, const int32_t region[6], int64_t counts[5], double values[4],
                              int64_t hist[64]);

/* Releases the handle and its device memory. */
int exaca_destroy(exaca_state *h);

/* RunProgram_Reduced(int id, int np, std::string InputFile) (src/runCA.hpp:11): the whole
 * program for one GPU -- read deck, initialise, run, write <base>.vtk, <base>_Misorientations.vtk
 * and <base>.log. np must be 1 (one GPU holds the whole domain). */
int exaca_run_program(int id, int np, const char *input_file);

/* Host-only view of the initial state exaca_create would upload (no device needed): the
 * reference's host arrays after NucleiInit (src/runCA.cpp:90-195). dims[0..7]: X, Y, nz, nzActive,
 * nx, ny, nucleation sites, initially active cells. With dst == NULL only dims is filled; field
 * is one of CELLTYPE, GRAINID, CRITTIMESTEP, LAYERID, NUCLEI_LOC, NUCLEI_TIME. */
int exaca_host_state(const char *input_file, const char *data_dir, int field, void *dst, size_t dst_bytes,
                     int64_t dims[8]);

/* ---- Ensembles: many domains resident on one GPU, stepped together (no reference counterpart;
 * the reference runs one domain per program). Each domain evolves exactly as it does alone on the
 * plain path: its own cycle, frontier and 1000-step termination check, decided on the device. ---- */
typedef struct exaca_prepared exaca_prepared;
typedef struct exaca_ensemble exaca_ensemble;

/* Host-side initialisation of one type-C deck with a random-stream seed (thread-safe, no GPU). */
int exaca_prepare(const char *input_file, const char *data_dir, uint64_t seed, exaca_prepared **out);
/* out[0..3]: cells, device bytes needed, nucleation sites, liquid cells at the start. */
int exaca_prepared_info(const exaca_prepared *pr, int64_t out[4]);
int exaca_prepared_free(exaca_prepared *pr);
/* An ensemble of `slots` domain slots sharing one device arena of arena_bytes (the admission
 * budget). */
int exaca_ensemble_create(int slots, size_t arena_bytes, exaca_ensemble **out);
/* Uploads a prepared domain into a free slot; EXACA_ERR_STATE when no slot or memory is free. */
int exaca_ensemble_admit(exaca_ensemble *e, const exaca_prepared *pr, int *slot);
/* Frees a slot and its memory. */
int exaca_ensemble_release(exaca_ensemble *e, int slot);
/* Enqueues `steps` more steps of every slot (rounded up to 50-step graph replays; asynchronous). */
int exaca_ensemble_advance(exaca_ensemble *e, int steps);
/* Waits for the queued steps; final_cycles[slots] receives each slot's final cycle, 0 while it
 * runs, -1 when empty. errors[slots] (may be NULL) receives each slot's device error flags (the
 * flags of exaca_last_error's device messages; 0: none). A domain that raises one is stopped on
 * the device at the end of that step: its final cycle is then that step and its result is void. */
int exaca_ensemble_poll(exaca_ensemble *e, int32_t *final_cycles, int32_t *errors);
/* Workgroups per slot for the frontier kernels (multiple of 8, 8..256). */
int exaca_ensemble_set_blocks(exaca_ensemble *e, int blocks);
/* Online autotuner: runs chunks_per_candidate chunks per workgroup-count candidate on the live
 * ensemble and keeps the one with the least device time per active-cell update. */
int exaca_ensemble_autotune(exaca_ensemble *e, int chunks_per_candidate, int *chosen);
/* out[0..7]: final cycle (0 while running), nucleation events, frontier-size sum, captures,
 * solidifications, device error flag, free arena bytes, steps enqueued so far. */
int exaca_ensemble_query(exaca_ensemble *e, int slot, int64_t out[8]);
/* One per-cell field of one slot (whole domain, the layout of exaca_download). */
int exaca_ensemble_download(exaca_ensemble *e, int slot, int field, void *dst, size_t dst_bytes);
int exaca_ensemble_destroy(exaca_ensemble *e);

/* Message of the last failure on this thread. */
const char *exaca_last_error(void);

/* Timed stepping (no reference counterpart): runs nsteps further steps of `layer` like
 * exaca_run_layer, without graphs, with HIP events on the handle's stream around every kernel.
 * ms[0..2] receive the summed device time of the nucleation, claim and apply kernels; counts
 * [0..4] the frontier-size sum, captures, solidifications, kernel launches per kind and the number
 * of steps executed (checks and cycle jumps are applied as in exaca_run_layer). */
int exaca_profile_steps(exaca_state *h, int layer, int nsteps, double ms[3], int64_t counts[5]);

/* As exaca_profile_steps with every kernel timed on its own: ms[0..5] = nucleation, claim, the
 * claims' winners, apply, settle and the rebuild of the next step's list. */
int exaca_profile_kernels(exaca_state *h, int layer, int nsteps, double ms[6], int64_t counts[5]);

/* HBM calibration (no reference counterpart): a plain device copy of `bytes` bytes repeated
 * `reps` times; *ms receives the average time of one copy. */
int exaca_copy_probe(size_t bytes, int reps, double *ms);

/* Self test of the device capture geometry (no reference counterpart): n rows of
 * (orientation, cxold, cyold, czold, xp, yp, zp) -> n rows of (cx, cy, cz, dl, cdl[26]),
 * using the orientation table of h. in/out are host arrays of 7n / 30n doubles. */
int exaca_geometry_probe(exaca_state *h, const double *in, double *out, int n);

#ifdef __cplusplus
}
#endif
#endif
