/*
 * exaca_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C, serial restatement of the ExaCA (Kokkos, Aug 2021) solidification path, used as the
 * parity checker for the MI355X implementation. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library; the product never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the reference
 * repository root).
 */
#ifndef EXACA_ORACLE_H
#define EXACA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sim oracle_sim;

/* Field selectors for oracle_get_field (all whole-domain, X*Y*nz, reference linear order). */
enum {
    ORACLE_CELLTYPE = 0,     /* int32 */
    ORACLE_GRAINID = 1,      /* int32 */
    ORACLE_CRITTIMESTEP = 2, /* int32 */
    ORACLE_LAYERID = 3,      /* int32 */
    ORACLE_UNDERCOOLING = 4, /* float, UndercoolingCurrent */
    ORACLE_UCHANGE = 5,      /* float, UndercoolingChange */
    ORACLE_MELTED = 6,       /* int32 0/1 */
    ORACLE_DIAGLEN = 7,      /* float, active domain (X*Y*nzActive) */
    ORACLE_DOCENTER = 8,     /* float[3], active domain */
    ORACLE_CRITDIAG = 9,     /* float[26], active domain */
    ORACLE_NUCLEI_LOC = 10,  /* int32[P] */
    ORACLE_NUCLEI_TIME = 11, /* int32[P] */
};

/* Parse deck + initialise (np = 1). Returns 0 on success; message via oracle_last_error(). */
int oracle_load(const char *deck_path, const char *data_dir, oracle_sim **out);
/* The same with an ensemble instance's random streams (type C): mt19937_64(seed) for the substrate
 * and the nucleation sites, default_random_engine(seed % 2147483646 + 1) for the nucleation
 * undercoolings; seed 0 is oracle_load (the reference's fixed streams). */
int oracle_load_seeded(const char *deck_path, const char *data_dir, uint64_t seed, oracle_sim **out);
/* dims: [0]=X(MyXSlices) [1]=Y(MyYSlices) [2]=nz [3]=nzActive [4]=nx [5]=ny [6]=P (nuclei)
 *       [7]=cycle [8]=nn [9]=sum of frontier sizes [10]=captures [11]=solidifications
 *       [12]=layers [13]=current layer [14]=ZBound_Low [15]=ZBound_High
 *       [16]=nuclei or captures that fell outside the active window (0 in a well-formed run) */
void oracle_dims(const oracle_sim *s, int64_t dims[17]);
/* Scalar parameters: [0]=A [1]=B [2]=C (normalised) [3]=D [4]=deltax [5]=deltat [6]=G [7]=R
 *                    [8]=UndercoolingChange of the first liquid cell */
void oracle_params(const oracle_sim *s, double out[9]);
int oracle_get_field(const oracle_sim *s, int which, void *dst);
/* One time step at `cycle` on layer 0: Nucleation then CellCapture (src/runCA.cpp:295-311). */
void oracle_step(oracle_sim *s, int cycle);
/* IntermediateOutputAndCheck (src/CAupdate.cpp:732-822); counts[5] as in the reference. */
void oracle_check(oracle_sim *s, int *cycle, uint64_t counts[5], int *xswitch, int verbose);
/* The layer loop (src/runCA.cpp:283-342) from the current cycle; with max_cycles == 0 it goes on
 * through every remaining layer (the transitions of src/runCA.cpp:343-418), with max_cycles < 0 it
 * ends with the current layer, otherwise it stops after max_cycles in the current layer. Returns
 * the final cycle (of the last layer run).
 * capture_seconds receives the CellCapture wall time. */
int oracle_run(oracle_sim *s, int max_cycles, int verbose, double *capture_seconds);
/* DomainShiftAndResize + LayerSetup (src/CAinitialize.cpp:2772-2982) and the cycle reset: the next
 * layer's active window and cell-centred octahedra for its active cells. */
int oracle_next_layer(oracle_sim *s);
/* vtk + misorientation vtk + log, formats of src/CAprint.cpp:285-571 (timings written as 0). */
int oracle_write_outputs(const oracle_sim *s, const char *deck_path);
void oracle_free(oracle_sim *s);
const char *oracle_last_error(void);

/* ---- unit-level restatements used to pin the oracle ---- */
void oracle_mt64_draws(uint64_t seed, int n, uint64_t *out);
void oracle_uniform_draws(uint64_t seed, int n, double *out);   /* mt19937_64 + U(0,1) */
void oracle_minstd_draws(int n, uint32_t *out);                 /* default_random_engine */
void oracle_normal_draws(double mean, double sd, int n, double *out);
/* Nucleation / GrainInit critical diagonal lengths for one orientation (src/CAupdate.cpp:112-194). */
void oracle_seed_cdl(const float guv[9], float out[26]);
/* CellCapture's new decentred octahedron (src/CAupdate.cpp:353-563).
 * in: guv[9], capturer float centre (as double), new cell centre (xp,yp,zp).
 * out: centre[3] (double, before the float store), diag (float DL), cdl[26]. */
void oracle_capture_octahedron(const float guv[9], double cxold, double cyold, double czold, double xp,
                               double yp, double zp, double centre[3], float *diag, float cdl[26]);
/* Growth increment min(0.045, max(0, V(min(210, U)))) (src/CAupdate.cpp:257-261). */
double oracle_growth_increment(double A, double B, double C, double D, float U);

/* A simulation of a type-C deck started from another run's state at the end of step `cycle`
 * (the MI355X path's download): ct_raw (cell type | 0x10 for cell-centred octahedra, uint8 per
 * cell), GrainID, DiagonalLength (float) and octahedron centres (double[3]) per cell, and the P
 * nucleation sites. The temperature fields are the oracle's own; UndercoolingCurrent and the
 * critical lengths are derived as the reference accumulates and computes them. */
int oracle_load_state(const char *deck, const char *data_dir, int cycle, const uint8_t *ct_raw, const int32_t *gid,
                      const float *dl, const double *centre, int P, const int32_t *nloc, const int32_t *ntime,
                      oracle_sim **out);
/* nthreads > 1: CellCapture runs in parallel as the reference's OpenMP backend does (atomic
 * frontier appends, compare-exchange captures): a timing baseline whose results depend on thread
 * timing. 0 or 1: the serial restatement. */
void oracle_set_threads(oracle_sim *s, int nthreads);

/* ---- test hooks: hand-built states for the capture-order pins ----
 * oracle_set_field overwrites a whole field (the layouts of oracle_get_field; not the nucleus
 * arrays). oracle_set_orientation replaces one row of the orientation table (GrainUnitVector).
 * oracle_set_mutation plants a mistake in CellCapture's pass 2 so that a test can show the pins
 * fail it: ORACLE_MUT_DESCENDING visits the steering vector in descending index order,
 * ORACLE_MUT_LCOUNT does not count a liquid neighbour that the cell captures as liquid,
 * ORACLE_MUT_GROW_NEW lets a cell captured in a step grow in that step. 0 restores the
 * restatement. */
enum { ORACLE_MUT_NONE = 0, ORACLE_MUT_DESCENDING = 1, ORACLE_MUT_LCOUNT = 2, ORACLE_MUT_GROW_NEW = 3 };
int oracle_set_field(oracle_sim *s, int which, const void *src);
int oracle_set_orientation(oracle_sim *s, int o, const float guv[9]);
void oracle_set_mutation(oracle_sim *s, int mutation);

#ifdef __cplusplus
}
#endif
#endif
