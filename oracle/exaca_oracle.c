/*
 * exaca_oracle.c -- TEST INFRASTRUCTURE ONLY (parity checker; never part of the product path).
 *
 * A serial, plain-C restatement of the ExaCA-Kokkos (Aug 2021) solidification path as the
 * reference runs it on one MPI rank with the Kokkos Serial backend:
 *   deck parsing                 src/CAinitialize.cpp:24-576
 *   np = 1 mesh + stencil tables src/CAinitialize.cpp:579-745, 1003-1020; src/CAfunctions.cpp:34-213
 *   temperature init (C, S)      src/CAinitialize.cpp:1025-1124
 *   orientations                 src/CAinitialize.cpp:1408-1456
 *   substrate (C, S)             src/CAinitialize.cpp:1460-1732
 *   walls / active cells (S)     src/CAinitialize.cpp:1736-1830
 *   grain + nuclei init          src/CAinitialize.cpp:1834-2248, 2764-2767
 *   time loop                    src/runCA.cpp:283-342
 *   Nucleation, CellCapture      src/CAupdate.cpp:67-728
 *   termination check            src/CAupdate.cpp:732-822
 *   vtk / log output             src/CAprint.cpp:285-571
 * The libstdc++ (GCC 11) <random> streams the reference draws from (std::mt19937_64,
 * std::uniform_real_distribution<double>, std::default_random_engine = minstd_rand0,
 * std::normal_distribution<double>) are restated here as well; glibc libm supplies pow/sqrt/
 * sqrtf/log/acos exactly as it does for the reference.
 *
 * Deliberate differences from the reference, all outside its observable np = 1 behaviour:
 *   - 64-bit cell indices (the reference's int indices overflow above a 432^3 deck);
 *   - arrays the reference allocates without initialising are zero-filled (a large fresh
 *     allocation reads as zero, which is what the reference relies on for spot-melt decks);
 *   - decks with more than one layer, problem type R and Y/N print options are rejected.
 * Compile with -ffp-contract=off: the reference's arithmetic has no fused multiply-adds.
 */
#include "exaca_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Cell types, src/CAtypes.hpp:11 */
enum { WALL = 0, SOLID = 1, ACTIVE = 2, TEMP_UPDATE = 3, TEMP_INIT = 4, LIQUID = 5 };
#define NORIENT 10000

static char g_err[1024];
const char *oracle_last_error(void) { return g_err; }
static int fail(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

/* ======================================================================================
 * libstdc++ random streams
 * ====================================================================================== */
typedef struct {
    uint64_t x[312];
    int i;
} mt64_t;

/* std::mt19937_64 seeding: x0 = s, x_i = f*(x_{i-1} ^ (x_{i-1} >> 62)) + i. */
static void mt64_seed(mt64_t *g, uint64_t s) {
    g->x[0] = s;
    for (int k = 1; k < 312; k++) g->x[k] = 6364136223846793005ULL * (g->x[k - 1] ^ (g->x[k - 1] >> 62)) + (uint64_t)k;
    g->i = 312;
}
static uint64_t mt64_next(mt64_t *g) {
    const uint64_t hi = 0xFFFFFFFF80000000ULL, lo = 0x7FFFFFFFULL, a = 0xB5026F5AA96619E9ULL;
    if (g->i >= 312) {
        for (int k = 0; k < 312; k++) {
            uint64_t y = (g->x[k] & hi) | (g->x[(k + 1) % 312] & lo);
            g->x[k] = g->x[(k + 156) % 312] ^ (y >> 1) ^ ((y & 1ULL) ? a : 0ULL);
        }
        g->i = 0;
    }
    uint64_t z = g->x[g->i++];
    z ^= (z >> 29) & 0x5555555555555555ULL;
    z ^= (z << 17) & 0x71D67FFFEDA60000ULL;
    z ^= (z << 37) & 0xFFF7EEE000000000ULL;
    z ^= z >> 43;
    return z;
}
/* generate_canonical<double,53>(mt19937_64): one draw, divided by 2^64, clamped below 1. */
static double mt64_canonical(mt64_t *g) {
    double sum = (double)mt64_next(g);
    double tmp = (double)(1.0L * 18446744073709551616.0L);
    double r = sum / tmp;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

/* std::minstd_rand0, default seed 1. */
typedef struct {
    uint32_t x;
} minstd_t;
static uint32_t minstd_next(minstd_t *g) {
    g->x = (uint32_t)(((uint64_t)g->x * 16807ULL) % 2147483647ULL);
    return g->x;
}
/* generate_canonical<double,53>(minstd_rand0): range r = 2^31-2 gives floor(log2 r) = 30, so two
 * draws are combined: (u1-1) + (u2-1)*r, divided by r*r (long double product rounded to double). */
static double minstd_canonical(minstd_t *g) {
    const long double r = 2147483646.0L;
    double sum = 0.0, tmp = 1.0;
    for (int k = 2; k != 0; --k) {
        sum += (double)(minstd_next(g) - 1u) * tmp;
        tmp = (double)((long double)tmp * r);
    }
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}
/* normal_distribution<double> (Marsaglia polar method with one cached value). */
typedef struct {
    minstd_t g;
    int have;
    double saved;
} normal_t;
static double normal_next(normal_t *n, double mean, double sd) {
    double ret;
    if (n->have) {
        n->have = 0;
        ret = n->saved;
    }
    else {
        double x, y, r2;
        do {
            x = 2.0 * minstd_canonical(&n->g) - 1.0;
            y = 2.0 * minstd_canonical(&n->g) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        double mult = sqrt(-2 * log(r2) / r2);
        n->saved = x * mult;
        n->have = 1;
        ret = y * mult;
    }
    return ret * sd + mean;
}

void oracle_mt64_draws(uint64_t seed, int n, uint64_t *out) {
    mt64_t g;
    mt64_seed(&g, seed);
    for (int i = 0; i < n; i++) out[i] = mt64_next(&g);
}
void oracle_uniform_draws(uint64_t seed, int n, double *out) {
    mt64_t g;
    mt64_seed(&g, seed);
    for (int i = 0; i < n; i++) out[i] = mt64_canonical(&g) * (1.0 - 0.0) + 0.0;
}
void oracle_minstd_draws(int n, uint32_t *out) {
    minstd_t g = {1u};
    for (int i = 0; i < n; i++) out[i] = minstd_next(&g);
}
void oracle_normal_draws(double mean, double sd, int n, double *out) {
    normal_t nd;
    nd.g.x = 1u;
    nd.have = 0;
    nd.saved = 0.0;
    for (int i = 0; i < n; i++) out[i] = normal_next(&nd, mean, sd);
}

/* ======================================================================================
 * Simulation state
 * ====================================================================================== */
struct oracle_sim {
    /* deck (src/CAinitialize.cpp:232-576) */
    char type[16];
    int decomp;
    double A, B, C, D, FR, dx, NMax, dTN, dTs, G, R, dt, fract;
    double nread;
    char material[4096], path_out[4096], base[4096], orient_path[8192], sub_file[8192];
    int nx, ny, nz, nratio, nsx, nsy, soff, srad, nlayers, lheight, use_sub_file;
    float spacing;
    int print_mis, print_full, print_debug;
    /* np = 1 decomposition */
    int X, Y, xoff, yoff, zlow, zhigh, nzA;
    long long ndom, nact;
    int nbx[26], nby[26], nbz[26];
    int itlist[9][26];
    /* fields */
    int32_t *ct, *gid, *cts, *lid, *melted;
    float *U, *dU, *DL, *DOC, *CDL;
    float guv[9 * NORIENT];
    int32_t gorient[NORIENT];
    int P;
    int32_t *nloc, *ntime;
    int32_t *steer;
    int cycle;
    long long nn;
    uint64_t sum_front, captures, solidifications;
    int mutation; /* test hook: a planted mistake in pass 2 (ORACLE_MUT_*), 0 for the restatement */
    int nthreads; /* > 1: CellCapture runs as the reference's OpenMP backend does (timing only) */
    uint64_t seed; /* random-stream seed of a type-C initial state (0: the reference's own streams) */
    int layer;     /* the layer being solidified (src/runCA.cpp:283) */
    long long nn_layer; /* nucleation events before this layer (the reference's nn restarts per layer) */
    long long outside;  /* nuclei or captures that fell outside the active window (the reference would
                           index its window arrays out of bounds there); skipped and counted */
};

/* ---------------- deck parsing (src/CAinitialize.cpp:24-134) ---------------- */
static int read_line(FILE *f, char *buf, size_t n) {
    if (!fgets(buf, (int)n, f)) {
        buf[0] = 0;
        return 0;
    }
    size_t L = strlen(buf);
    if (L && buf[L - 1] == '\n') buf[L - 1] = 0;
    return 1;
}
static void skip_header(FILE *f) {
    char line[4096];
    while (read_line(f, line, sizeof line))
        if (strcmp(line, "*****") == 0) break;
}
static void strip_ws(const char *src, char *dst) {
    for (; *src; src++)
        if (!isspace((unsigned char)*src)) *dst++ = *src;
    *dst = 0;
}
/* getKey: the reference's deprecated-key loop exits as soon as any deprecated substring is
 * absent, so a key is always returned as read (src/CAinitialize.cpp:36-62). */
static int get_key(FILE *f, char *line, size_t n, char *key, const char **after) {
    read_line(f, line, n);
    char *c = strchr(line, ':');
    if (!c) {
        strcpy(key, line);
        return fail("Input \"%s\" must be separated from value by \":\".", line);
    }
    size_t kl = (size_t)(c - line);
    memcpy(key, line, kl);
    key[kl] = 0;
    *after = c + 1;
    return 0;
}
static int parse_input(FILE *f, const char *want, char *val) {
    char line[4096], key[4096];
    const char *after;
    if (get_key(f, line, sizeof line, key, &after)) return -1;
    if (!strstr(key, want)) return fail("Required input not present: \"%s\" not found in the input file", want);
    strip_ws(after, val);
    return 0;
}
static int parse_input2(FILE *f, const char *k1, const char *k2, char *val, int *which) {
    char line[4096], key[4096];
    const char *after;
    if (get_key(f, line, sizeof line, key, &after)) return -1;
    if (strstr(key, k1))
        *which = 1;
    else if (strstr(key, k2))
        *which = 2;
    else
        return fail("Required input not present: Neither %s nor %s was found in the input file", k1, k2);
    strip_ws(after, val);
    return 0;
}
static int parse_bool(FILE *f, const char *want, int *out) {
    char val[4096];
    if (parse_input(f, want, val)) return -1;
    if (strcmp(val, "N") == 0)
        *out = 0;
    else if (strcmp(val, "Y") == 0)
        *out = 1;
    else
        return fail("Input \"%s\" must be \"Y\" or \"N\".", want);
    return 0;
}
static int to_int(const char *v, int *out) {
    char *end;
    long x = strtol(v, &end, 10);
    if (end == v) return fail("could not convert \"%s\" to an integer", v);
    *out = (int)x;
    return 0;
}
static int file_exists(const char *p) {
    FILE *f = fopen(p, "r");
    if (!f) return 0;
    fclose(f);
    return 1;
}

#define TRY(x)                                                                                               \
    do {                                                                                                     \
        if ((x) != 0) goto bad;                                                                              \
    } while (0)

static int read_deck(oracle_sim *s, const char *deck, const char *data_dir) {
    char val[4096];
    FILE *in = fopen(deck, "r"), *mat = NULL;
    if (!in) return fail("Could not locate/open Input file");
    skip_header(in);
    TRY(parse_input(in, "Problem type", s->type));
    TRY(parse_input(in, "Decomposition strategy", val));
    TRY(to_int(val, &s->decomp));
    TRY(parse_input(in, "Material", s->material));
    {
        char mpath[4096];
        snprintf(mpath, sizeof mpath, "%s/Materials/%s", data_dir, s->material);
        mat = fopen(mpath, "r");
        if (!mat) {
            fail("Could not locate/open Materials file");
            goto bad;
        }
        skip_header(mat);
        TRY(parse_input(mat, "A", val));
        s->A = atof(val);
        TRY(parse_input(mat, "B", val));
        s->B = atof(val);
        TRY(parse_input(mat, "C", val));
        s->C = atof(val);
        TRY(parse_input(mat, "D", val));
        s->D = atof(val);
        TRY(parse_input(mat, "Alloy freezing range", val));
        s->FR = atof(val);
        fclose(mat);
        mat = NULL;
    }
    TRY(parse_input(in, "Cell size", val));
    s->dx = atof(val) * pow(10, -6);
    TRY(parse_input(in, "Heterogeneous nucleation density", val));
    s->nread = atof(val);
    s->NMax = s->nread * pow(10, 12);
    TRY(parse_input(in, "Mean nucleation undercooling", val));
    s->dTN = atof(val);
    TRY(parse_input(in, "Standard deviation of nucleation undercooling", val));
    s->dTs = atof(val);
    TRY(parse_input(in, "Path to output", s->path_out));
    TRY(parse_input(in, "Output file base name", s->base));
    TRY(parse_input(in, "File of grain orientations", val));
    snprintf(s->orient_path, sizeof s->orient_path, "%s/Substrate/%s", data_dir, val);
    if (!file_exists(s->orient_path)) {
        fail("Could not locate/open Substrate file");
        goto bad;
    }
    if (strcmp(s->type, "C") == 0) {
        s->nlayers = 1;
        TRY(parse_input(in, "Thermal gradient", val));
        s->G = atof(val);
        TRY(parse_input(in, "Cooling rate", val));
        s->R = atof(val);
        TRY(parse_input(in, "Domain size in x", val));
        TRY(to_int(val, &s->nx));
        s->nx += 2;
        TRY(parse_input(in, "Domain size in y", val));
        TRY(to_int(val, &s->ny));
        s->ny += 2;
        TRY(parse_input(in, "Domain size in z", val));
        TRY(to_int(val, &s->nz));
        s->nz += 2;
        s->lheight = s->nz;
        TRY(parse_input(in, "\"N\"", val));
        TRY(to_int(val, &s->nratio));
        s->dt = s->dx / (s->nratio * (s->R / s->G));
        TRY(parse_input(in, "Fraction of surface sites active", val));
        s->fract = atof(val);
    }
    else if (strcmp(s->type, "S") == 0) {
        int which = 0;
        TRY(parse_input(in, "Thermal gradient", val));
        s->G = atof(val);
        TRY(parse_input(in, "Cooling rate", val));
        s->R = atof(val);
        TRY(parse_input(in, "Number of spots in x", val));
        TRY(to_int(val, &s->nsx));
        TRY(parse_input(in, "Number of spots in y", val));
        TRY(to_int(val, &s->nsy));
        TRY(parse_input(in, "Offset between spot centers", val));
        float off = (float)atof(val);
        s->soff = (int)(off * 1.0 * pow(10, -6) / s->dx);
        TRY(parse_input(in, "Radii of spots", val));
        float rad = (float)atof(val);
        s->srad = (int)(rad * 1.0 * pow(10, -6) / s->dx);
        TRY(parse_input(in, "Number of layers", val));
        TRY(to_int(val, &s->nlayers));
        TRY(parse_input(in, "Offset between layers", val));
        TRY(to_int(val, &s->lheight));
        TRY(parse_input(in, "\"N\"", val));
        TRY(to_int(val, &s->nratio));
        s->dt = s->dx / (s->nratio * (s->R / s->G));
        TRY(parse_input2(in, "Path to substrate file", "Substrate grain spacing", val, &which));
        if (which == 1) {
            char name[4096];
            TRY(parse_input(in, "Substrate file name", name));
            snprintf(s->sub_file, sizeof s->sub_file, "%s/%s", val, name);
            s->use_sub_file = 1;
            if (!file_exists(s->sub_file)) {
                fail("Could not locate/open Substrate file");
                goto bad;
            }
        }
        else {
            s->spacing = (float)atof(val);
            s->use_sub_file = 0;
        }
        s->nz = s->srad + 3 + (s->nlayers - 1) * s->lheight;
        s->nx = 4 + 2 * s->srad + s->soff * (s->nsx - 1);
        s->ny = 4 + 2 * s->srad + s->soff * (s->nsy - 1);
    }
    else {
        fail("problem type \"%s\" is outside the restated path (only C and S)", s->type);
        goto bad;
    }
    {
        int which = 0;
        TRY(parse_input2(in, "Print file of grain orientations", "Output data printing options", val, &which));
        if (which == 1) {
            int old[6] = {0, 0, 0, 0, 0, 0};
            if (strcmp(val, "Y") == 0) old[0] = 1;
            TRY(parse_bool(in, "Print csv file of grain id values", &old[1]));
            TRY(parse_bool(in, "Print csv file of ExaConstit-formatted grain id values", &old[2]));
            TRY(parse_bool(in, "Print Paraview vtk file of grain misorientation values", &old[3]));
            TRY(parse_bool(in, "Print file of grain area values", &old[4]));
            TRY(parse_bool(in, "Print file of weighted grain area value", &old[5]));
            s->print_debug = 0;
            s->print_mis = old[3];
            s->print_full = (old[0] + old[1] + old[2] + old[4] + old[5]) > 0;
        }
        else {
            int a = 0, b = 0;
            TRY(parse_bool(in, "Paraview vtk file of grain misorientation", &s->print_mis));
            TRY(parse_bool(in, "Paraview vtk file of all ExaCA data", &s->print_full));
            TRY(parse_bool(in, "Print data for main Kokkos views", &a));
            TRY(parse_bool(in, "Print data for all main data structures", &b));
            s->print_debug = b ? 2 : (a ? 1 : 0);
        }
    }
    fclose(in);
    if (s->nlayers < 1) return fail("at least one layer is needed");
    return 0;
bad:
    if (mat) fclose(mat);
    fclose(in);
    return -1;
}

/* ---------------- mesh + stencil tables (src/CAinitialize.cpp:590-745) ---------------- */
static void mesh_init(oracle_sim *s) {
    /* 26 neighbour offsets: 0-8 at dy = -1, 9-16 in the y plane, 17-25 at dy = +1. */
    static const int off[26][3] = {{0, -1, 0},  {1, -1, 0},  {-1, -1, 0}, {0, -1, 1},  {0, -1, -1}, {-1, -1, 1},
                                   {1, -1, 1},  {-1, -1, -1}, {1, -1, -1}, {0, 0, 1},   {0, 0, -1},  {1, 0, 1},
                                   {-1, 0, 1},  {1, 0, -1},  {-1, 0, -1}, {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},
                                   {1, 1, 0},   {-1, 1, 0},  {0, 1, 1},   {0, 1, -1},  {1, 1, 1},   {-1, 1, 1},
                                   {1, 1, -1},  {-1, 1, -1}};
    for (int n = 0; n < 26; n++) {
        s->nbx[n] = off[n][0];
        s->nby[n] = off[n][1];
        s->nbz[n] = off[n][2];
    }
    /* Edge-case lists: case 0 all; 1 no dy=-1; 2 no dy=+1; 3 no dx=-1; 4 no dx=+1; 5-8 corners. */
    for (int c = 0; c < 9; c++) {
        int pos = 0;
        for (int n = 0; n < 26; n++) {
            int ok = 1;
            if ((c == 1 || c == 5 || c == 6) && s->nby[n] == -1) ok = 0;
            if ((c == 2 || c == 7 || c == 8) && s->nby[n] == 1) ok = 0;
            if ((c == 3 || c == 5 || c == 7) && s->nbx[n] == -1) ok = 0;
            if ((c == 4 || c == 6 || c == 8) && s->nbx[n] == 1) ok = 0;
            if (ok) s->itlist[c][pos++] = n;
        }
        for (; pos < 26; pos++) s->itlist[c][pos] = 0;
    }
    /* np = 1: one slab of nx+2 by ny+2 including the ghost ring, offset -1 (src/CAfunctions.cpp:34-166). */
    s->X = s->nx + 2;
    s->Y = s->ny + 2;
    s->xoff = -1;
    s->yoff = -1;
}
/* FindItBounds (src/CAfunctions.cpp:169-213). */
static int it_bounds(int x, int y, int X, int Y) {
    if (y == 0) return x == 0 ? 5 : (x == X - 1 ? 6 : 1);
    if (y == Y - 1) return x == 0 ? 7 : (x == X - 1 ? 8 : 2);
    return x == 0 ? 3 : (x == X - 1 ? 4 : 0);
}
static int list_len(int b) { return b == 0 ? 26 : (b > 4 ? 11 : 17); }

static int alloc_fields(oracle_sim *s) {
    s->ndom = (long long)s->X * s->Y * s->nz;
    s->ct = calloc((size_t)s->ndom, 4);
    s->gid = calloc((size_t)s->ndom, 4);
    s->cts = calloc((size_t)s->ndom, 4);
    s->lid = calloc((size_t)s->ndom, 4);
    s->melted = calloc((size_t)s->ndom, 4);
    s->U = calloc((size_t)s->ndom, 4);
    s->dU = calloc((size_t)s->ndom, 4);
    if (!s->ct || !s->gid || !s->cts || !s->lid || !s->melted || !s->U || !s->dU) return fail("out of host memory");
    return 0;
}

/* ---------------- temperature fields (src/CAinitialize.cpp:1025-1124) ---------------- */
static void temp_init_C(oracle_sim *s) {
    s->zlow = 0;
    s->zhigh = s->nz - 1;
    s->nzA = s->nz;
    for (int k = 0; k < s->nz; k++)
        for (int i = 0; i < s->X; i++)
            for (int j = 0; j < s->Y; j++) {
                long long p = ((long long)k * s->X + i) * s->Y + j;
                s->U[p] = 0;
                s->lid[p] = 0;
                s->dU[p] = (float)(s->R * s->dt);
                s->cts[p] = (int)(((k - 1) * s->G * s->dx) / (s->R * s->dt));
                int gx = i + s->xoff, gy = j + s->yoff;
                s->melted[p] = (gx > -1 && gx < s->nx && gy > -1 && gy < s->ny && k > 0 && k < s->nz - 1);
            }
}
static void temp_init_S(oracle_sim *s) {
    s->zlow = 0;
    s->zhigh = s->srad + 1;
    s->nzA = s->srad + 2;
    for (long long p = 0; p < s->ndom; p++) {
        s->U[p] = 0;
        s->melted[p] = 0;
        s->lid[p] = -1;
    }
    int nspots = s->nsx * s->nsy;
    float iso = (float)((s->R / s->G) * s->dt / s->dx);
    int between = (int)(s->srad / iso + (s->FR / s->R) / s->dt);
    for (int layer = 0; layer < s->nlayers; layer++)
        for (int n = 0; n < nspots; n++) {
            int xs = 2 + s->srad + (n % s->nsx) * s->soff;
            int ys = 2 + s->srad + (n / s->nsx) * s->soff;
            int zs = s->srad + 2 + s->lheight * layer;
            for (int k = 0; k < s->nzA; k++) {
                float dz = (float)(zs - (k + s->lheight * layer));
                for (int i = 0; i < s->X; i++) {
                    float ddx = (float)(xs - (i + s->xoff));
                    for (int j = 0; j < s->Y; j++) {
                        float ddy = (float)(ys - (j + s->yoff));
                        float d = sqrtf(ddx * ddx + ddy * ddy + dz * dz);
                        if (d <= s->srad) {
                            long long p = ((long long)(k + layer * s->lheight) * s->X + i) * s->Y + j;
                            s->cts[p] = 1 + (int)(((float)s->srad - d) / iso) + between * n;
                            s->dU[p] = (float)(s->R * s->dt);
                            s->lid[p] = layer;
                            s->melted[p] = 1;
                        }
                    }
                }
            }
        }
}

/* ---------------- orientations (src/CAinitialize.cpp:1408-1456) ---------------- */
static int orient_init(oracle_sim *s) {
    FILE *f = fopen(s->orient_path, "r");
    if (!f) return fail("could not open %s", s->orient_path);
    char line[4096];
    read_line(f, line, sizeof line);
    for (int i = 0; i < NORIENT; i++) {
        if (!read_line(f, line, sizeof line)) break;
        int comp = 0, uv = 0;
        char *p = line;
        while (1) {
            char *c = strchr(p, ',');
            if (c) *c = 0;
            if (9 * i + 3 * uv + comp < 9 * NORIENT) s->guv[9 * i + 3 * uv + comp] = (float)atof(p);
            if (++comp > 2) {
                comp = 0;
                uv++;
            }
            if (!c) break;
            p = c + 1;
        }
    }
    fclose(f);
    for (int h = 0; h < NORIENT; h++) s->gorient[h] = h;
    return 0;
}

/* ---------------- substrates (src/CAinitialize.cpp:1460-1732) ---------------- */
static void substrate_C(oracle_sim *s) {
    mt64_t g;
    mt64_seed(&g, s->seed); /* rank id 0 (or the instance seed) */
    for (int k = 0; k < s->nz; k++)
        for (int i = 0; i < s->X; i++)
            for (int j = 0; j < s->Y; j++) {
                int gx = i + s->xoff, gy = j + s->yoff;
                long long p = ((long long)k * s->X + i) * s->Y + j;
                if (gx == -1 || gx == s->nx || gy == -1 || gy == s->ny || k == 0 || k == s->nz - 1) {
                    s->ct[p] = WALL;
                    s->gid[p] = 0;
                }
                else if (k == 1 && i > 0 && i < s->X - 1 && j > 0 && j < s->Y - 1) {
                    double r = mt64_canonical(&g);
                    s->ct[p] = (r < s->fract) ? ACTIVE : LIQUID;
                }
                else
                    s->ct[p] = LIQUID;
            }
    int next = 1;
    for (long long p = 0; p < s->ndom; p++)
        if (s->ct[p] == ACTIVE) s->gid[p] = next++;
}

static int substrate_from_file(oracle_sim *s) {
    FILE *f = fopen(s->sub_file, "r");
    if (!f) return fail("Could not locate/open Substrate file");
    char line[4096];
    int dims[3];
    for (int d = 0; d < 3; d++) {
        read_line(f, line, sizeof line);
        char *e = strchr(line, '=');
        dims[d] = atoi(e ? e + 1 : line);
    }
    int nzS = dims[0], nyS = dims[1], nxS = dims[2];
    if (nzS < s->nz) {
        fclose(f);
        return fail("Error: only %d layers of substrate data are present in the file %s ; at least %d layers of "
                    "substrate data are required to simulate the specified solidification problem",
                    nzS, s->sub_file, s->nz);
    }
    for (int k = 0; k < nzS && k != s->nz; k++)
        for (int j = 0; j < nyS; j++)
            for (int i = 0; i < nxS; i++) {
                read_line(f, line, sizeof line);
                if (i >= s->xoff && i < s->xoff + s->X && j >= s->yoff && j < s->yoff + s->Y) {
                    long long p = ((long long)k * s->X + (i - s->xoff)) * s->Y + (j - s->yoff);
                    s->gid[p] = (s->cts[p] == 0) ? atoi(line) : 0;
                }
            }
    fclose(f);
    return 0;
}

static int substrate_from_spacing(oracle_sim *s) {
    mt64_t g;
    mt64_seed(&g, 1); /* std::mt19937_64 gen(1.0) */
    double prob = (s->dx * s->dx * s->dx) / (s->spacing * s->spacing * s->spacing * pow(10, -18));
    long long cap = (long long)s->nx * s->ny * s->nzA, nb = 0;
    int *bx = malloc(sizeof(int) * (size_t)(cap + 1)), *by = malloc(sizeof(int) * (size_t)(cap + 1)),
        *bz = malloc(sizeof(int) * (size_t)(cap + 1));
    if (!bx || !by || !bz) return fail("out of host memory");
    for (int k = 1; k <= s->nzA; k++)
        for (int i = 1; i < s->nx - 1; i++)
            for (int j = 1; j < s->ny - 1; j++) {
                double r = mt64_canonical(&g);
                if (r < prob) {
                    bx[nb] = i;
                    by[nb] = j;
                    bz[nb] = k;
                    nb++;
                }
            }
    long long nactive = (long long)s->X * s->Y * s->nzA;
    for (int k = 1; k < s->nzA; k++)
        for (int i = 0; i < s->X; i++)
            for (int j = 0; j < s->Y; j++) {
                int gx = i + s->xoff, gy = j + s->yoff;
                long long p = ((long long)k * s->X + i) * s->Y + j;
                if (s->cts[p] == 0) {
                    float best = (float)nactive;
                    long long bi = -1;
                    for (long long n = 0; n < nb; n++) {
                        float ddx = (float)abs(bx[n] - gx), ddy = (float)abs(by[n] - gy), ddz = (float)abs(bz[n] - k);
                        float d = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
                        if (d < best) {
                            bi = n;
                            best = d;
                        }
                    }
                    s->gid[p] = (int32_t)(bi + 1);
                }
                else
                    s->gid[p] = 0;
            }
    for (int k = s->nzA; k < s->nz; k++)
        for (int i = 1; i < s->X - 1; i++)
            for (int j = 1; j < s->Y - 1; j++) {
                long long p = ((long long)k * s->X + i) * s->Y + j;
                if (s->cts[p] != 0) s->gid[p] = 0;
            }
    int next = (int)nb + 1;
    for (long long p = (long long)s->X * s->Y * s->nzA; p < s->ndom; p++)
        if (s->cts[p] == 0) s->gid[p] = next++;
    free(bx);
    free(by);
    free(bz);
    return 0;
}

/* ActiveCellWallInit (src/CAinitialize.cpp:1736-1830), ExtraWalls = false. */
static void active_wall_init(oracle_sim *s) {
    const int X = s->X, Y = s->Y;
    for (int k = 0; k < s->nz; k++)
        for (int i = 0; i < X; i++)
            for (int j = 0; j < Y; j++) {
                int gx = i + s->xoff, gy = j + s->yoff;
                long long p = ((long long)k * X + i) * Y + j;
                if (gx == -1 || gx == s->nx || gy == -1 || gy == s->ny || k == 0 || k == s->nz - 1) {
                    s->ct[p] = WALL;
                    s->gid[p] = 0;
                }
            }
    for (int k = 1; k < s->nz - 1; k++)
        for (int j = 0; j < Y; j++)
            for (int i = 0; i < X; i++) {
                long long p = ((long long)k * X + i) * Y + j;
                s->ct[p] = (s->cts[p] == 0) ? SOLID : LIQUID;
            }
    for (int k = 1; k < s->nz - 1; k++)
        for (int j = 0; j < Y; j++)
            for (int i = 0; i < X; i++) {
                long long p = ((long long)k * X + i) * Y + j;
                if (s->cts[p] != 0) continue;
                int b = it_bounds(i, j, X, Y);
                for (int ll = 0; ll < list_len(b); ll++) {
                    int l = s->itlist[b][ll];
                    long long q = ((long long)(k + s->nbz[l]) * X + (i + s->nbx[l])) * Y + (j + s->nby[l]);
                    if (s->cts[q] > 0 && s->ct[q] != WALL) {
                        s->ct[p] = ACTIVE;
                        s->dU[p] = s->dU[q];
                        break;
                    }
                }
            }
}

/* The cell-centred octahedron's 26 critical diagonal lengths (src/CAupdate.cpp:112-194 and the
 * identical GrainInit loop, src/CAinitialize.cpp:1963-2042, where x0 reduces to the offset). */
static void seed_cdl(const float *g, const int *nbx, const int *nby, const int *nbz, float *out) {
    for (int n = 0; n < 26; n++) {
        double x0 = nbx[n], y0 = nby[n], z0 = nbz[n];
        double mag0 = pow(pow(x0, 2.0) + pow(y0, 2.0) + pow(z0, 2.0), 0.5);
        double a1 = (g[0] * x0 + g[1] * y0 + g[2] * z0) / mag0;
        double d1x, d1y, d1z, d2x, d2y, d2z, d3x, d3y, d3z;
        if (a1 < 0) {
            d1x = g[0];
            d1y = g[1];
            d1z = g[2];
        }
        else {
            d1x = -g[0];
            d1y = -g[1];
            d1z = -g[2];
        }
        double a2 = (g[3] * x0 + g[4] * y0 + g[5] * z0) / mag0;
        if (a2 < 0) {
            d2x = g[3];
            d2y = g[4];
            d2z = g[5];
        }
        else {
            d2x = -g[3];
            d2y = -g[4];
            d2z = -g[5];
        }
        double a3 = (g[6] * x0 + g[7] * y0 + g[8] * z0) / mag0;
        if (a3 < 0) {
            d3x = g[6];
            d3y = g[7];
            d3z = g[8];
        }
        else {
            d3x = -g[6];
            d3y = -g[7];
            d3z = -g[8];
        }
        double u1[3] = {d2x - d1x, d2y - d1y, d2z - d1z};
        double u2[3] = {d3x - d1x, d3y - d1y, d3z - d1z};
        double uu[3] = {u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2], u1[0] * u2[1] - u1[1] * u2[0]};
        double nd = sqrt(uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2]);
        double nx_ = uu[0] / nd, ny_ = uu[1] / nd, nz_ = uu[2] / nd;
        double t = (nx_ * x0 + ny_ * y0 + nz_ * z0) / (nx_ * d1x + ny_ * d1y + nz_ * d1z);
        out[n] = (float)pow(pow(t * d1x, 2.0) + pow(t * d1y, 2.0) + pow(t * d1z, 2.0), 0.5);
    }
}
static const int k_nbx[26] = {0, 1, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, 0, 1, -1, 0, 0, 1, -1, 1, -1};
static const int k_nby[26] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1};
static const int k_nbz[26] = {0, 0, 0, 1, -1, 1, 1, -1, -1, 1, -1, 1, 1, -1, -1, 0, 0, 0, 0, 0, 1, -1, 1, 1, -1, -1};
void oracle_seed_cdl(const float guv[9], float out[26]) { seed_cdl(guv, k_nbx, k_nby, k_nbz, out); }

/* GrainOrientation((abs(GrainID) - 1) % NGrainOrientations). A GrainID of 0 (never carried by a
 * growing or printed melted cell in a completed run) would index -1 in the reference; 0 is used. */
static int orientation_of(const oracle_sim *s, int gid) {
    int k = (abs(gid) - 1) % NORIENT;
    return s->gorient[k < 0 ? 0 : k];
}

/* GrainInit (src/CAinitialize.cpp:1834-2206), np = 1, first layer. */
static int grain_init(oracle_sim *s) {
    const int X = s->X, Y = s->Y;
    mt64_t g;
    mt64_seed(&g, s->seed);
    double bulk = s->NMax * s->dx * s->dx * s->dx;
    s->P = 0;
    for (int k = 1; k < s->nz - 1; k++)
        for (int j = 1; j < Y - 1; j++)
            for (int i = 1; i < X - 1; i++) {
                long long p = ((long long)k * X + i) * Y + j;
                if (s->cts[p] != 0) {
                    double r = mt64_canonical(&g);
                    if (r < bulk) {
                        s->P++;
                        s->ct[p] = TEMP_INIT;
                    }
                }
            }
    s->nact = (long long)X * Y * s->nzA;
    s->DL = calloc((size_t)s->nact, 4);
    s->DOC = calloc((size_t)s->nact * 3, 4);
    s->CDL = calloc((size_t)s->nact * 26, 4);
    s->steer = calloc((size_t)s->nact, 4);
    if (!s->DL || !s->DOC || !s->CDL || !s->steer) return fail("out of host memory");
    int ncounter = -1;
    for (int gz = 1; gz < s->nz - 1; gz++)
        for (int rx = 0; rx < X; rx++)
            for (int ry = 0; ry < Y; ry++) {
                long long pg = ((long long)gz * X + rx) * Y + ry;
                if (s->ct[pg] == ACTIVE) {
                    if (gz <= s->zhigh) {
                        long long pa = ((long long)(gz - s->zlow) * X + rx) * Y + ry;
                        int o = orientation_of(s, s->gid[pg]);
                        s->DL[pa] = 0.01f;
                        s->DOC[3 * pa] = (float)(rx + s->xoff + 0.5);
                        s->DOC[3 * pa + 1] = (float)(ry + s->yoff + 0.5);
                        s->DOC[3 * pa + 2] = (float)(gz + 0.5);
                        seed_cdl(&s->guv[9 * o], s->nbx, s->nby, s->nbz, &s->CDL[26 * pa]);
                    }
                }
                else if (s->ct[pg] == TEMP_INIT) {
                    s->gid[pg] = ncounter--;
                }
            }
    /* Liquid cells with no non-solid, non-wall neighbour become solid, in index order. */
    for (int rz = 1; rz < s->nz - 1; rz++)
        for (int rx = 1; rx < X - 1; rx++)
            for (int ry = 1; ry < Y - 1; ry++) {
                long long p = ((long long)rz * X + rx) * Y + ry;
                if (s->ct[p] != LIQUID) continue;
                int b = it_bounds(rx, ry, X, Y), cnt = 0;
                for (int ll = 0; ll < list_len(b); ll++) {
                    int l = s->itlist[b][ll];
                    long long q = ((long long)(rz + s->nbz[l]) * X + (rx + s->nbx[l])) * Y + (ry + s->nby[l]);
                    if (s->ct[q] != SOLID && s->ct[q] != WALL) cnt++;
                }
                if (cnt == 0) s->ct[p] = SOLID;
            }
    return 0;
}

/* NucleiInit (src/CAinitialize.cpp:2213-2248, 2764-2767), np = 1. */
static int nuclei_init(oracle_sim *s) {
    normal_t nd;
    nd.g.x = (uint32_t)(s->seed % 2147483646u + 1u); /* default seed 1 (or from the instance seed) */
    nd.have = 0;
    nd.saved = 0;
    s->nloc = calloc((size_t)(s->P > 0 ? s->P : 1), 4);
    s->ntime = calloc((size_t)(s->P > 0 ? s->P : 1), 4);
    if (!s->nloc || !s->ntime) return fail("out of host memory");
    int ne = 0;
    for (long long p = 0; p < s->ndom; p++) {
        if (s->ct[p] != TEMP_INIT) continue;
        s->nloc[ne] = (int32_t)p;
        double und = normal_next(&nd, s->dTN, s->dTs);
        int t = (int)(s->cts[p] + round(und / s->dU[p]));
        s->ntime[ne] = s->cts[p] > t ? s->cts[p] : t;
        ne++;
    }
    for (long long p = 0; p < s->ndom; p++)
        if (s->ct[p] == TEMP_INIT) s->ct[p] = LIQUID;
    return 0;
}

static void critical_lengths(const float *g, double xp, double yp, double zp, double cx, double cy, double cz,
                             const int *nbx, const int *nby, const int *nbz, float *cdl_out);

/* A simulation whose cell state is taken from another run at the end of step `cycle`: the deck,
 * mesh, orientations and temperature fields are the oracle's own (TempInit sets CritTimeStep,
 * LayerID, UndercoolingChange and Melted); CellType, GrainID, the nucleation sites and the active
 * cells' octahedra come from the caller. UndercoolingCurrent is what the reference's pass 1
 * accumulates (src/CAupdate.cpp:235-237): the float sum of cycle - CritTimeStep increments for a
 * liquid or active cell past its liquidus time. The critical lengths are recomputed from the
 * centre (double, as computed at capture) or, for a cell-centred octahedron (bit 0x10 of ct_raw),
 * from the orientation (src/CAupdate.cpp:112-194, 524-563), as the reference computed them. Type
 * C only (no cycle jump; one UndercoolingChange shared by every cell). */
int oracle_load_state(const char *deck, const char *data_dir, int cycle, const uint8_t *ct_raw, const int32_t *gid,
                      const float *dl, const double *centre, int P, const int32_t *nloc, const int32_t *ntime,
                      oracle_sim **out) {
    oracle_sim *s = calloc(1, sizeof *s);
    if (!s) return fail("out of host memory");
    *out = NULL;
    if (read_deck(s, deck, data_dir)) goto bad;
    if (strcmp(s->type, "C") != 0) {
        fail("state import covers directional solidification (type C) decks");
        goto bad;
    }
    mesh_init(s);
    if (alloc_fields(s)) goto bad;
    temp_init_C(s);
    if (orient_init(s)) goto bad;
    s->nact = (long long)s->X * s->Y * s->nzA;
    s->DL = calloc((size_t)s->nact, 4);
    s->DOC = calloc((size_t)s->nact * 3, 4);
    s->CDL = calloc((size_t)s->nact * 26, 4);
    s->steer = calloc((size_t)s->nact, 4);
    s->P = P;
    s->nloc = calloc((size_t)(P > 0 ? P : 1), 4);
    s->ntime = calloc((size_t)(P > 0 ? P : 1), 4);
    if (!s->DL || !s->DOC || !s->CDL || !s->steer || !s->nloc || !s->ntime) {
        fail("out of host memory");
        goto bad;
    }
    if (P > 0) {
        memcpy(s->nloc, nloc, (size_t)P * 4);
        memcpy(s->ntime, ntime, (size_t)P * 4);
    }
    {
        /* U_k, the float sum of k increments of the shared UndercoolingChange */
        const float du = (float)(s->R * s->dt);
        int kmax = 0;
        for (long long p = 0; p < s->ndom; p++) {
            int k = cycle - s->cts[p];
            if (k > kmax) kmax = k;
        }
        float *uk = calloc((size_t)kmax + 1, sizeof(float));
        if (!uk) {
            fail("out of host memory");
            goto bad;
        }
        for (int k = 1; k <= kmax; k++) uk[k] = uk[k - 1] + du * (float)1;
        const long long XY = (long long)s->X * s->Y;
        for (long long p = 0; p < s->ndom; p++) {
            const int t = ct_raw[p] & 0x0F;
            s->ct[p] = t;
            s->gid[p] = gid[p];
            if ((t == LIQUID || t == ACTIVE) && cycle > s->cts[p]) s->U[p] = uk[cycle - s->cts[p]];
            if (t != ACTIVE) continue;
            const long long pa = p - (long long)s->zlow * XY;
            const int rz = (int)(pa / XY);
            const long long rem = pa % XY;
            const int rx = (int)(rem / s->Y), ry = (int)(rem % s->Y);
            const int o = orientation_of(s, gid[p]);
            s->DL[pa] = dl[p];
            s->DOC[3 * pa] = (float)centre[3 * p];
            s->DOC[3 * pa + 1] = (float)centre[3 * p + 1];
            s->DOC[3 * pa + 2] = (float)centre[3 * p + 2];
            if (ct_raw[p] & 0x10)
                seed_cdl(&s->guv[9 * o], s->nbx, s->nby, s->nbz, &s->CDL[26 * pa]);
            else
                critical_lengths(&s->guv[9 * o], (rx + s->xoff) + 0.5, (ry + s->yoff) + 0.5, (rz + s->zlow) + 0.5,
                                 centre[3 * p], centre[3 * p + 1], centre[3 * p + 2], s->nbx, s->nby, s->nbz,
                                 &s->CDL[26 * pa]);
        }
        free(uk);
    }
    s->A = s->A * s->dt / s->dx;
    s->B = s->B * s->dt / s->dx;
    s->C = s->C * s->dt / s->dx;
    s->cycle = cycle;
    *out = s;
    return 0;
bad:
    oracle_free(s);
    return -1;
}

int oracle_load(const char *deck, const char *data_dir, oracle_sim **out) {
    return oracle_load_seeded(deck, data_dir, 0, out);
}

/* An ensemble instance's streams: mt19937_64(seed) for the substrate and the nucleation sites,
 * default_random_engine(seed % 2147483646 + 1) for the nucleation undercoolings (type C). */
int oracle_load_seeded(const char *deck, const char *data_dir, uint64_t seed, oracle_sim **out) {
    oracle_sim *s = calloc(1, sizeof *s);
    if (!s) return fail("out of host memory");
    *out = NULL;
    s->seed = seed;
    if (read_deck(s, deck, data_dir)) goto bad;
    if (seed && strcmp(s->type, "C") != 0) {
        fail("seeded initial states cover directional solidification (type C) decks");
        goto bad;
    }
    mesh_init(s);
    if (alloc_fields(s)) goto bad;
    if (strcmp(s->type, "S") == 0)
        temp_init_S(s);
    else
        temp_init_C(s);
    if (orient_init(s)) goto bad;
    if (strcmp(s->type, "C") == 0)
        substrate_C(s);
    else {
        if (s->use_sub_file) {
            if (substrate_from_file(s)) goto bad;
        }
        else if (substrate_from_spacing(s))
            goto bad;
        active_wall_init(s);
    }
    if (grain_init(s)) goto bad;
    if (nuclei_init(s)) goto bad;
    s->A = s->A * s->dt / s->dx;
    s->B = s->B * s->dt / s->dx;
    s->C = s->C * s->dt / s->dx;
    s->cycle = 0;
    *out = s;
    return 0;
bad:
    oracle_free(s);
    return -1;
}

void oracle_free(oracle_sim *s) {
    if (!s) return;
    free(s->ct);
    free(s->gid);
    free(s->cts);
    free(s->lid);
    free(s->melted);
    free(s->U);
    free(s->dU);
    free(s->DL);
    free(s->DOC);
    free(s->CDL);
    free(s->nloc);
    free(s->ntime);
    free(s->steer);
    free(s);
}

void oracle_dims(const oracle_sim *s, int64_t d[17]) {
    d[16] = s->outside;
    d[12] = s->nlayers;
    d[13] = s->layer;
    d[14] = s->zlow;
    d[15] = s->zhigh;
    d[0] = s->X;
    d[1] = s->Y;
    d[2] = s->nz;
    d[3] = s->nzA;
    d[4] = s->nx;
    d[5] = s->ny;
    d[6] = s->P;
    d[7] = s->cycle;
    d[8] = s->nn;
    d[9] = (int64_t)s->sum_front;
    d[10] = (int64_t)s->captures;
    d[11] = (int64_t)s->solidifications;
}
void oracle_params(const oracle_sim *s, double o[9]) {
    o[0] = s->A;
    o[1] = s->B;
    o[2] = s->C;
    o[3] = s->D;
    o[4] = s->dx;
    o[5] = s->dt;
    o[6] = s->G;
    o[7] = s->R;
    o[8] = 0;
    for (long long p = 0; p < s->ndom; p++)
        if (s->ct[p] == LIQUID) {
            o[8] = s->dU[p];
            break;
        }
}
int oracle_get_field(const oracle_sim *s, int which, void *dst) {
    size_t n = (size_t)s->ndom * 4;
    switch (which) {
    case ORACLE_CELLTYPE: memcpy(dst, s->ct, n); return 0;
    case ORACLE_GRAINID: memcpy(dst, s->gid, n); return 0;
    case ORACLE_CRITTIMESTEP: memcpy(dst, s->cts, n); return 0;
    case ORACLE_LAYERID: memcpy(dst, s->lid, n); return 0;
    case ORACLE_UNDERCOOLING: memcpy(dst, s->U, n); return 0;
    case ORACLE_UCHANGE: memcpy(dst, s->dU, n); return 0;
    case ORACLE_MELTED: memcpy(dst, s->melted, n); return 0;
    case ORACLE_DIAGLEN: memcpy(dst, s->DL, (size_t)s->nact * 4); return 0;
    case ORACLE_DOCENTER: memcpy(dst, s->DOC, (size_t)s->nact * 12); return 0;
    case ORACLE_CRITDIAG: memcpy(dst, s->CDL, (size_t)s->nact * 104); return 0;
    case ORACLE_NUCLEI_LOC: memcpy(dst, s->nloc, (size_t)s->P * 4); return 0;
    case ORACLE_NUCLEI_TIME: memcpy(dst, s->ntime, (size_t)s->P * 4); return 0;
    }
    return fail("unknown field %d", which);
}

/* ---- test hooks: hand-built states and planted mistakes (tests/ only) ---- */
int oracle_set_field(oracle_sim *s, int which, const void *src) {
    size_t n = (size_t)s->ndom * 4;
    switch (which) {
    case ORACLE_CELLTYPE: memcpy(s->ct, src, n); return 0;
    case ORACLE_GRAINID: memcpy(s->gid, src, n); return 0;
    case ORACLE_CRITTIMESTEP: memcpy(s->cts, src, n); return 0;
    case ORACLE_LAYERID: memcpy(s->lid, src, n); return 0;
    case ORACLE_UNDERCOOLING: memcpy(s->U, src, n); return 0;
    case ORACLE_UCHANGE: memcpy(s->dU, src, n); return 0;
    case ORACLE_MELTED: memcpy(s->melted, src, n); return 0;
    case ORACLE_DIAGLEN: memcpy(s->DL, src, (size_t)s->nact * 4); return 0;
    case ORACLE_DOCENTER: memcpy(s->DOC, src, (size_t)s->nact * 12); return 0;
    case ORACLE_CRITDIAG: memcpy(s->CDL, src, (size_t)s->nact * 104); return 0;
    }
    return fail("field %d cannot be set", which);
}
int oracle_set_orientation(oracle_sim *s, int o, const float guv[9]) {
    if (o < 0 || o >= NORIENT) return fail("orientation %d out of range", o);
    memcpy(&s->guv[9 * o], guv, 9 * sizeof(float));
    return 0;
}
void oracle_set_mutation(oracle_sim *s, int mutation) { s->mutation = mutation; }

/* ======================================================================================
 * The time step
 * ====================================================================================== */
/* Nucleation (src/CAupdate.cpp:67-200), executed in nucleus order as the Serial backend does. */
static void nucleation(oracle_sim *s, int cycle, int layer) {
    const int X = s->X, Y = s->Y;
    for (int n = 0; n < s->P; n++) {
        long long pg = s->nloc[n];
        if (!(cycle >= s->ntime[n] && s->ct[pg] == LIQUID && s->lid[pg] <= layer)) continue;
        s->ct[pg] = TEMP_UPDATE;
        int gz = (int)(pg / ((long long)X * Y));
        long long rem = pg % ((long long)X * Y);
        int rx = (int)(rem / Y), ry = (int)(rem % Y), rz = gz - s->zlow;
        if (rz < 0 || rz >= s->nzA) {
            s->ct[pg] = LIQUID;
            s->outside++;
            continue;
        }
        long long pa = ((long long)rz * X + rx) * Y + ry;
        if (rx > 0 && rx < X - 1 && ry > 0 && ry < Y - 1) s->nn++;
        int o = orientation_of(s, s->gid[pg]);
        s->DL[pa] = 0.01f;
        s->DOC[3 * pa] = (float)(rx + s->xoff + 0.5);
        s->DOC[3 * pa + 1] = (float)(ry + s->yoff + 0.5);
        s->DOC[3 * pa + 2] = (float)(gz + 0.5);
        seed_cdl(&s->guv[9 * o], s->nbx, s->nby, s->nbz, &s->CDL[26 * pa]);
        s->ct[pg] = ACTIVE;
    }
}

static inline double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */

/* Growth of one octahedron per step (src/CAupdate.cpp:257-261). */
double oracle_growth_increment(double A, double B, double C, double D, float U) {
    double u = U;
    u = dmin(210.0, u);
    double v = A * pow(u, 3.0) + B * pow(u, 2.0) + C * u + D;
    v = dmax(0.0, v);
    return dmin(0.045, v);
}

/* The captured cell's decentred octahedron (src/CAupdate.cpp:360-563). */
static void capture_octahedron(const float *g, double cxold, double cyold, double czold, double xp, double yp,
                               double zp, const int *nbx, const int *nby, const int *nbz, double c_out[3],
                               float *dl_out, float *cdl_out) {
    double x0 = xp - cxold, y0 = yp - cyold, z0 = zp - czold;
    double mag0 = sqrtf((float)(x0 * x0 + y0 * y0 + z0 * z0));
    double a1 = (g[0] * x0 + g[1] * y0 + g[2] * z0) / mag0;
    double a2 = (g[3] * x0 + g[4] * y0 + g[5] * z0) / mag0;
    double a3 = (g[6] * x0 + g[7] * y0 + g[8] * z0) / mag0;
    int s1 = 2 * (a1 < 0) - 1, s2 = 2 * (a2 < 0) - 1, s3 = 2 * (a3 < 0) - 1;
    double d1x = g[0] * (float)s1, d1y = g[1] * (float)s1, d1z = g[2] * (float)s1;
    double d2x = g[3] * (float)s2, d2y = g[4] * (float)s2, d2z = g[5] * (float)s2;
    double d3x = g[6] * (float)s3, d3y = g[7] * (float)s3, d3z = g[8] * (float)s3;
    double u1[3] = {d2x - d1x, d2y - d1y, d2z - d1z};
    double u2[3] = {d3x - d1x, d3y - d1y, d3z - d1z};
    double uu[3] = {u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2], u1[0] * u2[1] - u1[1] * u2[0]};
    double nd = sqrtf((float)(uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2]));
    double n0 = uu[0] / nd, n1 = uu[1] / nd, n2 = uu[2] / nd;
    double t = (n0 * x0 + n1 * y0 + n2 * z0) / (n0 * d1x + n1 * d1y + n2 * d1z);
    double tx[3] = {cxold + t * d1x, cxold + t * d2x, cxold + t * d3x};
    double ty[3] = {cyold + t * d1y, cyold + t * d2y, cyold + t * d3y};
    double tz[3] = {czold + t * d1z, czold + t * d2z, czold + t * d3z};
    double dc[3];
    for (int q = 0; q < 3; q++)
        dc[q] = sqrtf((float)(((tx[q] - xp) * (tx[q] - xp)) + ((ty[q] - yp) * (ty[q] - yp)) +
                              ((tz[q] - zp) * (tz[q] - zp))));
    int bx = dc[0] < dc[1], by = dc[1] < dc[2], bz = dc[2] < dc[0];
    int idx = 2 * (bz - by) * bz + (by - bx) * by;
    double mind = dc[idx], xc = tx[idx], yc = ty[idx], zc = tz[idx];
    double x1 = tx[(idx + 1) % 3], y1 = ty[(idx + 1) % 3], z1 = tz[(idx + 1) % 3];
    double x2 = tx[(idx + 2) % 3], y2 = ty[(idx + 2) % 3], z2 = tz[(idx + 2) % 3];
    double D1 = sqrtf((float)(((xp - x2) * (xp - x2)) + ((yp - y2) * (yp - y2)) + ((zp - z2) * (zp - z2))));
    double D2 = sqrtf((float)(((xc - x2) * (xc - x2)) + ((yc - y2) * (yc - y2)) + ((zc - z2) * (zc - z2))));
    double D3 = sqrtf((float)(((xp - x1) * (xp - x1)) + ((yp - y1) * (yp - y1)) + ((zp - z1) * (zp - z1))));
    double D4 = sqrtf((float)(((xc - x1) * (xc - x1)) + ((yc - y1) * (yc - y1)) + ((zc - z1) * (zc - z1))));
    double I1 = 0, I2 = D2, J1 = 0, J2 = D4;
    if (mind != 0) {
        I1 = D1 * ((xp - x2) * (xc - x2) + (yp - y2) * (yc - y2) + (zp - z2) * (zc - z2)) / (D1 * D2);
        I2 = D2 - I1;
        J1 = D3 * ((xp - x1) * (xc - x1) + (yp - y1) * (yc - y1) + (zp - z1) * (zc - z1)) / (D3 * D4);
        J2 = D4 - J1;
    }
    double L12 = 0.5 * (dmin(I1, sqrt(3.0)) + dmin(I2, sqrt(3.0)));
    double L13 = 0.5 * (dmin(J1, sqrt(3.0)) + dmin(J2, sqrt(3.0)));
    double newdl = sqrt(2.0) * dmax(L12, L13);
    *dl_out = (float)newdl;
    double cd[3] = {xc - cxold, yc - cyold, zc - czold};
    double cn = sqrt(cd[0] * cd[0] + cd[1] * cd[1] + cd[2] * cd[2]);
    double cx = xc - newdl * (cd[0] / cn), cy = yc - newdl * (cd[1] / cn), cz = zc - newdl * (cd[2] / cn);
    c_out[0] = cx;
    c_out[1] = cy;
    c_out[2] = cz;
    critical_lengths(g, xp, yp, zp, cx, cy, cz, nbx, nby, nbz, cdl_out);
}
/* The 26 critical diagonal lengths of an octahedron centred at (cx, cy, cz) around the cell
 * centre (xp, yp, zp): the largest |F_q . (x_n - c)| over the four face normals
 * (src/CAupdate.cpp:524-563), stored as float. */
static void critical_lengths(const float *g, double xp, double yp, double zp, double cx, double cy, double cz,
                             const int *nbx, const int *nby, const int *nbz, float *cdl_out) {
    double Fx[4], Fy[4], Fz[4];
    Fx[0] = g[0] + g[3] + g[6];
    Fx[1] = g[0] - g[3] + g[6];
    Fx[2] = g[0] + g[3] - g[6];
    Fx[3] = g[0] - g[3] - g[6];
    Fy[0] = g[1] + g[4] + g[7];
    Fy[1] = g[1] - g[4] + g[7];
    Fy[2] = g[1] + g[4] - g[7];
    Fy[3] = g[1] - g[4] - g[7];
    Fz[0] = g[2] + g[5] + g[8];
    Fz[1] = g[2] - g[5] + g[8];
    Fz[2] = g[2] + g[5] - g[8];
    Fz[3] = g[2] - g[5] - g[8];
    for (int n = 0; n < 26; n++) {
        double ex = xp + nbx[n] - cx, ey = yp + nby[n] - cy, ez = zp + nbz[n] - cz;
        double D[4];
        for (int q = 0; q < 4; q++) D[q] = ex * Fx[q] + ey * Fy[q] + ez * Fz[q];
        cdl_out[n] = (float)dmax(dmax(fabs(D[0]), fabs(D[1])), dmax(fabs(D[2]), fabs(D[3])));
    }
}
void oracle_capture_octahedron(const float guv[9], double cxold, double cyold, double czold, double xp,
                               double yp, double zp, double centre[3], float *diag, float cdl[26]) {
    capture_octahedron(guv, cxold, cyold, czold, xp, yp, zp, k_nbx, k_nby, k_nbz, centre, diag, cdl);
}

/* CellCapture (src/CAupdate.cpp:204-728): pass 1 over the active domain, then pass 2 over the
 * steering vector in ascending order, each capture visible to later cells at once. */
static void cell_capture(oracle_sim *s, int cycle, int layer) {
    const int X = s->X, Y = s->Y;
    const long long XY = (long long)X * Y;
    long long nsteer = 0;
    for (long long pa = 0; pa < s->nact; pa++) {
        long long pg = pa + (long long)s->zlow * XY;
        int c = s->ct[pg];
        if (s->lid[pg] <= layer && c != SOLID && cycle > s->cts[pg]) {
            s->U[pg] += s->dU[pg] * (float)((c == LIQUID) + (c == ACTIVE));
            if (c == ACTIVE) s->steer[nsteer++] = (int32_t)pa;
        }
    }
    s->sum_front += (uint64_t)nsteer;
    for (long long qq = 0; qq < nsteer; qq++) {
        /* planted mistake ORACLE_MUT_DESCENDING: visit the steering vector from its end */
        const long long q = s->mutation == ORACLE_MUT_DESCENDING ? nsteer - 1 - qq : qq;
        long long pa = s->steer[q];
        int rz = (int)(pa / XY);
        long long rem = pa % XY;
        int rx = (int)(rem / Y), ry = (int)(rem % Y), gz = rz + s->zlow;
        long long pg = ((long long)gz * X + rx) * Y + ry;
        s->DL[pa] = (float)((double)s->DL[pa] + oracle_growth_increment(s->A, s->B, s->C, s->D, s->U[pg]));
        int lcount = 0;
        int b = it_bounds(rx, ry, X, Y);
        for (int ll = 0; ll < list_len(b); ll++) {
            int l = s->itlist[b][ll];
            int mx = rx + s->nbx[l], my = ry + s->nby[l], mz = rz + s->nbz[l];
            if (mz >= s->nzA) continue;
            if (mz < 0) { /* below the window: never liquid in a well-formed run */
                if (s->ct[((long long)(mz + s->zlow) * X + mx) * Y + my] == LIQUID) s->outside++;
                continue;
            }
            long long na = ((long long)mz * X + mx) * Y + my;
            long long ng = ((long long)(mz + s->zlow) * X + mx) * Y + my;
            const int liquid = s->ct[ng] == LIQUID;
            /* planted mistake ORACLE_MUT_LCOUNT: a neighbour the cell captures now does not count */
            if (liquid && !(s->mutation == ORACLE_MUT_LCOUNT && s->DL[pa] >= s->CDL[26 * pa + l])) lcount = 1;
            if (!(s->DL[pa] >= s->CDL[26 * pa + l] && liquid)) continue;
            s->ct[ng] = TEMP_UPDATE; /* the Serial backend's compare-exchange always succeeds here */
            int h = s->gid[pg];
            int o = orientation_of(s, h);
            s->gid[ng] = h;
            double cxold = s->DOC[3 * pa], cyold = s->DOC[3 * pa + 1], czold = s->DOC[3 * pa + 2];
            double xp = (rx + s->xoff + s->nbx[l]) + 0.5, yp = (ry + s->yoff + s->nby[l]) + 0.5,
                   zp = (gz + s->nbz[l]) + 0.5;
            double c[3];
            float dl;
            capture_octahedron(&s->guv[9 * o], cxold, cyold, czold, xp, yp, zp, s->nbx, s->nby, s->nbz, c, &dl,
                               &s->CDL[26 * na]);
            s->DL[na] = dl;
            s->DOC[3 * na] = (float)c[0];
            s->DOC[3 * na + 1] = (float)c[1];
            s->DOC[3 * na + 2] = (float)c[2];
            s->ct[ng] = ACTIVE;
            s->captures++;
            /* planted mistake ORACLE_MUT_GROW_NEW: the new cell joins this step's steering vector */
            if (s->mutation == ORACLE_MUT_GROW_NEW) s->steer[nsteer++] = (int32_t)na;
        }
        if (lcount == 0) {
            s->ct[pg] = SOLID;
            s->solidifications++;
        }
    }
}

/* CellCapture as the reference's Kokkos OpenMP backend executes it (src/CAupdate.cpp:204-728):
 * pass 1 in parallel, each active cell appended to the steering vector with an atomic fetch-add
 * (:238-240, so in no fixed order); pass 2 in parallel over the steering vector, a capture taken
 * by compare-exchange of the neighbour's type from Liquid to TemporaryUpdate (:336-352), the
 * neighbour types read while other threads update them (:333-334). Which of two cells reaching
 * one liquid cell in a step captures it then depends on thread timing, so results are not
 * reproducible: this is the timing baseline, never the parity checker. */
static void cell_capture_omp(oracle_sim *s, int cycle, int layer) {
    const int X = s->X, Y = s->Y;
    const long long XY = (long long)X * Y;
    long long nsteer = 0;
    uint64_t ncap = 0, nsol = 0;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (long long pa = 0; pa < s->nact; pa++) {
        long long pg = pa + (long long)s->zlow * XY;
        int c = s->ct[pg];
        if (s->lid[pg] <= layer && c != SOLID && cycle > s->cts[pg]) {
            s->U[pg] += s->dU[pg] * (float)((c == LIQUID) + (c == ACTIVE));
            if (c == ACTIVE) s->steer[__atomic_fetch_add(&nsteer, 1, __ATOMIC_RELAXED)] = (int32_t)pa;
        }
    }
    s->sum_front += (uint64_t)nsteer;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) reduction(+ : ncap, nsol)
    for (long long q = 0; q < nsteer; q++) {
        long long pa = s->steer[q];
        int rz = (int)(pa / XY);
        long long rem = pa % XY;
        int rx = (int)(rem / Y), ry = (int)(rem % Y), gz = rz + s->zlow;
        long long pg = ((long long)gz * X + rx) * Y + ry;
        s->DL[pa] = (float)((double)s->DL[pa] + oracle_growth_increment(s->A, s->B, s->C, s->D, s->U[pg]));
        int lcount = 0;
        int b = it_bounds(rx, ry, X, Y);
        for (int ll = 0; ll < list_len(b); ll++) {
            int l = s->itlist[b][ll];
            int mx = rx + s->nbx[l], my = ry + s->nby[l], mz = rz + s->nbz[l];
            if (mz >= s->nzA) continue;
            if (mz < 0) { /* below the window: never liquid in a well-formed run */
                if (s->ct[((long long)(mz + s->zlow) * X + mx) * Y + my] == LIQUID) s->outside++;
                continue;
            }
            long long na = ((long long)mz * X + mx) * Y + my;
            long long ng = ((long long)(mz + s->zlow) * X + mx) * Y + my;
            const int liquid = __atomic_load_n(&s->ct[ng], __ATOMIC_RELAXED) == LIQUID;
            if (liquid) lcount = 1;
            if (!(s->DL[pa] >= s->CDL[26 * pa + l] && liquid)) continue;
            int32_t expected = LIQUID;
            if (!__atomic_compare_exchange_n(&s->ct[ng], &expected, TEMP_UPDATE, 0, __ATOMIC_ACQ_REL,
                                             __ATOMIC_RELAXED))
                continue;
            int h = s->gid[pg];
            int o = orientation_of(s, h);
            s->gid[ng] = h;
            double cxold = s->DOC[3 * pa], cyold = s->DOC[3 * pa + 1], czold = s->DOC[3 * pa + 2];
            double xp = (rx + s->xoff + s->nbx[l]) + 0.5, yp = (ry + s->yoff + s->nby[l]) + 0.5,
                   zp = (gz + s->nbz[l]) + 0.5;
            double c[3];
            float dl;
            capture_octahedron(&s->guv[9 * o], cxold, cyold, czold, xp, yp, zp, s->nbx, s->nby, s->nbz, c, &dl,
                               &s->CDL[26 * na]);
            s->DL[na] = dl;
            s->DOC[3 * na] = (float)c[0];
            s->DOC[3 * na + 1] = (float)c[1];
            s->DOC[3 * na + 2] = (float)c[2];
            __atomic_store_n(&s->ct[ng], ACTIVE, __ATOMIC_RELEASE);
            ncap++;
        }
        if (lcount == 0) {
            __atomic_store_n(&s->ct[pg], SOLID, __ATOMIC_RELAXED);
            nsol++;
        }
    }
    s->captures += ncap;
    s->solidifications += nsol;
}

void oracle_set_threads(oracle_sim *s, int nthreads) { s->nthreads = nthreads; }

static void capture_step(oracle_sim *s, int cycle, int layer) {
    if (s->nthreads > 1)
        cell_capture_omp(s, cycle, layer);
    else
        cell_capture(s, cycle, layer);
}

void oracle_step(oracle_sim *s, int cycle) {
    nucleation(s, cycle, s->layer);
    capture_step(s, cycle, s->layer);
    s->cycle = cycle;
}

/* The transition to the next layer (src/runCA.cpp:343-418): DomainShiftAndResize
 * (src/CAinitialize.cpp:2772-2832) moves the active window's bottom to just below the lowest
 * active cell and its top to the highest cell of the next layer; LayerSetup (:2835-2982) zeroes
 * the window's octahedron arrays and gives every active cell in it a cell-centred octahedron
 * (half-diagonal 0.01, critical lengths of its orientation). The cycle count restarts at 0. */
int oracle_next_layer(oracle_sim *s) {
    if (s->layer + 1 >= s->nlayers) return fail("no layer after layer %d", s->layer);
    const long long XY = (long long)s->X * s->Y;
    int newmin = 0x7FFFFFFF, newmax = (int)0x80000000;
    for (long long p = 0; p < s->ndom; p++) {
        const int z = (int)(p / XY);
        if (s->ct[p] == ACTIVE && z < newmin) newmin = z;
        if (s->lid[p] == s->layer + 1 && z > newmax) newmax = z;
    }
    s->zlow = newmin - 1;
    s->zhigh = newmax;
    s->nzA = s->zhigh - s->zlow + 1;
    if (s->zlow < 0 || s->nzA < 1 || s->zhigh >= s->nz) return fail("no active window for layer %d", s->layer + 1);
    s->nact = XY * s->nzA;
    free(s->DL);
    free(s->DOC);
    free(s->CDL);
    free(s->steer);
    s->DL = calloc((size_t)s->nact, 4);
    s->DOC = calloc((size_t)s->nact * 3, 4);
    s->CDL = calloc((size_t)s->nact * 26, 4);
    s->steer = calloc((size_t)s->nact, 4);
    if (!s->DL || !s->DOC || !s->CDL || !s->steer) return fail("out of host memory");
    for (long long pa = 0; pa < s->nact; pa++) {
        const long long pg = pa + (long long)s->zlow * XY;
        if (s->ct[pg] != ACTIVE) continue;
        const int rz = (int)(pa / XY);
        const long long rem = pa % XY;
        const int rx = (int)(rem / s->Y), ry = (int)(rem % s->Y);
        const int o = orientation_of(s, s->gid[pg]);
        s->DL[pa] = 0.01f;
        s->DOC[3 * pa] = (float)(rx + s->xoff + 0.5);
        s->DOC[3 * pa + 1] = (float)(ry + s->yoff + 0.5);
        s->DOC[3 * pa + 2] = (float)(rz + s->zlow + 0.5);
        seed_cdl(&s->guv[9 * o], s->nbx, s->nby, s->nbz, &s->CDL[26 * pa]);
    }
    s->layer++;
    s->cycle = 0;
    s->nn_layer = s->nn;
    return 0;
}

void oracle_check(oracle_sim *s, int *cycle, uint64_t counts[5], int *xswitch, int verbose) {
    const int layer = s->layer;
    uint64_t c[5] = {0, 0, 0, 0, 0};
    for (long long p = 0; p < s->ndom; p++) {
        if (s->lid[p] == layer) {
            if (s->ct[p] == LIQUID) {
                if (s->cts[p] > *cycle)
                    c[0]++;
                else
                    c[1]++;
            }
            else if (s->ct[p] == ACTIVE)
                c[2]++;
            else if (s->ct[p] == SOLID)
                c[3]++;
        }
        else if (s->ct[p] == LIQUID)
            c[4]++;
    }
    memcpy(counts, c, sizeof c);
    if (verbose)
        printf("cycle = %d Superheated liquid cells = %lu Undercooled liquid cells = %lu Number of nucleation "
               "events this layer %lld Remaining liquid cells in future layers of this simulation = %lu\n",
               *cycle, (unsigned long)c[0], (unsigned long)c[1], s->nn - s->nn_layer, (unsigned long)c[4]);
    if (verbose) fflush(stdout);
    if (c[0] + c[1] == 0) *xswitch = 1;
    if (*xswitch == 0 && strcmp(s->type, "S") == 0 && c[1] == 0) {
        unsigned long next = (unsigned long)-1;
        const long long XY = (long long)s->X * s->Y;
        for (long long pa = 0; pa < s->nact; pa++) {
            long long pg = pa + (long long)s->zlow * XY;
            unsigned long v = (unsigned long)(long)s->cts[pg];
            if (s->ct[pg] == LIQUID && s->lid[pg] == layer && v < next) next = v;
        }
        if ((next - (unsigned long)(long)*cycle) > 5000) {
            *cycle = (int)(next - 1);
            s->cycle = *cycle;
            if (verbose) printf("Jumping to cycle %d\n", *cycle + 1);
            if (verbose) fflush(stdout);
        }
    }
}

int oracle_run(oracle_sim *s, int max_cycles, int verbose, double *capture_seconds) {
    int cycle = s->cycle, xswitch = 0;
    double tcap = 0;
    for (;;) {
        cycle++;
        nucleation(s, cycle, s->layer);
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        capture_step(s, cycle, s->layer);
        clock_gettime(CLOCK_MONOTONIC, &t1);
        tcap += (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
        s->cycle = cycle;
        if (cycle % 1000 == 0) {
            uint64_t counts[5];
            oracle_check(s, &cycle, counts, &xswitch, verbose);
            s->cycle = cycle;
        }
        if (max_cycles > 0 && cycle >= max_cycles) break;
        if (xswitch) {
            /* a whole run goes on through every layer (src/runCA.cpp:283-418) */
            if (max_cycles != 0 || s->layer + 1 >= s->nlayers) break;
            if (oracle_next_layer(s)) break;
            if (verbose) printf("New layer %d, active window z = %d .. %d\n", s->layer, s->zlow, s->zhigh);
            cycle = 0;
            xswitch = 0;
        }
    }
    if (capture_seconds) *capture_seconds = tcap;
    return cycle;
}

/* ======================================================================================
 * Output (src/CAprint.cpp:285-571), np = 1, whole domain gathered as (k, i, j) with the ghost
 * ring dropped and k = 0, nz-1 left at zero.
 * ====================================================================================== */
static int32_t whole(const oracle_sim *s, const int32_t *f, int k, int i, int j) {
    if (k < 1 || k > s->nz - 2) return 0;
    return f[((long long)k * s->X + (i + 1)) * s->Y + (j + 1)];
}
int oracle_write_outputs(const oracle_sim *s, const char *deck_path) {
    char fname[4096];
    if (s->print_mis) {
        snprintf(fname, sizeof fname, "%s%s_Misorientations.vtk", s->path_out, s->base);
        FILE *f = fopen(fname, "w");
        if (!f) return fail("cannot write %s", fname);
        fprintf(f, "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET STRUCTURED_POINTS\n");
        fprintf(f, "DIMENSIONS %d %d %d\nORIGIN 0 0 0\nSPACING 1 1 1\n", s->nx - 2, s->ny - 2, s->nz - 2);
        fprintf(f, "POINT_DATA %d\nSCALARS Angle_z int 1\nLOOKUP_TABLE default\n",
                (s->nx - 2) * (s->ny - 2) * (s->nz - 2));
        int *rounded = malloc(sizeof(int) * NORIENT);
        for (int n = 0; n < NORIENT; n++) {
            double amin = 62.7;
            for (int ll = 0; ll < 3; ll++) {
                double a = fabs((180 / M_PI) * acos(s->guv[9 * n + 3 * ll + 2]));
                if (a < amin) amin = a;
            }
            rounded[n] = (int)round(amin);
        }
        for (int k = 1; k < s->nz - 1; k++) {
            for (int j = 1; j < s->ny - 1; j++)
                for (int i = 1; i < s->nx - 1; i++) {
                    if (whole(s, s->melted, k, i, j) == 0)
                        fprintf(f, "200 ");
                    else {
                        int g = whole(s, s->gid, k, i, j);
                        int o = orientation_of(s, g);
                        fprintf(f, "%d ", g < 0 ? rounded[o] + 100 : rounded[o]);
                    }
                }
            fprintf(f, "\n");
        }
        free(rounded);
        fclose(f);
    }
    if (s->print_mis || s->print_full) {
        /* Without the full-output flag the reference names this file _debug.vtk and writes only
         * LayerID, which it never gathered (its whole-domain view is empty); zeros are written. */
        snprintf(fname, sizeof fname, "%s%s%s.vtk", s->path_out, s->base, s->print_full ? "" : "_debug");
        FILE *f = fopen(fname, "w");
        if (!f) return fail("cannot write %s", fname);
        fprintf(f, "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET STRUCTURED_POINTS\n");
        fprintf(f, "DIMENSIONS %d %d %d\nORIGIN 0 0 0\nSPACING 1 1 1\n", s->nx, s->ny, s->nz);
        fprintf(f, "POINT_DATA %d\n", s->nx * s->ny * s->nz);
        const int32_t *fields[3] = {s->lid, s->gid, s->melted};
        const char *names[3] = {"LayerID", "GrainID", "Melted"};
        int nf = s->print_full ? 3 : 1;
        for (int q = 0; q < nf; q++) {
            fprintf(f, "SCALARS %s int 1\nLOOKUP_TABLE default\n", names[q]);
            for (int k = 0; k < s->nz; k++) {
                for (int j = 0; j < s->ny; j++)
                    for (int i = 0; i < s->nx; i++) {
                        int v = s->print_full ? whole(s, fields[q], k, i, j) : 0;
                        if (q == 2) v = v ? 1 : 0;
                        fprintf(f, "%d ", v);
                    }
                fprintf(f, "\n");
            }
        }
        fclose(f);
    }
    snprintf(fname, sizeof fname, "%s%s.log", s->path_out, s->base);
    FILE *f = fopen(fname, "w");
    if (!f) return fail("cannot write %s", fname);
    fprintf(f, "log file for a simulation run with input file %s  run on 1 MPI ranks, output written at cycle %d\n",
            deck_path, s->cycle);
    fprintf(f, "This simulation took 0 seconds to run, with the init/run/output breakdown as 0/0/0\n");
    fprintf(f, "This simulation was type: %s\n", s->type);
    fprintf(f, "Domain size in x: %d\nDomain size in y: %d\nDomain size in z: %d\n", s->nx, s->ny, s->nz);
    fprintf(f, "Cell size: %g microns\n", s->dx);
    fprintf(f, "Time step: %g microseconds\n", s->dt);
    fprintf(f,
            "Nucleation density was %g m^-3 , mean nucleation undercooling was %g K, and standard deviation of "
            "nucleation undercooling was %g K\n",
            s->NMax, s->dTN, s->dTs);
    fprintf(f, "Interfacial response function parameters used were %g , %g , %g , %g , and the alloy freezing range "
               "was %g\n",
            s->A, s->B, s->C, s->D, s->FR);
    if (strcmp(s->type, "C") == 0) {
        fprintf(f, "The thermal gradient was %g K/m, and the cooling rate %g K/s\n", s->G, s->R);
        fprintf(f, "The fraction of surface sites active was %g\n", s->fract);
    }
    else {
        fprintf(f, "A total of %d in X and %d in Y were considered\n", s->nsx, s->nsy);
        fprintf(f, "The spots were offset by %d microns, and had radii of %d microns\n", s->soff, s->srad);
        fprintf(f, "This pattern was repeated for %d layers\n", s->nlayers);
    }
    fprintf(f, "The decomposition scheme used was: %d\n", s->decomp);
    fprintf(f, "Rank 0 contained %d cells in x , %d cells in y; subdomain was offset by %d in x , %d in y\n", s->X,
            s->Y, s->xoff, s->yoff);
    fprintf(f, "Max/min rank time initializing data  = 0 / 0 s\n");
    fprintf(f, "Max/min rank time in CA nucleation   = 0 / 0 s\n");
    fprintf(f, "Max/min rank time in CA cell capture = 0 / 0 s\n");
    fprintf(f, "Max/min rank time in CA ghosting     = 0 / 0 s\n");
    fprintf(f, "Max/min rank time exporting data     = 0 / 0 s\n\n");
    fclose(f);
    return 0;
}
