// ca_kernels.hip -- the CA time step as memory-bound CDNA4 kernels over the cells due this step.
//
// One step (reference: src/runCA.cpp:291-311) is six launches on one stream (five when no
// nucleation site can be due):
//   k_nucleate  activates the pre-sampled nuclei whose time has come (src/CAupdate.cpp:67-200)
//   k_claim     for every cell due this step: the growth since it was last looked at, a claim on
//               every liquid neighbour it reaches (atomicMin of its index into the neighbour's
//               claim word), and the step at which it is due next; a cell past its liquidus time
//               with no liquid neighbour turns solid
//   k_win       the claimants that won a claim (their index is in the claimed cell's word)
//   k_apply     the winner of each claimed cell builds the new decentred octahedron, stores its
//               26 critical lengths, and takes the cell out of its neighbours' liquid masks
//   k_settle    a cell whose liquid neighbours were all taken by lower-indexed cells turns solid
//   k_rebuild   the list of the cells due at the next step, in index order, from its calendar slot
// A cell is due when something can happen to it: it is new, its liquidus time has passed, or its
// half-diagonal reaches the least critical length of its liquid neighbours; a neighbour's capture
// or nucleation that empties its mask is handled where it happens. Between those steps its growth
// is a fixed recurrence, applied exactly when it is next looked at.
// Claim-then-apply reproduces the reference's Serial backend exactly: there pass 2 visits active
// cells in ascending index order and every capture is visible at once, so a liquid cell goes to
// the lowest-indexed cell that reaches it, and cell j still "sees" a liquid neighbour only if no
// lower-indexed cell took it (src/CAupdate.cpp:264-352, 722-725).
//
// Every floating-point expression keeps the reference's operation order and float/double types;
// the library is compiled with -ffp-contract=off and correctly rounded fp32 divide/sqrt.
#include <hip/hip_runtime.h>

#include "ca_kernels.hpp"
#include "growth.hpp"
#include "threshold.hpp"

namespace exaca {
namespace {

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max

__device__ __forceinline__ int orientation_of(int gid) { return (abs(gid) - 1) % kNumOrientations; }

// Steps executed in (t, cycle] of one layer whose cycle jumps are entries j0 .. j1-1: the reference
// adds the undercooling change once per executed step while a cell is liquid or active past its
// liquidus time (src/CAupdate.cpp:235-237); cycle jumps (src/CAupdate.cpp:812-814) skip steps
// J+1 .. T-1.
__device__ __forceinline__ int executed_in(int t, int cycle, const StepCtl *ctl, int j0, int j1) {
    int n = cycle - t;
    for (int q = j0; q < j1; q++) {
        const int lo = max(ctl->jump_from[q] + 1, t + 1);
        const int hi = min(ctl->jump_to[q] - 1, cycle);
        if (hi >= lo) n -= hi - lo + 1;
    }
    return n;
}

struct Coord {
    int x, y, z;
};
// Unsigned division by a runtime-uniform divisor as a multiply-high and shifts (the round-up
// method; exact for every 32-bit dividend). mul/shift come from Divisor::make on the host side of
// the kernel (uniform, so they live in scalar registers).
struct Divisor {
    uint32_t d, mul, sh1, sh2;
    __host__ __device__ static Divisor make(uint32_t d) {
        Divisor r;
        r.d = d;
        uint32_t l = 0;
        while ((1ull << l) < d) l++;
        r.mul = static_cast<uint32_t>(((1ull << 32) * ((1ull << l) - d)) / d + 1);
        r.sh1 = l > 0 ? 1 : 0;
        r.sh2 = l > 0 ? l - 1 : 0;
        return r;
    }
    __device__ __forceinline__ uint32_t div(uint32_t n) const {
        const uint32_t t = __umulhi(n, mul);
        return (t + ((n - t) >> sh1)) >> sh2;
    }
};
// Offsets of neighbour l (a per-lane value) decoded from the packed constants.
__device__ __forceinline__ int nb_dx(int l) { return static_cast<int>((kPackX >> (2 * l)) & 3u) - 1; }
__device__ __forceinline__ int nb_dy(int l) { return static_cast<int>((kPackY >> (2 * l)) & 3u) - 1; }
__device__ __forceinline__ int nb_dz(int l) { return static_cast<int>((kPackZ >> (2 * l)) & 3u) - 1; }

struct Grid {
    Divisor xy, y;
    int X, Y, nzA;
    int off[26];  // linear offsets of the 26 neighbours, for loops unrolled over l
    // linear offset of neighbour l for a per-lane l
    __device__ __forceinline__ int offset(int l) const {
        return nb_dz(l) * static_cast<int>(xy.d) + nb_dx(l) * Y + nb_dy(l);
    }
};
__device__ __forceinline__ Grid grid_of(const DomainView &d) {
    Grid g;
    g.X = d.X;
    g.Y = d.Y;
    g.nzA = d.nzA;
    g.xy = Divisor::make(static_cast<uint32_t>(d.X) * static_cast<uint32_t>(d.Y));
    g.y = Divisor::make(static_cast<uint32_t>(d.Y));
#pragma unroll
    for (int l = 0; l < 26; l++) g.off[l] = (kNbZ[l] * d.X + kNbX[l]) * d.Y + kNbY[l];
    return g;
}
__device__ __forceinline__ Coord decode(uint32_t p, const Grid &g) {
    Coord c;
    const uint32_t z = g.xy.div(p);
    const uint32_t rem = p - z * g.xy.d;
    const uint32_t x = g.y.div(rem);
    c.z = static_cast<int>(z);
    c.x = static_cast<int>(x);
    c.y = static_cast<int>(rem - x * g.y.d);
    return c;
}
__device__ __forceinline__ Coord decode(uint32_t p, int X, int Y) {
    const uint32_t XY = static_cast<uint32_t>(X) * static_cast<uint32_t>(Y);
    Coord c;
    c.z = static_cast<int>(p / XY);
    const uint32_t rem = p - static_cast<uint32_t>(c.z) * XY;
    c.x = static_cast<int>(rem / static_cast<uint32_t>(Y));
    c.y = static_cast<int>(rem - static_cast<uint32_t>(c.x) * static_cast<uint32_t>(Y));
    return c;
}
// Neighbours inside the rank's slab and below the active-domain top; the reference's edge lists
// (src/CAupdate.cpp:276-318) and its MyNeighborZ < nzActive test (:328). z >= 0 always holds for
// an active cell (plane 0 is a wall); it is tested for memory safety.
__device__ __forceinline__ uint32_t valid_neighbours(const Coord &c, int X, int Y, int nzA) {
    if (c.x > 0 && c.x < X - 1 && c.y > 0 && c.y < Y - 1 && c.z > 0 && c.z < nzA - 1) return (1u << 26) - 1;
    uint32_t m = 0;
#pragma unroll
    for (int l = 0; l < 26; l++) {
        const int x = c.x + kNbX[l], y = c.y + kNbY[l], z = c.z + kNbZ[l];
        if (x >= 0 && x < X && y >= 0 && y < Y && z >= 0 && z < nzA) m |= 1u << l;
    }
    return m;
}

// Face normals of the octahedron as the reference forms them: float sums of unit-vector
// components (src/CAupdate.cpp:526-551).
struct Faces {
    double fx[4], fy[4], fz[4];
};
__device__ __forceinline__ Faces faces_of(const float *g) {
    Faces f;
    f.fx[0] = g[0] + g[3] + g[6];
    f.fx[1] = g[0] - g[3] + g[6];
    f.fx[2] = g[0] + g[3] - g[6];
    f.fx[3] = g[0] - g[3] - g[6];
    f.fy[0] = g[1] + g[4] + g[7];
    f.fy[1] = g[1] - g[4] + g[7];
    f.fy[2] = g[1] + g[4] - g[7];
    f.fy[3] = g[1] - g[4] - g[7];
    f.fz[0] = g[2] + g[5] + g[8];
    f.fz[1] = g[2] - g[5] + g[8];
    f.fz[2] = g[2] + g[5] - g[8];
    f.fz[3] = g[2] - g[5] - g[8];
    return f;
}
// Critical diagonal length of neighbour l for an octahedron centred at c around the cell centre
// (xp, yp, zp): the largest |F_q . (x_l - c)| (src/CAupdate.cpp:553-563), rounded to float as
// the reference stores it.
__device__ __forceinline__ float critical_length(const Faces &f, double xp, double yp, double zp, double cx,
                                                 double cy, double cz, int l) {
    const double ex = xp + nb_dx(l) - cx, ey = yp + nb_dy(l) - cy, ez = zp + nb_dz(l) - cz;
    const double d0 = ex * f.fx[0] + ey * f.fy[0] + ez * f.fz[0];
    const double d1 = ex * f.fx[1] + ey * f.fy[1] + ez * f.fz[1];
    const double d2 = ex * f.fx[2] + ey * f.fy[2] + ez * f.fz[2];
    const double d3 = ex * f.fx[3] + ey * f.fy[3] + ez * f.fz[3];
    return static_cast<float>(dmax(dmax(fabs(d0), fabs(d1)), dmax(fabs(d2), fabs(d3))));
}

// The centre of cell q's octahedron record (the last 24 bytes of its line).
__device__ __forceinline__ void set_centre(const DomainView &d, uint32_t q, double x, double y, double z) {
    Octa &o = d.octa[q];
    o.x = x;
    o.y = y;
    o.z = z;
}

struct Octahedron {
    double cx, cy, cz, dl;
};
__device__ __forceinline__ double sel3(int i, double a, double b, double c) { return i == 0 ? a : (i == 1 ? b : c); }

// The captured cell's decentred octahedron (src/CAupdate.cpp:360-515): the capturing face is
// found from the signs of the unit-vector projections, the face corner nearest the new cell
// centre is the new octahedron's tip, and its half-diagonal follows from the projections of the
// centre onto the two face edges through that corner.
__device__ Octahedron new_octahedron(const float *g, double cxold, double cyold, double czold, double xp, double yp,
                                     double zp) {
    const double x0 = xp - cxold, y0 = yp - cyold, z0 = zp - czold;
    const double mag0 = sqrtf(static_cast<float>(x0 * x0 + y0 * y0 + z0 * z0));
    const double a1 = (g[0] * x0 + g[1] * y0 + g[2] * z0) / mag0;
    const double a2 = (g[3] * x0 + g[4] * y0 + g[5] * z0) / mag0;
    const double a3 = (g[6] * x0 + g[7] * y0 + g[8] * z0) / mag0;
    const float s1 = (a1 < 0) ? 1.0f : -1.0f, s2 = (a2 < 0) ? 1.0f : -1.0f, s3 = (a3 < 0) ? 1.0f : -1.0f;
    const double d1x = g[0] * s1, d1y = g[1] * s1, d1z = g[2] * s1;
    const double d2x = g[3] * s2, d2y = g[4] * s2, d2z = g[5] * s2;
    const double d3x = g[6] * s3, d3y = g[7] * s3, d3z = g[8] * s3;
    const double u10 = d2x - d1x, u11 = d2y - d1y, u12 = d2z - d1z;
    const double u20 = d3x - d1x, u21 = d3y - d1y, u22 = d3z - d1z;
    const double uu0 = u11 * u22 - u12 * u21, uu1 = u12 * u20 - u10 * u22, uu2 = u10 * u21 - u11 * u20;
    const double nd = sqrtf(static_cast<float>(uu0 * uu0 + uu1 * uu1 + uu2 * uu2));
    const double n0 = uu0 / nd, n1 = uu1 / nd, n2 = uu2 / nd;
    const double t = (n0 * x0 + n1 * y0 + n2 * z0) / (n0 * d1x + n1 * d1y + n2 * d1z);
    const double tx0 = cxold + t * d1x, ty0 = cyold + t * d1y, tz0 = czold + t * d1z;
    const double tx1 = cxold + t * d2x, ty1 = cyold + t * d2y, tz1 = czold + t * d2z;
    const double tx2 = cxold + t * d3x, ty2 = cyold + t * d3y, tz2 = czold + t * d3z;
    const double dc0 = sqrtf(static_cast<float>(((tx0 - xp) * (tx0 - xp)) + ((ty0 - yp) * (ty0 - yp)) +
                                                ((tz0 - zp) * (tz0 - zp))));
    const double dc1 = sqrtf(static_cast<float>(((tx1 - xp) * (tx1 - xp)) + ((ty1 - yp) * (ty1 - yp)) +
                                                ((tz1 - zp) * (tz1 - zp))));
    const double dc2 = sqrtf(static_cast<float>(((tx2 - xp) * (tx2 - xp)) + ((ty2 - yp) * (ty2 - yp)) +
                                                ((tz2 - zp) * (tz2 - zp))));
    const int bx = dc0 < dc1, by = dc1 < dc2, bz = dc2 < dc0;
    const int idx = 2 * (bz - by) * bz + (by - bx) * by;
    const int i1 = (idx + 1) % 3, i2 = (idx + 2) % 3;
    const double mind = sel3(idx, dc0, dc1, dc2);
    const double xc = sel3(idx, tx0, tx1, tx2), yc = sel3(idx, ty0, ty1, ty2), zc = sel3(idx, tz0, tz1, tz2);
    const double x1 = sel3(i1, tx0, tx1, tx2), y1 = sel3(i1, ty0, ty1, ty2), z1 = sel3(i1, tz0, tz1, tz2);
    const double x2 = sel3(i2, tx0, tx1, tx2), y2 = sel3(i2, ty0, ty1, ty2), z2 = sel3(i2, tz0, tz1, tz2);
    const double D1 =
        sqrtf(static_cast<float>(((xp - x2) * (xp - x2)) + ((yp - y2) * (yp - y2)) + ((zp - z2) * (zp - z2))));
    const double D2 =
        sqrtf(static_cast<float>(((xc - x2) * (xc - x2)) + ((yc - y2) * (yc - y2)) + ((zc - z2) * (zc - z2))));
    const double D3 =
        sqrtf(static_cast<float>(((xp - x1) * (xp - x1)) + ((yp - y1) * (yp - y1)) + ((zp - z1) * (zp - z1))));
    const double D4 =
        sqrtf(static_cast<float>(((xc - x1) * (xc - x1)) + ((yc - y1) * (yc - y1)) + ((zc - z1) * (zc - z1))));
    double I1 = 0, I2 = D2, J1 = 0, J2 = D4;
    if (mind != 0) {
        I1 = D1 * ((xp - x2) * (xc - x2) + (yp - y2) * (yc - y2) + (zp - z2) * (zc - z2)) / (D1 * D2);
        I2 = D2 - I1;
        J1 = D3 * ((xp - x1) * (xc - x1) + (yp - y1) * (yc - y1) + (zp - z1) * (zc - z1)) / (D3 * D4);
        J2 = D4 - J1;
    }
    const double sqrt3 = sqrt(3.0);
    const double L12 = 0.5 * (dmin(I1, sqrt3) + dmin(I2, sqrt3));
    const double L13 = 0.5 * (dmin(J1, sqrt3) + dmin(J2, sqrt3));
    const double newdl = sqrt(2.0) * dmax(L12, L13);
    const double c0 = xc - cxold, c1 = yc - cyold, c2 = zc - czold;
    const double cn = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    Octahedron o;
    o.cx = xc - newdl * (c0 / cn);
    o.cy = yc - newdl * (c1 / cn);
    o.cz = zc - newdl * (c2 / cn);
    o.dl = newdl;
    return o;
}

// Wave64 inclusive prefix sum of v; returns the inclusive sum and sets total to the wave sum.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v, uint32_t &total) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t n = __shfl_up(v, o, 64);
        if (lane >= o) v += n;
    }
    total = __shfl(v, 63, 64);
    return v;
}

__device__ __forceinline__ unsigned long long block_sum(unsigned long long v, unsigned long long *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int q = 0; q < kBlock / 64; q++) t += sh[q];
    return t;
}

// Contiguous slice of the frontier for workgroup b of the nb serving one domain. Workgroups are
// dispatched round-robin over the 8 XCDs, so consecutive slices are given to workgroups on the
// same XCD: neighbouring frontier cells (and their neighbour reads and claims) then share one L2.
__device__ __forceinline__ void frontier_slice(uint32_t count, uint32_t cap, uint32_t b, uint32_t nb,
                                               uint32_t &begin, uint32_t &end) {
    const uint32_t n = count < cap ? count : cap;  // an overflowed list is flagged, never over-read
    const uint32_t logical = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    uint32_t chunk = (n + nb - 1) / nb;
    chunk = (chunk + kBlock - 1) / kBlock * kBlock;
    const unsigned long long b0 = static_cast<unsigned long long>(logical) * chunk;
    begin = b0 >= n ? n : static_cast<uint32_t>(b0);
    end = (b0 + chunk >= n) ? n : static_cast<uint32_t>(b0 + chunk);
}

// Wave-aggregated append: every lane with `want` gets its own slot of the list behind `counter`
// (global or LDS), at one atomic per wave. Every lane of the wave must reach the call.
__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t *counter) {
    const unsigned long long b = __ballot(want);
    if (b == 0) return 0;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(static_cast<long long>(b)) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(b)));
    base = __shfl(base, leader, 64);
    return base + static_cast<uint32_t>(__popcll(b & ((1ull << lane) - 1)));
}

// A cell record read as two 16-byte loads (one 32-byte sector).
struct CellWords {
    unsigned long long lm;
    float dl, thr;
    int32_t cts, gid;
    uint32_t win;
    int8_t lid;
    uint8_t ct;
    uint16_t flags;
};
__device__ __forceinline__ CellWords load_cell(const Cell *c) {
    const uint4 a = reinterpret_cast<const uint4 *>(c)[0];
    const uint4 b = reinterpret_cast<const uint4 *>(c)[1];
    CellWords w;
    w.lm = static_cast<unsigned long long>(a.x) | (static_cast<unsigned long long>(a.y) << 32);
    w.dl = __uint_as_float(a.z);
    w.thr = __uint_as_float(a.w);
    w.cts = static_cast<int32_t>(b.x);
    w.gid = static_cast<int32_t>(b.y);
    w.win = b.z;
    w.lid = static_cast<int8_t>(b.w & 0xFFu);
    w.ct = static_cast<uint8_t>((b.w >> 8) & 0xFFu);
    w.flags = static_cast<uint16_t>(b.w >> 16);
    return w;
}

// The low word of a cell's liquid-neighbour mask: bits 0-25 and the not-yet-known bit.
__device__ __forceinline__ uint32_t *mask_word(const DomainView &d, uint32_t q) {
    return reinterpret_cast<uint32_t *>(&d.cells[q].lm);
}
// The high word: serial + 1 of the last step in which a capture by a cell of index >= q took one
// of q's liquid neighbours (0: none yet). Written by the capture pass only where the low word is
// known, so the 64-bit stores of a mask build (which clear it) never meet it in one step.
__device__ __forceinline__ uint32_t *keep_word(const DomainView &d, uint32_t q) {
    return reinterpret_cast<uint32_t *>(&d.cells[q].lm) + 1;
}
constexpr uint32_t kMaskUnknown = static_cast<uint32_t>(kLmUnknown);
constexpr uint32_t kMaskNow = static_cast<uint32_t>(kLmNow);
// A liquid cell's words: the low word keeps bits 26-31 of its initial all-ones value (only bits
// 0-25 are ever cleared, by nucleation), bit 30 marking it liquid; the high word is its claim word:
// the lowest index of a cell that claimed it in the current step (kClaimNone: none). A claimed
// cell is captured in the same step and never turns liquid again, so a claim word is read only in
// the step that wrote it, and the capture overwrites both words.
constexpr uint32_t kLiquidMark = 1u << 30;
__device__ __forceinline__ uint32_t *claim_word(const DomainView &d, uint32_t q) { return keep_word(d, q); }
__device__ __forceinline__ bool is_unclaimed_liquid(unsigned long long lm) {
    return (static_cast<uint32_t>(lm) & kLiquidMark) && static_cast<uint32_t>(lm >> 32) == kClaimNone;
}

// ---- Growth between events. A cell's half-diagonal grows by grow[k] at the k-th executed step
// past its liquidus time (src/CAupdate.cpp:257-261). Between two events of a cell the growth is
// this fixed recurrence, so it is applied when the cell is next looked at, one float rounding per
// step exactly as the reference stores it. ----

// k of the step at `cycle` of the current layer for cell q: the steps executed past its liquidus
// time in every layer from its own LayerID on in which it lay inside that layer's active window
// (pass 1 visits the window only and accumulates only while LayerID <= layer, and the undercooling
// is never reset between layers; src/CAupdate.cpp:217-237, src/runCA.cpp:283-418). 0 while the cell
// waits for its liquidus time or its layer.
__device__ __forceinline__ int growth_index(const DomainView &d, int cts, int lid, uint32_t q, int cycle) {
    const StepCtl *c = d.ctl;
    const int L = c->layer;
    if (lid > L) return 0;
    int k = executed_in(cts, cycle, c, c->layer_jump0[L], c->njumps);
    k = k < 0 ? 0 : k;
    if (L > 0) {
        const int z = static_cast<int>(q / (static_cast<uint32_t>(d.X) * static_cast<uint32_t>(d.Y)));
        for (int l = lid < 0 ? 0 : lid; l < L; l++) {
            if (z < c->layer_zlo[l] || z > c->layer_zhi[l]) continue;
            const int e = executed_in(cts, c->layer_end[l], c, c->layer_jump0[l], c->layer_jump0[l + 1]);
            k += e < 0 ? 0 : e;
        }
    }
    return k;
}
// The growth increments: grow[k] below the plateau index ksat, the plateau value from there on.
// tab is the table or a copy of its first ksat entries in LDS (the capture pass stages one, as its
// replays walk the table one dependent step at a time).
struct Growth {
    const double *tab;
    int ksat;
    double gsat;
    __device__ __forceinline__ double inc(int k) const { return k < ksat ? tab[k] : gsat; }
};
__device__ __forceinline__ Growth growth_of(const DomainView &d) { return Growth{d.grow, d.ksat, d.grow[d.ngrow - 1]}; }
// The same with the table's first ksat entries copied to the workgroup's LDS buffer when they fit
// (the replays walk the table one dependent step at a time). Every thread of the workgroup calls it.
constexpr int kGrowStage = 4096;  // table entries staged in LDS (32 KB)
__device__ __forceinline__ Growth staged_growth(const DomainView &d, double *lds) {
    Growth g = growth_of(d);
    if (d.ksat <= kGrowStage) {
        for (int k = threadIdx.x; k < d.ksat; k += kBlock) lds[k] = d.grow[k];
        g.tab = lds;
    }
    __syncthreads();
    return g;
}

// dl after the growth of the steps with index k0+1 .. k1, one float rounding per step exactly as
// the reference stores it.
__device__ float replay_growth(const Growth &g, float dl, int k0, int k1) {
    int k = k0;
    for (; k < k1 && k + 1 < g.ksat; k++) dl = grow_step(dl, g.inc(k + 1));
    if (k < k1) dl = advance_const(dl, k1 - k, g.gsat);
    return dl;
}
// Steps ahead (1 .. kAlarmSlots - 1) at which dl, now at growth index k, first reaches thr; the
// horizon when it does not get there before.
__device__ uint32_t steps_until(const Growth &g, float dl, int k, float thr) {
    const int horizon = static_cast<int>(kAlarmSlots) - 1;
    int n = 0;
    while (n < horizon && k + n + 1 < g.ksat) {
        n++;
        dl = grow_step(dl, g.inc(k + n));
        if (dl >= thr) return static_cast<uint32_t>(n);
    }
    if (n == horizon) return static_cast<uint32_t>(horizon);
    return static_cast<uint32_t>(n + reach_const(dl, g.gsat, thr, horizon - n));
}

// ---- The calendar: cell q is processed at the step of serial s when bit q of slot s mod
// kAlarmSlots is set. ----
__device__ __forceinline__ size_t bitmap_words(const DomainView &d) { return static_cast<size_t>((d.ncell + 31) / 32); }
__device__ __forceinline__ void set_alarm(const DomainView &d, uint32_t serial, uint32_t q) {
    atomicOr(&d.alarms[(serial % kAlarmSlots) * alarm_stride(d.ncell) + (q >> 5)], 1u << (q & 31));
}

struct StepCounts {
    unsigned long long nsol, entered, neval, nmask;
};
// Workgroup b's value of a counter row for this step (thread 0 holds the block sum).
__device__ __forceinline__ void put_part(const DomainView &d, int row, uint32_t b, unsigned long long v,
                                         unsigned long long *sh) {
    const unsigned long long t = block_sum(v, sh);
    if (threadIdx.x == 0) d.part[static_cast<size_t>(row) * d.nparts + b] = t;
}
// Four rows at once (one pair of barriers); sh holds 4 * kBlock / 64 entries.
__device__ __forceinline__ void put_parts4(const DomainView &d, const int (&rows)[4], unsigned long long (&v)[4],
                                           uint32_t b, unsigned long long *sh) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[i] += __shfl_down(v[i], o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < 4; i++) sh[i * (kBlock / 64) + w] = v[i];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            unsigned long long t = 0;
            for (int q = 0; q < kBlock / 64; q++) t += sh[i * (kBlock / 64) + q];
            d.part[static_cast<size_t>(rows[i]) * d.nparts + b] = t;
        }
}

// Growth index that an active cell's dl includes, at a step of growth index k. A cell not yet
// looked at as an active cell (mask not yet known, or built by its capture: kFresh) holds its
// initial dl: it grows from the step it is first looked at on, so its base is k - 1 (k while it
// waits for its liquidus time or its layer: k then holds only earlier layers' steps, 0 in a
// single-layer run); otherwise the base is in win.
__device__ __forceinline__ int growth_base(const CellWords &w, int k, bool waiting) {
    if ((w.lm & kLmUnknown) || (w.flags & kFresh)) return waiting ? k : k - 1;
    return static_cast<int>(w.win);
}
// Cell r of known mask is past its liquidus time and has no liquid neighbour left: at its visit
// in this step's capture the reference grows it and turns it solid (src/CAupdate.cpp:722-725).
__device__ __forceinline__ void solidify(const DomainView &d, const Growth &g, uint32_t r, const CellWords &w, int k) {
    Cell &c = d.cells[r];
    c.dl = replay_growth(g, w.dl, growth_base(w, k, false), k);
    c.win = static_cast<uint32_t>(k);
    c.flags = w.flags & ~kFresh;
    c.ct = kSolid;
}

// Every body below works on one domain with workgroups b = 0 .. nb-1; the plain launches use the
// grid x dimension, the ensemble launches add one grid row per domain slot. A finished domain
// (device-side termination) keeps its state: its bodies return at once.

// Nucleation (src/CAupdate.cpp:67-200). A nucleus leaves the liquid masks of its neighbours; a
// neighbour whose mask this empties and that is past its liquidus time finds no liquid neighbour in
// this step's capture and turns solid here.
// Most steps have no site due: next_nuc[p], the least time of the sites not yet decided, lets every
// workgroup return at once (its counter rows zero). A step that examines sites sets the next step's
// value from what is left (next_nuc[1-p], reset to INT32_MAX by the capture pass of the step before);
// for a step that does not, the capture pass passes the value on. The step's control words are
// advanced by the capture pass too, so a step may leave this kernel out when no site can be due
// (the host knows the sites' times).
constexpr int kPartNucRows[4] = {kPartNucNn, kPartNucAct, kPartNucSol, kPartNucEntered};
__device__ __forceinline__ void nucleate_body(const DomainView &d, int p, int layer, uint32_t b, uint32_t nb) {
    __shared__ unsigned long long sh[4 * (kBlock / 64)];
    __shared__ int smin;
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    const int cycle = ctl->cycle[p];
    const int gate = ctl->next_nuc[p];
    if (cycle < gate) {
        if (threadIdx.x < 4) d.part[static_cast<size_t>(kPartNucRows[threadIdx.x]) * d.nparts + b] = 0;
        return;
    }
    if (threadIdx.x == 0) smin = 0x7FFFFFFF;
    __syncthreads();
    int tmin = 0x7FFFFFFF;  // least time of this thread's sites left undecided
    const Grid gr = grid_of(d);
    unsigned long long nn = 0, act = 0;
    StepCounts n{0, 0, 0, 0};
    for (int i = b * kBlock + threadIdx.x; i < d.nnuclei; i += nb * kBlock) {
        const int t = d.ntime[i];
        if (cycle < t) {
            tmin = t < tmin ? t : tmin;
            continue;
        }
        const uint32_t q = static_cast<uint32_t>(d.nloc[i]);
        const uint32_t tail = reinterpret_cast<const uint32_t *>(d.cells + q)[7];  // lid | ct << 8 | flags
        if (static_cast<int8_t>(tail & 0xFFu) > layer) {
            tmin = t < tmin ? t : tmin;
            continue;
        }
        // Past its time and in this layer, the site is decided now: a cell that is not liquid
        // never turns liquid again, so the site is not looked at after this step.
        d.ntime[i] = 0x7FFFFFFF;
        if (((tail >> 8) & 0xFFu) != kLiquid) continue;
        const Coord c = decode(q, gr);
        if (c.x > 0 && c.x < d.X - 1 && c.y > 0 && c.y < d.Y - 1) nn++;
        Cell &cq = d.cells[q];
        cq.dl = 0.01f;
        cq.thr = 0.0f;
        cq.lm = kLmUnknown;
        cq.flags = 0;
        cq.ct = kActive | kSeedBit;
        set_centre(d, q, (c.x + d.xoff) + 0.5, (c.y + d.yoff) + 0.5, c.z + 0.5);
        const uint32_t valid = valid_neighbours(c, gr.X, gr.Y, gr.nzA);
        // all updates in flight together, then the results
        uint32_t old[26];
#pragma unroll
        for (int l = 0; l < 26; l++)
            old[l] = ((valid >> l) & 1u) ? atomicAnd(mask_word(d, q + gr.off[l]), ~(1u << kOpposite.v[l])) : 0u;
        uint32_t emptied = 0;
#pragma unroll
        for (int l = 0; l < 26; l++)
            emptied |= static_cast<uint32_t>(((valid >> l) & 1u) &&
                                             (old[l] & (kMaskUnknown | kMaskNow)) == (1u << kOpposite.v[l])) << l;
        for (uint32_t m = emptied; m; m &= m - 1) {
            const uint32_t r = q + gr.offset(__builtin_ctz(m));
            const CellWords w = load_cell(d.cells + r);
            if ((w.ct & kTypeMask) != kActive || !(cycle > w.cts && w.lid <= layer)) continue;
            if (!(w.flags & kCounted)) n.entered++;
            solidify(d, growth_of(d), r, w, growth_index(d, w.cts, w.lid, r, cycle));
            d.cells[r].flags = (w.flags & ~kFresh) | kCounted;
            n.nsol++;
        }
        atomicMin(&ctl->elo, q);
        atomicMax(&ctl->ehi, q);
        act++;
        const uint32_t slot = atomicAdd(&ctl->count[p], 1u);
        if (slot < d.list_cap)
            d.list[p][slot] = q;
        else
            atomicOr(&ctl->error, 1);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int v = __shfl_xor(tmin, o, 64);
        tmin = v < tmin ? v : tmin;
    }
    if ((threadIdx.x & 63) == 0) atomicMin(&smin, tmin);
    unsigned long long v4[4] = {nn, act, n.nsol, n.entered};
    put_parts4(d, {kPartNucNn, kPartNucAct, kPartNucSol, kPartNucEntered}, v4, b, sh);
    if (threadIdx.x == 0 && smin != 0x7FFFFFFF) atomicMin(&ctl->next_nuc[1 - p], smin);
}

// Critical lengths of the liquid neighbours of frontier cell j (half-diagonal dlj): claims on
// those reached. thr becomes the least length of the others (the caller stores it), hint the
// record's note of which neighbour that is and of the next-least length (threshold_hint), and
// the claimed mask is returned. The 26 lengths are one aligned 128-byte row: the cell's own
// (stored by its capture) or, for a cell-centred octahedron, its orientation's row of the seed
// table. They are compared all at once and the claims issued after, so a lane's work does not
// depend on its liquid count.
__device__ uint32_t evaluate_claims(const DomainView &d, const Grid &gr, uint32_t j, float dlj, uint32_t liquid,
                                    int32_t gid, uint8_t ct, float &thr, uint16_t &hint) {
    const float *row = d.octa[j].cdl;
    if (ct & kSeedBit) {
        int o = orientation_of(gid);
        if (o < 0) {
            atomicOr(&d.ctl->error, 2);
            o = 0;
        }
        row = d.seed_cdl + static_cast<size_t>(kCdlStride) * o;
    }
    float cdl[26];
    const float4 *r4 = reinterpret_cast<const float4 *>(row);
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const float4 v = r4[i];
        cdl[4 * i] = v.x;
        cdl[4 * i + 1] = v.y;
        cdl[4 * i + 2] = v.z;
        cdl[4 * i + 3] = v.w;
    }
    const float2 v2 = reinterpret_cast<const float2 *>(row)[12];
    cdl[24] = v2.x;
    cdl[25] = v2.y;
    uint32_t claimed = 0;
    float next = __builtin_inff(), next2 = __builtin_inff();
    int at = -1;
#pragma unroll
    for (int l = 0; l < 26; l++) {
        const bool liq = (liquid >> l) & 1u;
        const bool reach = dlj >= cdl[l];
        claimed |= static_cast<uint32_t>(liq && reach) << l;
        const bool open = liq && !reach;
        const bool lower = open && cdl[l] < next;
        next2 = lower ? next : ((open && cdl[l] < next2) ? cdl[l] : next2);
        at = lower ? l : at;
        next = lower ? cdl[l] : next;
    }
    for (uint32_t m = claimed; m; m &= m - 1) atomicMin(claim_word(d, j + gr.offset(__builtin_ctz(m))), j);
    thr = next;
    hint = threshold_hint(at, next, next2);
    return claimed;
}

// Pass A over the cells due this step (the step's calendar slot plus this step's nuclei), one per
// lane: the liquidus-time gate, the growth since the cell was last looked at, and the claims.
// The liquid neighbours come from the cell's mask (lm), kept current by the cells that leave the
// liquid state; a new active cell builds it once from the 26 neighbour types. A cell past its
// liquidus time with no liquid neighbour turns solid here (no capture of this step can give it
// one).
// thr is a lower bound on the critical length of every neighbour that is still liquid: the
// smallest critical length over the liquid neighbours the cell did not claim when it last
// evaluated them. Liquid sets only shrink and a cell's octahedron centre is
// fixed, so no capture is possible before the half-diagonal reaches thr: the cell is scheduled for
// that step (or for the end of its wait for the liquidus time) and not looked at before, unless a
// neighbour's capture or nucleation empties its mask. A new active cell starts at thr = 0.
// Cells with claims are appended to the claim lists inside the workgroup's own slice of the step's
// list (an LDS counter, no global atomics), and the slice's count goes to slice_claims[b].

__device__ __forceinline__ void claim_body(const DomainView &d, int p, int layer, uint32_t b, uint32_t nb) {
    __shared__ unsigned long long sh[4 * (kBlock / 64)];
    __shared__ uint32_t nclaim;
    __shared__ double gtab[kGrowStage];
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    if (threadIdx.x == 0) nclaim = 0;
    const Growth g = staged_growth(d, gtab);
    const int cycle = ctl->cycle[p];
    const uint32_t serial = ctl->serial[p];
    if (b == 0 && threadIdx.x == 0) {
        // the next step's control words (read only by the kernels after this one); this step's
        // nucleation gate, passed on when no site was due, then reset for the step two on
        ctl->cycle[1 - p] = cycle + 1;
        ctl->serial[1 - p] = serial + 1;
        ctl->ncand = 0;
        ctl->rebuild_seq = ctl->rebuild_seq + 1;
        const int gate = ctl->next_nuc[p];
        if (cycle < gate) ctl->next_nuc[1 - p] = gate;
        ctl->next_nuc[p] = 0x7FFFFFFF;
    }
    uint32_t begin, end;
    frontier_slice(ctl->count[p], d.list_cap, b, nb, begin, end);
    const uint32_t *list = d.list[p];
    const Grid gr = grid_of(d);
    StepCounts n{0, 0, 0, 0};
    // software pipeline: the list entry two rounds ahead and the record one round ahead are loaded
    // while the current round is processed
    uint32_t jn = 0, jnn = 0;
    CellWords wn{};
    {
        const uint32_t i1 = begin + threadIdx.x, i2 = i1 + kBlock;
        if (i1 < end) jn = list[i1];
        if (i2 < end) jnn = list[i2];
        if (i1 < end) wn = load_cell(d.cells + jn);
    }
    for (uint32_t i0 = begin; i0 < end; i0 += kBlock) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t j = 0, claimed = 0;
        const uint32_t jc = jn;
        const CellWords wc = wn;
        {
            const uint32_t i3 = i + 2 * kBlock;
            jn = jnn;
            if (i + kBlock < end) wn = load_cell(d.cells + jn);
            if (i3 < end) jnn = list[i3];
        }
        if (i < end) {
            j = jc;
            const CellWords w = wc;
            if ((w.ct & kTypeMask) == kActive) {
                const bool waiting = !(cycle > w.cts && w.lid <= layer);
                const int k = growth_index(d, w.cts, w.lid, j, cycle);
                uint32_t liquid;
                const int kdl = growth_base(w, k, waiting);
                // the record's new words, written back as one sector at the end
                unsigned long long lm = w.lm;
                float dl = w.dl, thr = w.thr;
                uint32_t win = w.win;
                uint8_t ct = w.ct;
                uint16_t flags = w.flags;
                if (w.lm & kLmUnknown) {
                    const Coord c = decode(j, gr);
                    const uint32_t valid = valid_neighbours(c, gr.X, gr.Y, gr.nzA);
                    liquid = 0;
#pragma unroll
                    for (int l = 0; l < 26; l++) {
                        const uint32_t q = ((valid >> l) & 1u) ? j + gr.off[l] : j;
                        liquid |= static_cast<uint32_t>(d.cells[q].ct == kLiquid) << l;
                    }
                    lm = liquid;
                    n.nmask++;
                }
                else {
                    liquid = static_cast<uint32_t>(w.lm & kLmNow);
                }
                if (waiting) {
                    win = static_cast<uint32_t>(kdl);
                    flags = w.flags & ~kFresh;
                    const int ahead = w.cts + 1 - cycle;
                    const uint32_t wake = ahead < 1 ? 1u
                                                    : (ahead < static_cast<int>(kAlarmSlots) - 1 ? static_cast<uint32_t>(ahead)
                                                                                               : kAlarmSlots - 1);
                    set_alarm(d, serial + wake, j);
                }
                else {
                    if (!(w.flags & kCounted)) n.entered++;
                    flags = (w.flags & ~kFresh) | kCounted;
                    dl = replay_growth(g, w.dl, kdl, k);
                    win = static_cast<uint32_t>(k);
                    if (!liquid) {
                        ct = kSolid;
                        n.nsol++;
                    }
                    else {
                        if (dl >= thr) {
                            const uint32_t at = (w.flags >> 2) & 31u;
                            const float bound = hint_bound(thr, w.flags);
                            if (at && !((liquid >> (at - 1)) & 1u) && dl < bound) {
                                // the neighbour that set thr is gone and no other is reached
                                thr = bound;
                                flags &= ~kHintMask;
                            }
                            else {
                                uint16_t hint = 0;
                                claimed = evaluate_claims(d, gr, j, dl, liquid, w.gid, w.ct, thr, hint);
                                flags = (flags & ~kHintMask) | hint;
                                n.neval++;
                            }
                        }
                        // No neighbour left to reach (all claimed now): the captures of this step
                        // empty the mask and the settle pass decides the cell's fate.
                        if (thr < __builtin_inff()) set_alarm(d, serial + steps_until(g, dl, k, thr), j);
                    }
                }
                // Only this lane writes cell j in this pass (claims go to liquid cells' claim words,
                // mask builds read liquid cells' types), so the whole sector is written back.
                uint4 *r4 = reinterpret_cast<uint4 *>(d.cells + j);
                r4[0] = make_uint4(static_cast<uint32_t>(lm), static_cast<uint32_t>(lm >> 32), __float_as_uint(dl),
                                   __float_as_uint(thr));
                r4[1] = make_uint4(static_cast<uint32_t>(w.cts), static_cast<uint32_t>(w.gid), win,
                                   static_cast<uint32_t>(static_cast<uint8_t>(w.lid)) | (static_cast<uint32_t>(ct) << 8) |
                                       (static_cast<uint32_t>(flags) << 16));
            }
        }
        const uint32_t slot = begin + wave_append(claimed != 0, &nclaim);
        if (claimed) {
            d.claim_cell[slot] = j;
            d.claim_bits[slot] = claimed;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) d.slice_claims[b] = nclaim;
    unsigned long long v4[4] = {n.nsol, n.entered, n.neval, n.nmask};
    put_parts4(d, {kPartClaimSol, kPartClaimEntered, kPartClaimEval, kPartClaimMask}, v4, b, sh);
}

// Pass B1 over the cells with claims, one workgroup per pass-A slice: a claim whose word holds the
// claimant's own index is won (the lowest index that reached the cell). The slice's claimants that
// won anything are compacted in place, with the claims they won, and slice_claims[b] becomes their
// count. (Winners of a round go to positions below the round's end, whose entries every lane has
// read before the barrier; later rounds' entries lie above.) Each claimant needs only its claim
// words here: one round of every slice is resident at once, a short chain of dependent loads.
__device__ __forceinline__ void win_body(const DomainView &d, int p, uint32_t b, uint32_t nb) {
    __shared__ uint32_t nwin;
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    uint32_t beg, e;
    frontier_slice(ctl->count[p], d.list_cap, b, nb, beg, e);
    const uint32_t n = d.slice_claims[b];
    const Grid gr = grid_of(d);
    if (threadIdx.x == 0) nwin = 0;
    // two claimants per thread and round, their loads in flight together (a slice holds a little
    // more than one workgroup's worth)
    constexpr uint32_t kPer = 2;
    for (uint32_t k0 = 0; k0 < n; k0 += kPer * kBlock) {
        uint32_t j[kPer], cb[kPer], won[kPer];
#pragma unroll
        for (uint32_t e = 0; e < kPer; e++) {
            const uint32_t k = k0 + e * kBlock + threadIdx.x;
            j[e] = 0;
            cb[e] = 0;
            won[e] = 0;
            if (k < n) {
                j[e] = d.claim_cell[beg + k];
                cb[e] = d.claim_bits[beg + k];
            }
        }
#pragma unroll
        for (uint32_t e = 0; e < kPer; e++) {
            uint32_t cw[26];
#pragma unroll
            for (int l = 0; l < 26; l++) cw[l] = ((cb[e] >> l) & 1u) ? *claim_word(d, j[e] + gr.off[l]) : kClaimNone;
#pragma unroll
            for (int l = 0; l < 26; l++) won[e] |= static_cast<uint32_t>(cw[l] == j[e]) << l;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t e = 0; e < kPer; e++) {
            const uint32_t slot = wave_append(won[e] != 0, &nwin);
            if (won[e]) {
                d.claim_cell[beg + slot] = j[e];
                d.claim_bits[beg + slot] = won[e];
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) d.slice_claims[b] = nwin;
}

// Pass B2 over the winners: workgroup a takes the winners of the kApplyGroup slices a*kApplyGroup ..
// (a few dozen each). The captures of a round of winners are numbered by a block scan and built one
// per lane (a winner mostly has one; one lane per capture keeps every lane's chain of dependent
// accesses to one capture): the winner's octahedron gives the captured cell's
// (src/CAupdate.cpp:353-563, 717), the new cell is due at the next step, and the captured cell
// leaves the liquid masks of its neighbours. Only masks that still hold it are updated: the masks
// are read first, and a cell's bit is cleared by its capture alone, so the read decides exactly.
// A neighbour whose mask this empties is a solidification candidate of this step, staged per
// workgroup in LDS and written to the workgroup's own region of the candidate buffer (no shared
// counter); a wave that finds the stage full appends to the overflow list in the other list buffer.
__device__ __forceinline__ void apply_body(const DomainView &d, int p, uint32_t a, uint32_t nbc) {
    __shared__ unsigned long long sh[kBlock / 64];
    __shared__ uint32_t wave_tot[kBlock / 64];
    __shared__ uint32_t pre[kBlock + 1];
    __shared__ uint32_t won_cell[kBlock], won_bits[kBlock];
    __shared__ uint32_t cstage[kCandStage];
    __shared__ uint32_t ncstage, stage_end;
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    // stage capacity in use (kCandStage; less only when a caller narrows it)
    const uint32_t cap = d.cand_cap < kCandStage ? d.cand_cap : kCandStage;
    if (threadIdx.x == 0) {
        ncstage = 0;
        stage_end = cap;
    }
    uint32_t *overflow = d.list[1 - p];
    uint32_t beg[kApplyGroup], upto[kApplyGroup], n = 0;
#pragma unroll
    for (uint32_t g = 0; g < kApplyGroup; g++) {
        const uint32_t b = a * kApplyGroup + g;
        uint32_t e = 0, cnt = 0;
        beg[g] = 0;
        if (b < nbc) {
            frontier_slice(ctl->count[p], d.list_cap, b, nbc, beg[g], e);
            cnt = d.slice_claims[b];
        }
        n += cnt;
        upto[g] = n;
    }
    const uint32_t serial = ctl->serial[p];
    const Grid gr = grid_of(d);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long ncap = 0;
    uint32_t qlo = 0xFFFFFFFFu, qhi = 0;
    for (uint32_t k0 = 0; k0 < n; k0 += kBlock) {
        // This round's winners in LDS, and the block-wide exclusive prefix of their capture counts.
        const uint32_t k = k0 + threadIdx.x;
        uint32_t j = 0, won = 0;
        if (k < n) {
            uint32_t i = beg[0] + k;
#pragma unroll
            for (uint32_t g = 1; g < kApplyGroup; g++)
                if (k >= upto[g - 1]) i = beg[g] + (k - upto[g - 1]);
            j = d.claim_cell[i];
            won = d.claim_bits[i];
        }
        const uint32_t cnt = __builtin_popcount(won);
        uint32_t wtot;
        const uint32_t incl = wave_inclusive_scan(cnt, wtot);
        if (lane == 63) wave_tot[wave] = wtot;
        won_cell[threadIdx.x] = j;
        won_bits[threadIdx.x] = won;
        __syncthreads();
        uint32_t base = 0, total = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            if (w < wave) base += wave_tot[w];
            total += wave_tot[w];
        }
        pre[threadIdx.x] = base + incl - cnt;
        if (threadIdx.x == 0) pre[kBlock] = total;
        __syncthreads();
        // One capture per lane: item t belongs to the winner s with pre[s] <= t < pre[s + 1].
        for (uint32_t t0 = 0; t0 < total; t0 += kBlock) {
            const uint32_t t = t0 + threadIdx.x;
            uint32_t q = 0, emptied = 0;
            if (t < total) {
                uint32_t lo = 0, hi = kBlock;  // pre[lo] <= t < pre[hi]
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (pre[mid] <= t)
                        lo = mid;
                    else
                        hi = mid;
                }
                const uint32_t jc = won_cell[lo];
                uint32_t bits = won_bits[lo];
                for (uint32_t r = t - pre[lo]; r; r--) bits &= bits - 1;
                const int l = __builtin_ctz(bits);
                q = jc + gr.offset(l);
                // the winner's grain and centre (float, as the capture uses it), loaded with the
                // neighbours' words below
                const int h = d.cells[jc].gid;
                const Octa &e = d.octa[jc];
                const float wcx = static_cast<float>(e.x), wcy = static_cast<float>(e.y), wcz = static_cast<float>(e.z);
                const Coord c = decode(jc, gr);
                const Coord cqc{c.x + nb_dx(l), c.y + nb_dy(l), c.z + nb_dz(l)};
                const uint32_t valid = valid_neighbours(cqc, gr.X, gr.Y, gr.nzA);
                // the grain's orientation, loaded with the neighbours' words below rather than
                // after the mask atomics
                int o = orientation_of(h);
                if (o < 0) {
                    atomicOr(&ctl->error, 2);
                    o = 0;
                }
                float gv[9];
#pragma unroll
                for (int m = 0; m < 9; m++) gv[m] = d.guv[9 * o + m];
                // the neighbours' words: the low word is the mask (or liquid marker), the high word
                // of a liquid neighbour its claim word
                unsigned long long lw[26];
#pragma unroll
                for (int kk = 0; kk < 26; kk++)
                    lw[kk] = ((valid >> kk) & 1u) ? *reinterpret_cast<const unsigned long long *>(&d.cells[q + gr.off[kk]].lm)
                                                  : static_cast<unsigned long long>(kMaskUnknown);
                // The new cell's mask: its liquid neighbours at its first visit are those liquid now
                // and not claimed in this step (every claimed cell is captured in this step, and the
                // claims are final). Until its store below, the new cell's own words read as a
                // claimed liquid cell, which no other capture counts.
                uint32_t liquid = 0, mw[26];
#pragma unroll
                for (int kk = 0; kk < 26; kk++) {
                    liquid |= static_cast<uint32_t>(is_unclaimed_liquid(lw[kk])) << kk;
                    mw[kk] = static_cast<uint32_t>(lw[kk]);
                }
                // updates only where the mask still holds q, all in flight together; their results
                // are used after the octahedron is built, which hides their round trip
                uint32_t need = 0;
#pragma unroll
                for (int kk = 0; kk < 26; kk++) {
                    const uint32_t bit = 1u << kOpposite.v[kk];
                    need |= static_cast<uint32_t>((mw[kk] & (kMaskUnknown | bit)) == bit) << kk;
                }
#pragma unroll
                for (int kk = 0; kk < 26; kk++)
                    if ((need >> kk) & 1u) mw[kk] = atomicAnd(mask_word(d, q + gr.off[kk]), ~(1u << kOpposite.v[kk]));
                // a neighbour at or below the claimant's index still had q liquid at its visit
                // in the reference's index order: its keep word gets this step's tag (every writer
                // of a step writes the same value)
#pragma unroll
                for (int kk = 0; kk < 26; kk++)
                    if (((need >> kk) & 1u) && q + gr.off[kk] <= jc) keep_word(d, q + gr.off[kk])[0] = serial + 1;
                const double xp = (c.x + d.xoff + nb_dx(l)) + 0.5;
                const double yp = (c.y + d.yoff + nb_dy(l)) + 0.5;
                const double zp = (c.z + nb_dz(l)) + 0.5;
                const Octahedron oc = new_octahedron(gv, wcx, wcy, wcz, xp, yp, zp);
                Cell &cq = d.cells[q];
                cq.gid = h;
                cq.dl = static_cast<float>(oc.dl);
                cq.thr = 0.0f;
                cq.flags = kFresh;
                cq.ct = kActive;
                set_alarm(d, serial + 1, q);
                ncap++;
                qlo = q < qlo ? q : qlo;
                qhi = q > qhi ? q : qhi;
#pragma unroll
                for (int kk = 0; kk < 26; kk++) {
                    const uint32_t bit = 1u << kOpposite.v[kk];
                    emptied |= static_cast<uint32_t>(((need >> kk) & 1u) && (mw[kk] & (kMaskUnknown | kMaskNow)) == bit) << kk;
                }
                cq.lm = liquid;
                // The new octahedron's record: its 26 critical lengths, computed once here (its
                // evaluations read them back), and its centre; whole 16-byte pieces of one line.
                {
                    const Faces f = faces_of(gv);
                    float4 *o4 = reinterpret_cast<float4 *>(&d.octa[q]);
#pragma unroll
                    for (int i = 0; i < 6; i++)
                        o4[i] = make_float4(critical_length(f, xp, yp, zp, oc.cx, oc.cy, oc.cz, 4 * i),
                                            critical_length(f, xp, yp, zp, oc.cx, oc.cy, oc.cz, 4 * i + 1),
                                            critical_length(f, xp, yp, zp, oc.cx, oc.cy, oc.cz, 4 * i + 2),
                                            critical_length(f, xp, yp, zp, oc.cx, oc.cy, oc.cz, 4 * i + 3));
                    const uint2 xb = __builtin_bit_cast(uint2, oc.cx), yb = __builtin_bit_cast(uint2, oc.cy),
                                zb = __builtin_bit_cast(uint2, oc.cz);
                    uint4 *u4 = reinterpret_cast<uint4 *>(&d.octa[q]);
                    u4[6] = make_uint4(__float_as_uint(critical_length(f, xp, yp, zp, oc.cx, oc.cy, oc.cz, 24)),
                                       __float_as_uint(critical_length(f, xp, yp, zp, oc.cx, oc.cy, oc.cz, 25)), xb.x, xb.y);
                    u4[7] = make_uint4(yb.x, yb.y, zb.x, zb.y);
                }
            }
            // Candidates: staged in LDS (one LDS atomic per wave); a wave that finds the stage
            // full appends to the global list directly. Waves reserve disjoint stage ranges in
            // the order of their atomics, so every wave after the first that does not fit does
            // not fit either, and every range below that wave's base is filled: stage_end (the
            // least base of a wave that did not fit) bounds the entries copied out below.
            const uint32_t ce = __builtin_popcount(emptied);
            uint32_t ctot;
            const uint32_t cincl = wave_inclusive_scan(ce, ctot);
            if (ctot) {
                uint32_t cbase = 0;
                if (lane == 63) cbase = atomicAdd(&ncstage, ctot);
                cbase = __shfl(cbase, 63, 64);
                const bool staged = cbase + ctot <= cap;
                if (!staged) {
                    if (lane == 63) {
                        atomicMin(&stage_end, cbase);
                        cbase = atomicAdd(&ctl->ncand, ctot);
                    }
                    cbase = __shfl(cbase, 63, 64);
                }
                uint32_t pos = cbase + cincl - ce;
                for (uint32_t r = emptied; r; r &= r - 1, pos++) {
                    const uint32_t cell = q + gr.offset(__builtin_ctz(r));
                    if (staged)
                        cstage[pos] = cell;
                    else if (pos < d.list_cap)
                        overflow[pos] = cell;
                }
            }
        }
        __syncthreads();
    }
    // The staged candidates go to this workgroup's region (the barrier orders thread 0's
    // initialisation before these reads when the workgroup had no winners).
    __syncthreads();
    const uint32_t staged = ncstage < stage_end ? ncstage : stage_end;
    for (uint32_t i = threadIdx.x; i < staged; i += kBlock) d.cand[a * kCandStage + i] = cstage[i];
    if (threadIdx.x == 0) d.cand_count[a] = staged;
    put_part(d, kPartCap, a, ncap, sh);
    // New cells widen the range of active cells (least and greatest index of this workgroup).
    __shared__ uint32_t lohi[2 * (kBlock / 64)];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t a2 = __shfl_down(qlo, o, 64), b2 = __shfl_down(qhi, o, 64);
        qlo = a2 < qlo ? a2 : qlo;
        qhi = b2 > qhi ? b2 : qhi;
    }
    if (lane == 0) {
        lohi[2 * wave] = qlo;
        lohi[2 * wave + 1] = qhi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; w++) {
            qlo = lohi[2 * w] < qlo ? lohi[2 * w] : qlo;
            qhi = lohi[2 * w + 1] > qhi ? lohi[2 * w + 1] : qhi;
        }
        // the range rarely grows: most workgroups need no atomic
        if (qlo <= qhi) {
            if (qlo < ctl->elo) atomicMin(&ctl->elo, qlo);
            if (qhi > ctl->ehi) atomicMax(&ctl->ehi, qhi);
        }
    }
}

// The settle pass has one workgroup per apply workgroup: workgroup b takes the candidates of
// region b and a share of the overflow list.
__host__ __device__ constexpr uint32_t settle_blocks(uint32_t claim_blocks) { return apply_blocks(claim_blocks); }

// Pass C over the candidates: a cell whose every claim-time liquid neighbour was captured in this
// step by a lower-indexed cell has no liquid neighbour left when the reference visits it in index
// order, so it grows and turns solid now (src/CAupdate.cpp:333-334, 722-725); a capture by itself
// or by a later cell leaves it active, with no liquid neighbour: it turns solid at the next step. A
// cell not past its liquidus time is not visited this step.
// A candidate's mask was emptied by this step's captures, each of which tagged the candidate's
// keep word with this step when its claimant's index is >= the candidate's: the tag decides.
__device__ __forceinline__ void settle_body(const DomainView &d, int p, int layer, uint32_t b, uint32_t nb) {
    __shared__ unsigned long long sh[kBlock / 64];
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    const int cycle = ctl->cycle[p];
    const uint32_t serial = ctl->serial[p];
    const uint32_t nown = d.cand_count[b];
    const uint32_t nover = ctl->ncand < d.list_cap ? ctl->ncand : d.list_cap;
    const uint32_t *own = d.cand + b * kCandStage;
    const uint32_t *overflow = d.list[1 - p];
    unsigned long long nsol = 0;
    // the growth table is read from global memory: a candidate has mostly been growing on the
    // plateau, where the replay needs no table entry, so staging it would cost more than it saves
    const Growth g = growth_of(d);
    auto settle = [&](uint32_t r) {
        const CellWords w = load_cell(d.cells + r);
        if ((w.ct & kTypeMask) != kActive || !(cycle > w.cts && w.lid <= layer)) return;
        if (static_cast<uint32_t>(w.lm >> 32) == serial + 1) {
            set_alarm(d, serial + 1, r);
        }
        else {
            solidify(d, g, r, w, growth_index(d, w.cts, w.lid, r, cycle));
            nsol++;
        }
    };
    for (uint32_t i = threadIdx.x; i < nown; i += kBlock) settle(own[i]);
    for (uint32_t i = b * kBlock + threadIdx.x; i < nover; i += nb * kBlock) settle(overflow[i]);
    put_part(d, kPartSettleSol, b, nsol, sh);
}


// The list of the next step is every set bit of its calendar slot, in ascending cell index, over
// the chunks of the range of active cells; the slot is cleared as it is read. One pass: workgroup
// b of nbw - 1 takes units b, b + nbw - 1, ... of kUnitChunks chunks, counts a unit's bits, and
// finds the unit's offset by decoupled look-back over the units before it (each unit publishes its
// count, then its inclusive prefix, as one tagged 64-bit word); then it writes the unit's cells.
// A unit's count never waits on anything, and a unit waits only for lower units, of workgroups
// dispatched before it or of earlier rounds of workgroups that are resident with it (the plain
// launch has at most 513 workgroups, and at 3 workgroups per CU all are resident); each wait is
// bounded all the same (error 8). The last workgroup books the step meanwhile. (The range is a few
// thousand chunks at 800^3: several hundred units, one to two per CU; measured there, units of
// 4, 8, 16 chunks with up to 512, 256, 128 workgroups: 23.9, 29.5, 65.9 us.)
constexpr uint32_t kChunkWords = 8 * kBlock;  // thread t of a chunk holds its words 8t .. 8t + 7
constexpr uint32_t kChunkCells = 32 * kChunkWords;
constexpr uint32_t kUnitChunks = 4;
constexpr uint32_t kRebuildStage = 8192;  // list entries of a unit staged in LDS (32 KB)
constexpr unsigned long long kFlagCount = 1ull << 32;   // value = the chunk's own count
constexpr unsigned long long kFlagPrefix = 2ull << 32;  // value = count of the chunk and all before it
constexpr unsigned long long kFlagState = 3ull << 32;
constexpr unsigned long long kFlagTag = ~((1ull << 34) - 1);
constexpr uint32_t kLookBackSpins = 1u << 20;  // bound on the waits of one look-back (error 8 past it)

__device__ __forceinline__ bool chunk_range(const DomainView &d, uint32_t &c0, uint32_t &c1) {
    const uint32_t lo = d.ctl->elo;
    uint32_t hi = d.ctl->ehi;
    if (lo > hi) return false;
    if (hi >= static_cast<uint64_t>(d.ncell)) hi = static_cast<uint32_t>(d.ncell - 1);
    c0 = lo / kChunkCells;
    c1 = hi / kChunkCells;
    return true;
}
__device__ __forceinline__ uint32_t *next_slot(const DomainView &d, int p) {
    return d.alarms + ((d.ctl->serial[p] + 1) % kAlarmSlots) * alarm_stride(d.ncell);
}
// Flags of one rebuild carry its tag (the rebuild sequence number, advanced by the kernels that
// precede a rebuild: the capture pass, rewake and layer setup; tags cycle through 2^30 - 1 nonzero values), so the
// flags of earlier rebuilds need no clearing.
__device__ __forceinline__ unsigned long long flag_tag(uint32_t seq) {
    return static_cast<unsigned long long>(seq % ((1u << 30) - 1) + 1) << 34;
}
__device__ __forceinline__ void publish(unsigned long long *f, unsigned long long v) {
    __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long peek(unsigned long long *f) {
    return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Cells before unit u (u > 0), by one wave: windows of 64 predecessors, nearest first, summed up
// to the nearest one holding a prefix; unit 0 always holds one.
__device__ uint32_t look_back(const DomainView &d, uint32_t u, unsigned long long tag) {
    const int lane = threadIdx.x & 63;
    uint32_t excl = 0, spins = 0;
    long long j0 = static_cast<long long>(u) - 1;
    for (;;) {
        const long long j = j0 - lane;
        const unsigned long long f = j >= 0 ? peek(d.chunk_flag + j) : (tag | kFlagPrefix);
        const bool valid = (f & kFlagTag) == tag && (f & kFlagState) != 0;
        const unsigned long long pm = __ballot((f & kFlagTag) == tag && (f & kFlagState) == kFlagPrefix);
        const unsigned long long vm = __ballot(valid);
        const int np = pm ? __builtin_ctzll(pm) : 63;
        const unsigned long long need = np >= 63 ? ~0ull : ((2ull << np) - 1);
        if ((vm & need) != need) {
            if (++spins > kLookBackSpins) {
                if (lane == 0) atomicOr(&d.ctl->error, 8);
                return excl;
            }
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        uint32_t v = lane <= np ? static_cast<uint32_t>(f) : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        excl += v;
        if (pm) return excl;
        j0 -= 64;
    }
}

// The step's counters: the per-workgroup slots of each row summed over the grids that wrote them
// (nb capture workgroups, the apply and settle workgroups that go with them, nnb nucleation
// workgroups).
constexpr uint32_t kMaxPartsPerThread = 8;  // nparts <= 8 * kBlock
struct StepSums {
    unsigned long long v[kPartRows];
};
__device__ __forceinline__ StepSums sum_parts(const DomainView &d, uint32_t nb, uint32_t nnb) {
    __shared__ unsigned long long red[kPartRows][kBlock / 64];
    const uint32_t na = apply_blocks(nb);
    StepSums s;
#pragma unroll
    for (int r = 0; r < kPartRows; r++) {
        const uint32_t n = (r >= kPartNucSol && r <= kPartNucNn) ? nnb
                           : (r == kPartCap ? na : (r == kPartSettleSol ? settle_blocks(nb) : nb));
        const unsigned long long *row = d.part + static_cast<size_t>(r) * d.nparts;
        unsigned long long v = 0;
#pragma unroll
        for (uint32_t k = 0; k < kMaxPartsPerThread; k++) {
            const uint32_t i = threadIdx.x + k * kBlock;
            if (i < n) v += row[i];
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if ((threadIdx.x & 63) == 0) red[r][threadIdx.x >> 6] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPartRows; r++) {
        unsigned long long v = 0;
        for (int w = 0; w < kBlock / 64; w++) v += red[r][w];
        s.v[r] = v;
    }
    return s;
}

// Closing a rebuild (the list's length is set by the last unit, or here when there is none): with
// account the step's counters are booked and a domain with autostop ends at a check step (cycle a
// multiple of 1000) with no liquid left: liquid cells only ever leave the liquid state by capture
// or nucleation, so the reference's superheated + undercooled count (src/CAupdate.cpp:738-783) is
// the initial liquid count minus both. A rebuild between steps passes account 0.
__device__ __forceinline__ void set_count(const DomainView &d, int p, uint32_t total) {
    d.ctl->count[1 - p] = total < d.list_cap ? total : d.list_cap;
    if (total > d.list_cap) atomicOr(&d.ctl->error, 1);
}
__device__ __forceinline__ void close_rebuild(const DomainView &d, int p, int account, uint32_t nb, uint32_t nnb,
                                              uint32_t nunits) {
    StepCtl *ctl = d.ctl;
    StepSums sums;
    if (account) sums = sum_parts(d, nb, nnb);
    if (threadIdx.x != 0) return;
    if (!nunits) set_count(d, p, 0);
    if (!account) return;
    // Counters are read once and written once (every access is a dependent round trip for this
    // single thread).
    const unsigned long long *v = sums.v;
    const unsigned long long sol = v[kPartClaimSol] + v[kPartNucSol] + v[kPartSettleSol];
    StepStats st = *d.stats;
    const unsigned long long nonwaiting = ctl->nonwaiting + v[kPartClaimEntered] + v[kPartNucEntered];
    const unsigned long long active = ctl->nactive + v[kPartNucAct];
    st.solidified += sol;
    st.evaluated += v[kPartClaimEval];
    st.masks += v[kPartClaimMask];
    st.captures += v[kPartCap];
    st.nucleated += v[kPartNucNn];
    st.activated += v[kPartNucAct];
    // The reference's numSteer of this step: active cells past their liquidus time at the start of
    // its capture pass.
    st.frontier += nonwaiting;
    st.waiting += active - nonwaiting;
    *d.stats = st;
    ctl->nonwaiting = nonwaiting - sol;
    ctl->nactive = active + v[kPartCap] - sol;
    const int cycle = ctl->cycle[p];
    ctl->last_cycle = cycle;
    // A domain whose kernels raised an error flag stops here: its later steps would work on a
    // truncated list. The host learns of it at its next check (plain path) or poll (ensemble).
    const int32_t err = ctl->error;
    if (err) {
        ctl->done = 1;
        ctl->final_cycle = cycle;
        if (d.status_err) *d.status_err = err;
        if (d.status) *d.status = cycle;
        return;
    }
    if (d.autostop && cycle % 1000 == 0 &&
        d.liquid0 - static_cast<long long>(st.captures) - static_cast<long long>(st.activated) == 0) {
        ctl->done = 1;
        ctl->final_cycle = cycle;
        if (d.status) *d.status = cycle;
    }
}

__device__ __forceinline__ void rebuild_body(const DomainView &d, int p, int account, uint32_t nb, uint32_t nnb,
                                             uint32_t b, uint32_t nbw) {
    __shared__ uint32_t sh_excl;
    __shared__ uint32_t wave_tot[kUnitChunks][kBlock / 64];
    __shared__ uint32_t stage[kRebuildStage];
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    uint32_t c0 = 1, c1 = 0;
    const bool any = chunk_range(d, c0, c1);
    const uint32_t nunits = any ? (c1 - c0) / kUnitChunks + 1 : 0;
    if (b == nbw - 1) {
        close_rebuild(d, p, account, nb, nnb, nunits);
        return;
    }
    const unsigned long long tag = flag_tag(ctl->rebuild_seq);
    const uint32_t nwords = static_cast<uint32_t>(bitmap_words(d));
    uint32_t *bits = next_slot(d, p);
    uint32_t *next = d.list[1 - p];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t t = b; t < nunits; t += nbw - 1) {
        // the unit's words, all loads in flight together
        uint32_t w[kUnitChunks][8];
#pragma unroll
        for (uint32_t k = 0; k < kUnitChunks; k++) {
            const uint32_t c = c0 + t * kUnitChunks + k;
            const uint32_t w0 = c * kChunkWords + 8 * threadIdx.x;
            if (c <= c1 && w0 + 8 <= nwords) {
                const uint4 a = reinterpret_cast<const uint4 *>(bits + w0)[0];
                const uint4 b = reinterpret_cast<const uint4 *>(bits + w0)[1];
                w[k][0] = a.x, w[k][1] = a.y, w[k][2] = a.z, w[k][3] = a.w;
                w[k][4] = b.x, w[k][5] = b.y, w[k][6] = b.z, w[k][7] = b.w;
            }
            else {
#pragma unroll
                for (int i = 0; i < 8; i++) w[k][i] = (c <= c1 && w0 + i < nwords) ? bits[w0 + i] : 0u;
            }
        }
        uint32_t cnt[kUnitChunks], incl[kUnitChunks];
#pragma unroll
        for (uint32_t k = 0; k < kUnitChunks; k++) {
            cnt[k] = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) cnt[k] += __builtin_popcount(w[k][i]);
            uint32_t wtot;
            incl[k] = wave_inclusive_scan(cnt[k], wtot);
            if (lane == 63) wave_tot[k][wave] = wtot;
        }
        __syncthreads();
        // offset of this thread's cells within the unit (chunk by chunk, thread by thread)
        uint32_t off[kUnitChunks], agg = 0;
#pragma unroll
        for (uint32_t k = 0; k < kUnitChunks; k++) {
            uint32_t before = 0, tot = 0;
            for (int q = 0; q < kBlock / 64; q++) {
                if (q < wave) before += wave_tot[k][q];
                tot += wave_tot[k][q];
            }
            off[k] = agg + before + incl[k] - cnt[k];
            agg += tot;
        }
        if (wave == 0) {
            uint32_t excl = 0;
            if (t == 0) {
                if (lane == 0) publish(d.chunk_flag + t, tag | kFlagPrefix | agg);
            }
            else {
                if (lane == 0) publish(d.chunk_flag + t, tag | kFlagCount | agg);
                excl = look_back(d, t, tag);
                if (lane == 0) publish(d.chunk_flag + t, tag | kFlagPrefix | (excl + agg));
            }
            if (lane == 0) {
                sh_excl = excl;
                // the last unit knows the list's length
                if (t == nunits - 1) set_count(d, p, excl + agg);
            }
        }
        // the slot is cleared where it held bits; the cells go to LDS in list order when they fit
        const bool staged = agg <= kRebuildStage;
#pragma unroll
        for (uint32_t k = 0; k < kUnitChunks; k++) {
            const uint32_t c = c0 + t * kUnitChunks + k;
            const uint32_t w0 = c * kChunkWords + 8 * threadIdx.x;
            if (c <= c1 && w0 + 8 <= nwords) {
                if (w[k][0] | w[k][1] | w[k][2] | w[k][3]) reinterpret_cast<uint4 *>(bits + w0)[0] = make_uint4(0, 0, 0, 0);
                if (w[k][4] | w[k][5] | w[k][6] | w[k][7]) reinterpret_cast<uint4 *>(bits + w0)[1] = make_uint4(0, 0, 0, 0);
            }
            else {
#pragma unroll
                for (int i = 0; i < 8; i++)
                    if (w[k][i]) bits[w0 + i] = 0u;
            }
            if (staged) {
                uint32_t o = off[k];
#pragma unroll
                for (int i = 0; i < 8; i++)
                    for (uint32_t m = w[k][i]; m; m &= m - 1) stage[o++] = (w0 + i) * 32u + __builtin_ctz(m);
            }
        }
        __syncthreads();
        const uint32_t excl = sh_excl;
        if (staged) {
            for (uint32_t i = threadIdx.x; i < agg; i += kBlock)
                if (excl + i < d.list_cap) next[excl + i] = stage[i];
        }
        else {
#pragma unroll
            for (uint32_t k = 0; k < kUnitChunks; k++) {
                const uint32_t w0 = (c0 + t * kUnitChunks + k) * kChunkWords + 8 * threadIdx.x;
                uint32_t pos = excl + off[k];
#pragma unroll
                for (int i = 0; i < 8; i++)
                    for (uint32_t m = w[k][i]; m; m &= m - 1, pos++)
                        if (pos < d.list_cap) next[pos] = (w0 + i) * 32u + __builtin_ctz(m);
            }
        }
        __syncthreads();  // the stage and the scan words are reused by the next unit
    }
}

// Every active cell of the window due at the next step (after a cycle jump, whose skipped steps the
// calendar did not count). Active cells above the window (substrate cells next to a later layer's
// melt pool) are never stepped, as the reference's pass 1 never visits them.
__device__ __forceinline__ void rewake_body(const DomainView &d, int p_next, uint32_t b, uint32_t nb) {
    StepCtl *ctl = d.ctl;
    if (b == 0 && threadIdx.x == 0) ctl->rebuild_seq = ctl->rebuild_seq + 1;
    const uint32_t lo = ctl->elo, hi = ctl->ehi;
    if (lo > hi) return;
    const unsigned long long end = static_cast<unsigned long long>(d.X) * d.Y * d.nzA;
    const uint32_t serial = ctl->serial[p_next];
    for (unsigned long long q = lo + static_cast<unsigned long long>(b) * kBlock + threadIdx.x; q <= hi && q < end;
         q += static_cast<unsigned long long>(nb) * kBlock)
        if ((d.cells[q].ct & kTypeMask) == kActive) set_alarm(d, serial, static_cast<uint32_t>(q));
}

// The next layer's window (src/CAinitialize.cpp:2772-2832): the plane of the lowest active cell and
// the top plane of the next layer's cells. out[0] = least z (atomicMin), out[1] = greatest z.
__global__ __launch_bounds__(kBlock) void k_layer_bounds(DomainView d, int next_layer, int *out) {
    int zmin = 0x7FFFFFFF, zmax = -1;
    const Grid gr = grid_of(d);
    for (long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; q < d.ncell;
         q += static_cast<long long>(gridDim.x) * kBlock) {
        const uint32_t tail = reinterpret_cast<const uint32_t *>(d.cells + q)[7];
        const int z = static_cast<int>(gr.xy.div(static_cast<uint32_t>(q)));
        if (((tail >> 8) & kTypeMask) == kActive) zmin = z < zmin ? z : zmin;
        if (static_cast<int8_t>(tail & 0xFFu) == next_layer) zmax = z > zmax ? z : zmax;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int a = __shfl_xor(zmin, o, 64), b = __shfl_xor(zmax, o, 64);
        zmin = a < zmin ? a : zmin;
        zmax = b > zmax ? b : zmax;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&out[0], zmin);
        atomicMax(&out[1], zmax);
    }
}

// LayerSetup (src/CAinitialize.cpp:2835-2982) on the resident state: every active cell of the new
// window [zlo, nzA) gets a cell-centred octahedron (half-diagonal 0.01, centre at the cell centre,
// critical lengths from its orientation) and an unknown liquid mask (its neighbours changed with the
// layer), is counted afresh, and is due at the first step of the layer (the step of parity p_next).
// The calendar was cleared before; the list is rebuilt after. ctl->nactive and the range of active
// cells (ctl->elo, ctl->ehi, reset by the host) are accumulated here.
__global__ __launch_bounds__(kBlock) void k_layer_setup(DomainView d, int p_next) {
    __shared__ unsigned long long sh[kBlock / 64];
    StepCtl *ctl = d.ctl;
    const uint32_t serial = ctl->serial[p_next];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->rebuild_seq = ctl->rebuild_seq + 1;
    const Grid gr = grid_of(d);
    const long long begin = static_cast<long long>(d.X) * d.Y * d.zlo;
    const long long end = static_cast<long long>(d.X) * d.Y * d.nzA;
    unsigned long long n = 0;
    uint32_t qlo = 0xFFFFFFFFu, qhi = 0;
    for (long long q = begin + static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; q < end;
         q += static_cast<long long>(gridDim.x) * kBlock) {
        Cell &c = d.cells[q];
        if ((c.ct & kTypeMask) != kActive) continue;
        const Coord p = decode(static_cast<uint32_t>(q), gr);
        c.lm = kLmUnknown;
        c.dl = 0.01f;
        c.thr = 0.0f;
        c.win = 0;
        c.flags = 0;
        c.ct = kActive | kSeedBit;
        set_centre(d, q, (p.x + d.xoff) + 0.5, (p.y + d.yoff) + 0.5, p.z + 0.5);
        set_alarm(d, serial, static_cast<uint32_t>(q));
        n++;
        qlo = static_cast<uint32_t>(q) < qlo ? static_cast<uint32_t>(q) : qlo;
        qhi = static_cast<uint32_t>(q) > qhi ? static_cast<uint32_t>(q) : qhi;
    }
    const unsigned long long t = block_sum(n, sh);
    if (threadIdx.x == 0 && t) atomicAdd(&ctl->nactive, t);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t a = __shfl_xor(qlo, o, 64), b = __shfl_xor(qhi, o, 64);
        qlo = a < qlo ? a : qlo;
        qhi = b > qhi ? b : qhi;
    }
    if ((threadIdx.x & 63) == 0 && qlo <= qhi) {
        atomicMin(&ctl->elo, qlo);
        atomicMax(&ctl->ehi, qhi);
    }
}

// Plain launches: one domain, passed by value.
__global__ __launch_bounds__(kBlock) void k_nucleate(DomainView d, int p, int layer) {
    nucleate_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_claim(DomainView d, int p, int layer) {
    claim_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_win(DomainView d, int p) {
    win_body(d, p, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_apply(DomainView d, int p, int nbc) {
    apply_body(d, p, blockIdx.x, static_cast<uint32_t>(nbc));
}
__global__ __launch_bounds__(kBlock) void k_settle(DomainView d, int p, int layer) {
    settle_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_rebuild(DomainView d, int p, int account, int nb, int nnb) {
    rebuild_body(d, p, account, static_cast<uint32_t>(nb), static_cast<uint32_t>(nnb), blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_rewake(DomainView d, int p_next) {
    rewake_body(d, p_next, blockIdx.x, gridDim.x);
}

// Ensemble launches: grid row y serves domain slot y of a device-resident table; an empty slot has
// no control block.
__global__ __launch_bounds__(kBlock) void k_nucleate_n(const DomainView *v, int p, int layer) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) nucleate_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_claim_n(const DomainView *v, int p, int layer) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) claim_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_win_n(const DomainView *v, int p) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) win_body(d, p, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_apply_n(const DomainView *v, int p, int nbc) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) apply_body(d, p, blockIdx.x, static_cast<uint32_t>(nbc));
}
__global__ __launch_bounds__(kBlock) void k_settle_n(const DomainView *v, int p, int layer) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) settle_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_rebuild_n(const DomainView *v, int p, int nb, int nnb) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) rebuild_body(d, p, 1, static_cast<uint32_t>(nb), static_cast<uint32_t>(nnb), blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kBlock) void k_check(DomainView d, int cycle, int layer, CheckResult *out) {
    __shared__ unsigned long long sh[kBlock / 64];
    __shared__ unsigned long long shmin[kBlock / 64];
    unsigned long long c[5] = {0, 0, 0, 0, 0};
    unsigned long long next = ~0ull, amin = ~0ull, amax = 0;
    // the active window [ZBound_Low, ZBound_High] (src/CAupdate.cpp:792-808 reduces over it)
    const long long active_begin = static_cast<long long>(d.X) * d.Y * d.zlo;
    const long long active_end = static_cast<long long>(d.X) * d.Y * d.nzA;
    for (long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; q < d.ncell;
         q += static_cast<long long>(gridDim.x) * kBlock) {
        const int t = d.cells[q].ct & kTypeMask;
        if (t == kActive && q >= active_begin && q < active_end) {
            amin = static_cast<unsigned long long>(q) < amin ? static_cast<unsigned long long>(q) : amin;
            amax = static_cast<unsigned long long>(q) > amax ? static_cast<unsigned long long>(q) : amax;
        }
        if (d.cells[q].lid == layer) {
            if (t == kLiquid) {
                const int crit = d.cells[q].cts;
                if (crit > cycle)
                    c[0]++;
                else
                    c[1]++;
                const unsigned long long v = static_cast<unsigned long long>(static_cast<long long>(crit));
                if (q >= active_begin && q < active_end && v < next) next = v;
            }
            else if (t == kActive)
                c[2]++;
            else if (t == kSolid)
                c[3]++;
        }
        else if (t == kLiquid) {
            c[4]++;
        }
    }
    for (int k = 0; k < 5; k++) {
        const unsigned long long s = block_sum(c[k], sh);
        if (threadIdx.x == 0 && s) atomicAdd(&out->counts[k], s);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long v = __shfl_down(next, o, 64);
        next = v < next ? v : next;
    }
    if ((threadIdx.x & 63) == 0) shmin[threadIdx.x >> 6] = next;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = shmin[0];
        for (int q = 1; q < kBlock / 64; q++) m = shmin[q] < m ? shmin[q] : m;
        if (m != ~0ull) atomicMin(&out->next_cts, m);
    }
    // range of the active cells, one atomic per wave
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long a = __shfl_down(amin, o, 64), b = __shfl_down(amax, o, 64);
        amin = a < amin ? a : amin;
        amax = b > amax ? b : amax;
    }
    if ((threadIdx.x & 63) == 0 && amin <= amax) {
        atomicMin(&out->amin, amin);
        atomicMax(&out->amax, amax);
    }
}

__global__ __launch_bounds__(kBlock) void k_seed_init(DomainView d, const uint32_t *cells, uint32_t n) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t q = cells[i];
        const Coord c = decode(q, d.X, d.Y);
        d.cells[q].dl = 0.01f;
        set_centre(d, q, (c.x + d.xoff) + 0.5, (c.y + d.yoff) + 0.5, c.z + 0.5);
    }
}

// Grain statistics of a region of the whole-domain views (k, i, j), the layout the reference's vtk
// files and its analysis executable use (analysis/src/GAprint.cpp:17-108): the interior of this
// domain for z = 1 .. nz-2, zero on the bottom and top planes. Per cell: its GrainID is flagged in
// gflag (GrainID - gmin; plain stores: every writer writes 1), a negative one counts as nucleated,
// and a melted cell counts in the bin of its orientation (LDS bins flushed once per workgroup), in
// a second set of bins when it lies in the region's top plane.
__global__ __launch_bounds__(kBlock) void k_grain_stats(DomainView d, const uint8_t *melted, GrainRegion r, int gmin,
                                                        uint32_t ngid, uint32_t *gflag, unsigned long long *counts,
                                                        unsigned int *obins, unsigned int *otop) {
    __shared__ unsigned int bins[kNumOrientations];
    __shared__ unsigned long long sh[kBlock / 64];
    for (int o = threadIdx.x; o < kNumOrientations; o += kBlock) bins[o] = 0;
    __syncthreads();
    const long long ni = r.i1 - r.i0 + 1, nj = r.j1 - r.j0 + 1, nk = r.k1 - r.k0 + 1;
    const long long total = ni * nj * nk;
    unsigned long long nmelt = 0, nnuc = 0, nbad = 0;
    for (long long c = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; c < total;
         c += static_cast<long long>(gridDim.x) * kBlock) {
        const int j = r.j0 + static_cast<int>(c % nj);
        const int i = r.i0 + static_cast<int>((c / nj) % ni);
        const int k = r.k0 + static_cast<int>(c / (nj * ni));
        int g = 0;
        bool m = false;
        if (k >= 1 && k <= d.nz - 2) {
            const long long q = (static_cast<long long>(k) * d.X + (i + 1)) * d.Y + (j + 1);
            g = d.cells[q].gid;
            m = melted[q] != 0;
        }
        const long long f = static_cast<long long>(g) - gmin;
        if (f < 0 || f >= ngid) {
            nbad++;
            continue;
        }
        gflag[f] = 1u;
        nnuc += g < 0;
        if (m) {
            nmelt++;
            const int o = orientation_of(g) < 0 ? 0 : orientation_of(g);
            atomicAdd(&bins[o], 1u);
            if (k == r.k1) atomicAdd(&otop[o], 1u);
        }
    }
    const unsigned long long tm = block_sum(nmelt, sh);
    const unsigned long long tn = block_sum(nnuc, sh);
    const unsigned long long tb = block_sum(nbad, sh);
    if (threadIdx.x == 0) {
        if (tm) atomicAdd(&counts[0], tm);
        if (tn) atomicAdd(&counts[1], tn);
        if (tb) atomicAdd(&counts[2], tb);
    }
    __syncthreads();
    for (int o = threadIdx.x; o < kNumOrientations; o += kBlock)
        if (bins[o]) atomicAdd(&obins[o], bins[o]);
}

// The initial state of a directional-solidification domain built on the device
// (TempInit_DirSolidification, SubstrateInit_ConstrainedGrowth, GrainInit and NucleiInit of
// src/CAinitialize.cpp:1025-1055, 1460-1529, 1834-2248): every cell is a wall (the ghost ring and
// the bottom and top planes) or liquid, with its plane's CritTimeStep, LayerID 0 and the Melted
// flag; then the substrate's active cells (GrainIDs 1, 2, ... in list order, cell-centred octahedra
// of half-diagonal 0.01) and the nucleation sites (GrainIDs -1, -2, ... in list order) from the
// host's draws. No cell of such a domain is solid, so GrainInit's last pass (liquid cells without a
// non-solid, non-wall neighbour turn solid) changes nothing when the interior is at least two
// cells wide in x and y.
__global__ __launch_bounds__(kBlock) void k_init_directional(DomainView d, const int32_t *plane_cts, int nx, int ny,
                                                             uint8_t *melted) {
    const Grid gr = grid_of(d);
    for (long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; q < d.ncell;
         q += static_cast<long long>(gridDim.x) * kBlock) {
        const Coord c = decode(static_cast<uint32_t>(q), gr);
        const int gx = c.x + d.xoff, gy = c.y + d.yoff;
        const bool wall = gx == -1 || gx == nx || gy == -1 || gy == ny || c.z == 0 || c.z == d.nz - 1;
        const uint8_t ct = wall ? kWall : kLiquid;
        uint4 *r = reinterpret_cast<uint4 *>(d.cells + q);
        r[0] = make_uint4(wall ? kMaskUnknown : 0xFFFFFFFFu, wall ? 0u : kClaimNone, 0u, 0u);
        r[1] = make_uint4(static_cast<uint32_t>(plane_cts[c.z]), 0u, 0u, static_cast<uint32_t>(ct) << 8);
        set_centre(d, q, 0.0, 0.0, 0.0);
        if (melted) melted[q] = wall ? 0 : 1;
    }
}
__global__ __launch_bounds__(kBlock) void k_init_substrate(DomainView d, const uint32_t *cells, uint32_t n) {
    const Grid gr = grid_of(d);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t q = cells[i];
        const Coord c = decode(q, gr);
        Cell &r = d.cells[q];
        r.lm = kLmUnknown;
        r.dl = 0.01f;
        r.gid = static_cast<int32_t>(i + 1);
        r.ct = kActive | kSeedBit;
        set_centre(d, q, (c.x + d.xoff) + 0.5, (c.y + d.yoff) + 0.5, c.z + 0.5);
    }
}
__global__ __launch_bounds__(kBlock) void k_init_sites(DomainView d, const int32_t *sites, uint32_t n) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        d.cells[sites[i]].gid = -static_cast<int32_t>(i + 1);
}

// Initial cell records from the host's compact arrays: no octahedron, no claim, liquid mask not
// yet known; the centres start at zero.
__global__ __launch_bounds__(kBlock) void k_pack(DomainView d, const uint8_t *ct, const int8_t *lid,
                                                 const int32_t *gid, const int32_t *cts) {
    for (long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; q < d.ncell;
         q += static_cast<long long>(gridDim.x) * kBlock) {
        uint4 *c = reinterpret_cast<uint4 *>(d.cells + q);
        const bool liquid = (ct[q] & kTypeMask) == kLiquid;
        // a liquid cell: the liquid marker and no claim; any other cell: mask not yet known
        c[0] = make_uint4(liquid ? 0xFFFFFFFFu : kMaskUnknown, liquid ? kClaimNone : 0u, 0u, 0u);
        c[1] = make_uint4(static_cast<uint32_t>(cts[q]), static_cast<uint32_t>(gid[q]), 0u,
                          static_cast<uint32_t>(static_cast<uint8_t>(lid[q])) | (static_cast<uint32_t>(ct[q]) << 8));
        set_centre(d, q, 0.0, 0.0, 0.0);
    }
}

// One per-cell field in the caller's layout (int32 per cell unless noted).
__global__ __launch_bounds__(kBlock) void k_extract(DomainView d, int field, void *out) {
    for (long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; q < d.ncell;
         q += static_cast<long long>(gridDim.x) * kBlock) {
        const Cell &c = d.cells[q];
        switch (field) {
        case kFieldCellType: static_cast<int32_t *>(out)[q] = c.ct & kTypeMask; break;
        case kFieldCellTypeRaw: static_cast<uint8_t *>(out)[q] = c.ct; break;  // uint8
        case kFieldGrainId: static_cast<int32_t *>(out)[q] = c.gid; break;
        case kFieldDiagonal: {  // float, with the growth of the steps since the cell was last looked at
            float dl = c.dl;
            if ((c.ct & kTypeMask) == kActive && !(c.lm & kLmUnknown) && !(c.flags & kFresh))
                dl = replay_growth(growth_of(d), dl, static_cast<int>(c.win), growth_index(d, c.cts, c.lid, static_cast<uint32_t>(q), d.ctl->last_cycle));
            static_cast<float *>(out)[q] = dl;
            break;
        }
        case kFieldCritTimeStep: static_cast<int32_t *>(out)[q] = c.cts; break;
        case kFieldLayerId: static_cast<int32_t *>(out)[q] = c.lid; break;
        case kFieldCentre: {  // float x, y, z
            const Octa &e = d.octa[q];
            float *o = static_cast<float *>(out) + 3 * q;
            o[0] = static_cast<float>(e.x);
            o[1] = static_cast<float>(e.y);
            o[2] = static_cast<float>(e.z);
            break;
        }
        case kFieldCentreF64: {  // double x, y, z, as computed at capture
            const Octa &e = d.octa[q];
            double *o = static_cast<double *>(out) + 3 * q;
            o[0] = e.x;
            o[1] = e.y;
            o[2] = e.z;
            break;
        }
        default: break;
        }
    }
}

// in: n rows of (orientation, cxold, cyold, czold, xp, yp, zp); out: n rows of
// (cx, cy, cz, float(dl), 26 critical lengths recomputed from the new centre).
__global__ __launch_bounds__(kBlock) void k_geometry_probe(const float *guv, const double *in, double *out, int n) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    const double *a = in + 7 * r;
    const float *g = guv + 9 * static_cast<int>(a[0]);
    const Octahedron o = new_octahedron(g, a[1], a[2], a[3], a[4], a[5], a[6]);
    double *b = out + 30 * r;
    b[0] = o.cx;
    b[1] = o.cy;
    b[2] = o.cz;
    b[3] = static_cast<float>(o.dl);
    const Faces f = faces_of(g);
    for (int l = 0; l < 26; l++) b[4 + l] = critical_length(f, a[4], a[5], a[6], o.cx, o.cy, o.cz, l);
}

// Unique GrainIDs of a region: the flags k_grain_stats set, counted into *out.
__global__ __launch_bounds__(kBlock) void k_count_flags(const uint32_t *gflag, uint32_t n, unsigned long long *out) {
    __shared__ unsigned long long sh[kBlock / 64];
    unsigned long long c = 0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) c += gflag[i] != 0;
    const unsigned long long t = block_sum(c, sh);
    if (threadIdx.x == 0 && t) atomicAdd(out, t);
}

// Fill of n 32-bit words (a kernel rather than a memset node, so that a caller may capture it).
__global__ __launch_bounds__(kBlock) void k_fill(uint32_t *p, size_t n, uint32_t v) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kBlock)
        p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kBlock)
        dst[i] = src[i];
}

inline int grid_for(long long n, int cap) {
    long long b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    return static_cast<int>(b < cap ? b : cap);
}

}  // namespace

int nucleation_blocks(const DomainView &d) { return grid_for(d.nnuclei, 1024); }
void launch_nucleation(const DomainView &d, int parity, int layer, hipStream_t s) {
    k_nucleate<<<nucleation_blocks(d), kBlock, 0, s>>>(d, parity, layer);
}
void launch_claim(const DomainView &d, int parity, int layer, int blocks, hipStream_t s) {
    k_claim<<<blocks, kBlock, 0, s>>>(d, parity, layer);
}
// Workgroups of the list rebuild: at most one per unit of the whole domain, 512, and as many as
// the device holds at once (each takes every (n-1)-th unit: a unit of a second round waits for a
// unit of a higher workgroup of the first, which must then be resident), plus the one that books
// the step.
int rebuild_blocks(const DomainView &d, int blocks) {
    const uint32_t units = (d.nchunks + kUnitChunks - 1) / kUnitChunks;
    uint32_t cap = static_cast<uint32_t>(blocks < 512 ? blocks : 512);
    if (d.rebuild_cap && d.rebuild_cap < cap) cap = d.rebuild_cap;
    return static_cast<int>(units < cap ? units : cap) + 1;
}
uint32_t rebuild_resident(int cus) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rebuild, kBlock, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return static_cast<uint32_t>(per_cu) * static_cast<uint32_t>(cus);
}
uint32_t rebuild_units(long long ncell) {
    const unsigned long long words = (static_cast<unsigned long long>(ncell) + 31) / 32;
    const unsigned long long chunks = (words + kChunkWords - 1) / kChunkWords;
    return static_cast<uint32_t>((chunks + kUnitChunks - 1) / kUnitChunks);
}
void launch_apply(const DomainView &d, int parity, int layer, int blocks, bool nucleated, hipStream_t s) {
    const int cb = rebuild_blocks(d, blocks);
    k_win<<<blocks, kBlock, 0, s>>>(d, parity);
    k_apply<<<apply_blocks(blocks), kBlock, 0, s>>>(d, parity, blocks);
    k_settle<<<settle_blocks(blocks), kBlock, 0, s>>>(d, parity, layer);
    k_rebuild<<<cb, kBlock, 0, s>>>(d, parity, 1, blocks, nucleated ? grid_for(d.nnuclei, 1024) : 0);
}
void launch_layer_bounds(const DomainView &d, int next_layer, int *out, hipStream_t s) {
    k_layer_bounds<<<grid_for(d.ncell, 4096), kBlock, 0, s>>>(d, next_layer, out);
}
void launch_layer_setup(const DomainView &d, int parity_next, int blocks, hipStream_t s) {
    const int cb = rebuild_blocks(d, blocks);
    const long long n = static_cast<long long>(d.X) * d.Y * (d.nzA - d.zlo);
    k_layer_setup<<<grid_for(n, 4096), kBlock, 0, s>>>(d, parity_next);
    k_rebuild<<<cb, kBlock, 0, s>>>(d, 1 - parity_next, 0, blocks, 0);
}
void launch_rewake(const DomainView &d, int parity_next, int blocks, hipStream_t s) {
    const int cb = rebuild_blocks(d, blocks);
    k_rewake<<<blocks, kBlock, 0, s>>>(d, parity_next);
    k_rebuild<<<cb, kBlock, 0, s>>>(d, 1 - parity_next, 0, blocks, 0);
}
void launch_step_timed(const DomainView &d, int parity, int layer, int blocks, hipStream_t s, hipEvent_t *ev) {
    const int cb = rebuild_blocks(d, blocks);
    int e = 0;
    auto rec = [&]() { (void)hipEventRecord(ev[e++], s); };
    rec();
    k_nucleate<<<grid_for(d.nnuclei, 1024), kBlock, 0, s>>>(d, parity, layer);
    rec();
    rec();
    k_claim<<<blocks, kBlock, 0, s>>>(d, parity, layer);
    rec();
    rec();
    k_win<<<blocks, kBlock, 0, s>>>(d, parity);
    rec();
    rec();
    k_apply<<<apply_blocks(blocks), kBlock, 0, s>>>(d, parity, blocks);
    rec();
    rec();
    k_settle<<<settle_blocks(blocks), kBlock, 0, s>>>(d, parity, layer);
    rec();
    rec();
    k_rebuild<<<cb, kBlock, 0, s>>>(d, parity, 1, blocks, grid_for(d.nnuclei, 1024));
    rec();
}
void launch_step_n(const DomainView *views, int slots, int parity, int layer, int nuc_blocks, int blocks,
                   hipStream_t s) {
    k_nucleate_n<<<dim3(nuc_blocks, slots), kBlock, 0, s>>>(views, parity, layer);
    k_claim_n<<<dim3(blocks, slots), kBlock, 0, s>>>(views, parity, layer);
    k_win_n<<<dim3(blocks, slots), kBlock, 0, s>>>(views, parity);
    k_apply_n<<<dim3(apply_blocks(blocks), slots), kBlock, 0, s>>>(views, parity, blocks);
    k_settle_n<<<dim3(settle_blocks(blocks), slots), kBlock, 0, s>>>(views, parity, layer);
    // one round of workgroups covers every unit of an ensemble domain (at most kEnsembleUnits):
    // each waits only for workgroups of its row dispatched before it; plus the booking workgroup
    k_rebuild_n<<<dim3(kEnsembleUnits + 1, slots), kBlock, 0, s>>>(views, parity, blocks, nuc_blocks);
}
void launch_check(const DomainView &d, int cycle, int layer, CheckResult *out, hipStream_t s) {
    k_check<<<grid_for(d.ncell, 4096), kBlock, 0, s>>>(d, cycle, layer, out);
}
void launch_seed_init(const DomainView &d, const uint32_t *cells, uint32_t n, hipStream_t s) {
    k_seed_init<<<grid_for(n, 1024), kBlock, 0, s>>>(d, cells, n);
}
void launch_pack(const DomainView &d, const uint8_t *ct, const int8_t *lid, const int32_t *gid, const int32_t *cts,
                 hipStream_t s) {
    k_pack<<<grid_for(d.ncell, 1 << 16), kBlock, 0, s>>>(d, ct, lid, gid, cts);
}
void launch_grain_stats(const DomainView &d, const uint8_t *melted, const GrainRegion &r, int gmin, uint32_t ngid,
                        uint32_t *gflag, unsigned long long *counts, unsigned int *obins, unsigned int *otop,
                        hipStream_t s) {
    const long long total = static_cast<long long>(r.i1 - r.i0 + 1) * (r.j1 - r.j0 + 1) * (r.k1 - r.k0 + 1);
    k_grain_stats<<<grid_for(total, 2048), kBlock, 0, s>>>(d, melted, r, gmin, ngid, gflag, counts, obins, otop);
}
void launch_init_directional(const DomainView &d, const int32_t *plane_cts, int nx, int ny, uint8_t *melted,
                             const uint32_t *substrate, uint32_t nsub, const int32_t *sites, uint32_t nsites,
                             hipStream_t s) {
    k_init_directional<<<grid_for(d.ncell, 1 << 16), kBlock, 0, s>>>(d, plane_cts, nx, ny, melted);
    if (nsub) k_init_substrate<<<grid_for(nsub, 4096), kBlock, 0, s>>>(d, substrate, nsub);
    if (nsites) k_init_sites<<<grid_for(nsites, 4096), kBlock, 0, s>>>(d, sites, nsites);
}
void launch_extract(const DomainView &d, int field, void *out, hipStream_t s) {
    k_extract<<<grid_for(d.ncell, 1 << 16), kBlock, 0, s>>>(d, field, out);
}
void launch_count_flags(const uint32_t *gflag, uint32_t n, unsigned long long *out, hipStream_t s) {
    k_count_flags<<<grid_for(n, 1024), kBlock, 0, s>>>(gflag, n, out);
}
void launch_fill(uint32_t *p, size_t n, uint32_t v, hipStream_t s) {
    k_fill<<<grid_for(static_cast<long long>(n), 4096), kBlock, 0, s>>>(p, n, v);
}
void launch_copy(const uint4 *src, uint4 *dst, size_t n, hipStream_t s) {
    k_copy<<<grid_for(static_cast<long long>(n), 1 << 16), kBlock, 0, s>>>(src, dst, n);
}
void launch_geometry_probe(const float *guv, const double *in, double *out, int n, hipStream_t s) {
    k_geometry_probe<<<grid_for(n, 1 << 20), kBlock, 0, s>>>(guv, in, out, n);
}

}  // namespace exaca
