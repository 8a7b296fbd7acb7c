// ca_kernels.hip -- the CA time step as memory-bound CDNA4 kernels over a compacted frontier.
//
// One step (reference: src/runCA.cpp:291-311) is seven launches on one stream:
//   k_nucleate  activates the pre-sampled nuclei whose time has come (src/CAupdate.cpp:67-200)
//   k_claim     for every frontier cell past its liquidus time: growth of its octahedron and a
//               claim on every liquid neighbour it reaches, written as atomicMin of the
//               capturer's index into win[]; a cell with no liquid neighbour turns solid
//   k_apply     the winner of each claimed cell builds the new decentred octahedron and takes the
//               cell out of its neighbours' liquid masks
//   k_settle    a cell whose liquid neighbours were all taken by lower-indexed cells turns solid
//   k_bits_*    the next frontier: the active cells in index order, from a bitmap
// Claim-then-apply reproduces the reference's Serial backend exactly: there pass 2 visits active
// cells in ascending index order and every capture is visible at once, so a liquid cell goes to
// the lowest-indexed cell that reaches it, and cell j still "sees" a liquid neighbour only if no
// lower-indexed cell took it (src/CAupdate.cpp:264-352, 722-725).
//
// Every floating-point expression keeps the reference's operation order and float/double types;
// the library is compiled with -ffp-contract=off and correctly rounded fp32 divide/sqrt.
#include <hip/hip_runtime.h>

#include "ca_kernels.hpp"

namespace exaca {
namespace {

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max

__device__ __forceinline__ int orientation_of(int gid) { return (abs(gid) - 1) % kNumOrientations; }

// Steps executed in (t, cycle]: the reference adds the undercooling change once per executed step
// while a cell is liquid or active past its liquidus time (src/CAupdate.cpp:235-237); cycle jumps
// (src/CAupdate.cpp:812-814) skip steps J+1 .. T-1.
__device__ __forceinline__ int executed_since(int t, int cycle, const StepCtl *ctl, int njumps) {
    int n = cycle - t;
    for (int q = 0; q < njumps; q++) {
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
    int8_t lid;
    uint8_t ct;
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
    w.lid = static_cast<int8_t>(b.w & 0xFFu);
    w.ct = static_cast<uint8_t>((b.w >> 8) & 0xFFu);
    return w;
}

__device__ __forceinline__ void set_active_bit(const DomainView &d, uint32_t q) {
    atomicOr(&d.bits[q >> 5], 1u << (q & 31));
}
__device__ __forceinline__ void clear_active_bit(const DomainView &d, uint32_t q) {
    atomicAnd(&d.bits[q >> 5], ~(1u << (q & 31)));
}

// Every body below works on one domain with workgroups b = 0 .. nb-1; the plain launches use the
// grid x dimension, the ensemble launches add one grid row per domain slot. A finished domain
// (device-side termination) keeps its state: its bodies return at once.
__device__ __forceinline__ void nucleate_body(const DomainView &d, int p, int layer, uint32_t b, uint32_t nb) {
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    const int cycle = ctl->cycle[p];
    if (b == 0 && threadIdx.x == 0) {
        ctl->cycle[1 - p] = cycle + 1;
        ctl->count[1 - p] = 0;
        ctl->ncand = 0;
    }
    unsigned long long nn = 0, act = 0;
    for (int i = b * kBlock + threadIdx.x; i < d.nnuclei; i += nb * kBlock) {
        const uint32_t q = static_cast<uint32_t>(d.nloc[i]);
        if (!(cycle >= d.ntime[i] && d.cells[q].ct == kLiquid && d.cells[q].lid <= layer)) continue;
        const Grid gr = grid_of(d);
        const Coord c = decode(q, gr);
        if (c.x > 0 && c.x < d.X - 1 && c.y > 0 && c.y < d.Y - 1) nn++;
        d.cells[q].dl = 0.01f;
        d.cells[q].thr = 0.0f;
        d.centre[q] = Centre{(c.x + d.xoff) + 0.5, (c.y + d.yoff) + 0.5, c.z + 0.5, 0.0};
        d.cells[q].lm = kLmUnknown;
        d.cells[q].ct = kActive | kSeedBit;
        set_active_bit(d, q);
        // q is no longer liquid for any neighbour
        const uint32_t valid = valid_neighbours(c, gr.X, gr.Y, gr.nzA);
#pragma unroll
        for (int l = 0; l < 26; l++)
            if ((valid >> l) & 1u)
                atomicAnd(reinterpret_cast<uint32_t *>(&d.cells[q + gr.off[l]].lm), ~(1u << kOpposite.v[l]));
        atomicMin(&ctl->rlo[p], q);
        atomicMax(&ctl->rhi[p], q);
        act++;
        const uint32_t slot = atomicAdd(&ctl->count[p], 1u);
        if (slot < d.list_cap)
            d.list[p][slot] = q;
        else
            atomicOr(&ctl->error, 1);
    }
    if (nn) atomicAdd(&d.stats->nucleated, nn);
    if (act) atomicAdd(&d.stats->activated, act);
}

// Pass A, one frontier cell per lane: the liquidus-time gate, growth, and the claims.
// The liquid neighbours come from the cell's own mask (lm), kept current by the cells that leave
// the liquid state; a cell whose mask is not yet known (new this step) reads the 26 neighbour
// types once. A cell past its liquidus time with no liquid neighbour turns solid here (no capture
// of this step can give it one).
// thr is a lower bound on the critical length of every neighbour that is still liquid: the
// smallest critical length (or a proven lower bound of it) over the liquid neighbours the cell did
// not claim when it last evaluated them. Liquid sets only shrink and a cell's octahedron centre is
// fixed, so while the half-diagonal stays below thr no neighbour can be captured and the step
// needs neither the octahedron nor any critical length. A new active cell starts at thr = 0.
// Cells with claims are appended to the claim lists inside the workgroup's own frontier slice (an
// LDS counter, no global atomics), and the slice's count goes to slice_claims[b] for pass B.
//
// The common case (mask known, half-diagonal below thr) reads one record and writes the grown
// half-diagonal. The rare cases, a mask to build or critical lengths to evaluate (a few percent of
// the cells but present in nearly every wave), are queued in LDS and run afterwards with every lane
// busy, instead of holding their whole wave at one or two active lanes.

constexpr uint32_t kQueueCap = 4 * kBlock;
constexpr uint32_t kItemNewMask = 0xFFFFFFFFu;  // queue tag: mask not yet known (not a float bit pattern used)

struct ClaimCounts {
    unsigned long long front, nwait, nsol, neval, nmask;
};

// Critical lengths of the liquid neighbours of frontier cell j (half-diagonal dlj): claims on
// those reached; thr becomes the least length (or bound) of the others. Returns the claimed mask.
__device__ uint32_t evaluate_claims(const DomainView &d, const Grid &gr, uint32_t j, float dlj, uint32_t liquid,
                                    int32_t gid, uint8_t ct) {
    int o = orientation_of(gid);
    if (o < 0) {
        atomicOr(&d.ctl->error, 2);
        o = 0;
    }
    uint32_t claimed = 0;
    float next = __builtin_inff();
    if (ct & kSeedBit) {
        const float *tab = d.seed_cdl + 26 * o;
        for (uint32_t m = liquid; m; m &= m - 1) {
            const int l = __builtin_ctz(m);
            const float cdl = tab[l];
            if (dlj >= cdl) {
                atomicMin(&d.cells[j + gr.offset(l)].win, j);
                claimed |= 1u << l;
            }
            else {
                next = fminf(next, cdl);
            }
        }
    }
    else {
        const Coord c = decode(j, gr);
        const Centre e = d.centre[j];
        const Faces f = faces_of(d.guv + 9 * o);
        const double xp = (c.x + d.xoff) + 0.5, yp = (c.y + d.yoff) + 0.5, zp = c.z + 0.5;
        // The critical length of a neighbour at distance |e| from the centre is at least
        // sqrt(reach)|e| (reach from the orientation's deviation from orthonormality, with a
        // margin far above rounding): below that the exact length is not needed, and the bound,
        // rounded down to float, stands in for it in thr.
        const double reach = d.reach[o];
        const double dl2 = static_cast<double>(dlj) * static_cast<double>(dlj);
        for (uint32_t m = liquid; m; m &= m - 1) {
            const int l = __builtin_ctz(m);
            const double ex = xp + nb_dx(l) - e.x, ey = yp + nb_dy(l) - e.y, ez = zp + nb_dz(l) - e.z;
            const double r2 = reach * (ex * ex + ey * ey + ez * ez);
            if (dl2 < r2) {
                next = fminf(next, __double2float_rd(sqrt(r2)));
                continue;
            }
            const float cdl = critical_length(f, xp, yp, zp, e.x, e.y, e.z, l);
            if (dlj >= cdl) {
                atomicMin(&d.cells[j + gr.offset(l)].win, j);
                claimed |= 1u << l;
            }
            else {
                next = fminf(next, cdl);
            }
        }
    }
    d.cells[j].thr = next;
    return claimed;
}

// The half-diagonal after this step's growth (src/CAupdate.cpp:257-261), stored.
__device__ __forceinline__ float grow_cell(const DomainView &d, uint32_t j, int t, int cycle, int njumps, float dl0) {
    int k = executed_since(t, cycle, d.ctl, njumps);
    k = k < 0 ? 0 : (k < d.ngrow - 1 ? k : d.ngrow - 1);
    const float dlj = static_cast<float>(static_cast<double>(dl0) + d.grow[k]);
    d.cells[j].dl = dlj;
    return dlj;
}

__device__ __forceinline__ void solidify(const DomainView &d, uint32_t j) {
    d.cells[j].ct = kSolid;
    clear_active_bit(d, j);
}

__device__ __forceinline__ void claim_body(const DomainView &d, int p, int layer, uint32_t b, uint32_t nb) {
    __shared__ unsigned long long sh[kBlock / 64];
    __shared__ uint32_t nclaim, nq;
    __shared__ uint32_t q_cell[kQueueCap], q_tag[kQueueCap], q_liq[kQueueCap];
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    if (threadIdx.x == 0) {
        nclaim = 0;
        nq = 0;
    }
    __syncthreads();
    const int cycle = ctl->cycle[p];
    const int njumps = ctl->njumps;
    uint32_t begin, end;
    frontier_slice(ctl->count[p], d.list_cap, b, nb, begin, end);
    const uint32_t *list = d.list[p];
    const Grid gr = grid_of(d);
    ClaimCounts n{0, 0, 0, 0, 0};

    // Runs the queued cells, one per lane, and appends their claims.
    auto drain = [&]() {
        const uint32_t total = nq;
        for (uint32_t k0 = 0; k0 < total; k0 += kBlock) {
            const uint32_t k = k0 + threadIdx.x;
            uint32_t j = 0, claimed = 0;
            if (k < total) {
                j = q_cell[k];
                const uint32_t tag = q_tag[k];
                const CellWords cw = load_cell(d.cells + j);
                if (tag == kItemNewMask) {
                    const Coord c = decode(j, gr);
                    const uint32_t valid = valid_neighbours(c, gr.X, gr.Y, gr.nzA);
                    unsigned long long liq = 0;
#pragma unroll
                    for (int l = 0; l < 26; l++) {
                        const uint32_t q = ((valid >> l) & 1u) ? j + gr.off[l] : j;
                        liq |= static_cast<unsigned long long>(d.cells[q].ct == kLiquid) << l;
                    }
                    d.cells[j].lm = liq | (liq << 32);
                    n.nmask++;
                    const uint32_t liquid = static_cast<uint32_t>(liq);
                    if (!(cycle > cw.cts && cw.lid <= layer)) {
                        n.nwait++;
                    }
                    else {
                        n.front++;
                        const float dlj = grow_cell(d, j, cw.cts, cycle, njumps, cw.dl);
                        if (!liquid) {
                            solidify(d, j);
                            n.nsol++;
                        }
                        else if (dlj >= cw.thr) {
                            claimed = evaluate_claims(d, gr, j, dlj, liquid, cw.gid, cw.ct);
                            n.neval++;
                        }
                    }
                }
                else {
                    claimed = evaluate_claims(d, gr, j, __uint_as_float(tag), q_liq[k], cw.gid, cw.ct);
                    n.neval++;
                }
            }
            const uint32_t slot = begin + wave_append(claimed != 0, &nclaim);
            if (claimed) {
                d.claim_cell[slot] = j;
                d.claim_bits[slot] = claimed;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) nq = 0;
        __syncthreads();
    };

    for (uint32_t i0 = begin; i0 < end; i0 += kBlock) {
        const uint32_t i = i0 + threadIdx.x;
        bool defer = false;
        uint32_t j = 0, tag = 0, liquid = 0;
        if (i < end) {
            j = list[i];
            const CellWords cw = load_cell(d.cells + j);
            if (cw.lm & kLmUnknown) {
                defer = true;
                tag = kItemNewMask;
            }
            else {
                liquid = static_cast<uint32_t>(cw.lm & kLmNow);
                if (liquid != static_cast<uint32_t>(cw.lm >> 32))
                    d.cells[j].lm = static_cast<unsigned long long>(liquid) | (static_cast<unsigned long long>(liquid) << 32);
                if (!(cycle > cw.cts && cw.lid <= layer)) {
                    n.nwait++;
                }
                else {
                    n.front++;
                    const float dlj = grow_cell(d, j, cw.cts, cycle, njumps, cw.dl);
                    if (!liquid) {
                        solidify(d, j);
                        n.nsol++;
                    }
                    else if (dlj >= cw.thr) {
                        defer = true;
                        tag = __float_as_uint(dlj);
                    }
                }
            }
        }
        const uint32_t qs = wave_append(defer, &nq);
        if (defer) {
            q_cell[qs] = j;
            q_tag[qs] = tag;
            q_liq[qs] = liquid;
        }
        __syncthreads();
        if (nq > kQueueCap - kBlock) drain();
    }
    __syncthreads();
    if (nq) drain();
    __syncthreads();
    if (threadIdx.x == 0) d.slice_claims[b] = nclaim;
    unsigned long long t = block_sum(n.front, sh);
    if (threadIdx.x == 0 && t) atomicAdd(&d.stats->frontier, t);
    t = block_sum(n.nwait, sh);
    if (threadIdx.x == 0 && t) atomicAdd(&d.stats->waiting, t);
    t = block_sum(n.nsol, sh);
    if (threadIdx.x == 0 && t) atomicAdd(&d.stats->solidified, t);
    t = block_sum(n.neval, sh);
    if (threadIdx.x == 0 && t) atomicAdd(&d.stats->evaluated, t);
    t = block_sum(n.nmask, sh);
    if (threadIdx.x == 0 && t) atomicAdd(&d.stats->masks, t);
}

// Pass B over the cells with claims. Workgroup a takes the claims of the kApplyGroup slices
// a*kApplyGroup .. of pass A's nbc workgroups (a slice holds a few dozen claims: grouping fills the
// waves and keeps the whole grid resident). A claim whose word holds the claimant's own index is won
// (the lowest index that reached the cell). The won captures of a round of claimants are numbered
// by a block scan and then built one per lane: the winner's octahedron gives the captured cell's
// (src/CAupdate.cpp:353-563, 717), and the captured cell leaves the liquid masks of its
// neighbours. Only masks that still hold it are updated: the masks are read first, and a cell's
// bit is cleared by its capture alone, so the read decides exactly. A neighbour whose mask this
// empties is a solidification candidate of this step, appended (per wave) to the candidate list
// held in the other frontier buffer.
constexpr uint32_t kApplyGroup = 4;
__host__ __device__ constexpr uint32_t apply_blocks(uint32_t claim_blocks) {
    return (claim_blocks + kApplyGroup - 1) / kApplyGroup;
}

// The low word of a cell's liquid-neighbour mask: bits 0-25 and the not-yet-known bit.
__device__ __forceinline__ uint32_t *mask_word(const DomainView &d, uint32_t q) {
    return reinterpret_cast<uint32_t *>(&d.cells[q].lm);
}
constexpr uint32_t kMaskUnknown = static_cast<uint32_t>(kLmUnknown);
constexpr uint32_t kMaskNow = static_cast<uint32_t>(kLmNow);

__device__ __forceinline__ void apply_body(const DomainView &d, int p, uint32_t a, uint32_t nbc) {
    __shared__ unsigned long long sh[kBlock / 64];
    __shared__ uint32_t wave_tot[kBlock / 64];
    __shared__ uint32_t pre[kBlock + 1];
    __shared__ uint32_t won_cell[kBlock], won_bits[kBlock];
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
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
    const Grid gr = grid_of(d);
    uint32_t *cand = d.list[1 - p];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long ncap = 0;
    for (uint32_t k0 = 0; k0 < n; k0 += kBlock) {
        // Won claims of this round's claimants, and their block-wide exclusive prefix.
        const uint32_t k = k0 + threadIdx.x;
        uint32_t j = 0, won = 0;
        if (k < n) {
            uint32_t i = beg[0] + k;
#pragma unroll
            for (uint32_t g = 1; g < kApplyGroup; g++)
                if (k >= upto[g - 1]) i = beg[g] + (k - upto[g - 1]);
            j = d.claim_cell[i];
            for (uint32_t m = d.claim_bits[i]; m; m &= m - 1) {
                const int l = __builtin_ctz(m);
                won |= static_cast<uint32_t>(d.cells[j + gr.offset(l)].win == j) << l;
            }
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
        // One capture per lane: item t belongs to the claimant s with pre[s] <= t < pre[s + 1].
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
                const Coord c = decode(jc, gr);
                const int h = d.cells[jc].gid;
                int o = orientation_of(h);
                if (o < 0) {
                    atomicOr(&ctl->error, 2);
                    o = 0;
                }
                const Centre e = d.centre[jc];
                const double xp = (c.x + d.xoff + nb_dx(l)) + 0.5;
                const double yp = (c.y + d.yoff + nb_dy(l)) + 0.5;
                const double zp = (c.z + nb_dz(l)) + 0.5;
                const Octahedron oc = new_octahedron(d.guv + 9 * o, static_cast<float>(e.x), static_cast<float>(e.y),
                                                     static_cast<float>(e.z), xp, yp, zp);
                d.cells[q].gid = h;
                d.cells[q].dl = static_cast<float>(oc.dl);
                d.cells[q].thr = 0.0f;
                d.centre[q] = Centre{oc.cx, oc.cy, oc.cz, 0.0};
                d.cells[q].lm = kLmUnknown;
                d.cells[q].ct = kActive;
                set_active_bit(d, q);
                ncap++;
                const Coord cq{c.x + nb_dx(l), c.y + nb_dy(l), c.z + nb_dz(l)};
                const uint32_t valid = valid_neighbours(cq, gr.X, gr.Y, gr.nzA);
                uint32_t mw[26];
#pragma unroll
                for (int k = 0; k < 26; k++) mw[k] = ((valid >> k) & 1u) ? *mask_word(d, q + gr.off[k]) : kMaskUnknown;
#pragma unroll
                for (int k = 0; k < 26; k++) {
                    const uint32_t bit = 1u << kOpposite.v[k];
                    if ((mw[k] & (kMaskUnknown | bit)) == bit) {
                        const uint32_t old = atomicAnd(mask_word(d, q + gr.off[k]), ~bit);
                        emptied |= static_cast<uint32_t>((old & (kMaskUnknown | kMaskNow)) == bit) << k;
                    }
                }
            }
            // Candidates: one scan and one atomic per wave.
            const uint32_t ce = __builtin_popcount(emptied);
            uint32_t ctot;
            const uint32_t cincl = wave_inclusive_scan(ce, ctot);
            if (ctot) {
                uint32_t cbase = 0;
                if (lane == 63) cbase = atomicAdd(&ctl->ncand, ctot);
                cbase = __shfl(cbase, 63, 64);
                uint32_t pos = cbase + cincl - ce;
                for (uint32_t r = emptied; r; r &= r - 1, pos++)
                    if (pos < d.list_cap) cand[pos] = q + gr.offset(__builtin_ctz(r));
            }
        }
        __syncthreads();
    }
    const unsigned long long t = block_sum(ncap, sh);
    if (threadIdx.x == 0 && t) atomicAdd(&d.stats->captures, t);
}

// Pass C over the candidates: a cell whose every claim-time liquid neighbour was captured in this
// step by a lower-indexed cell has no liquid neighbour left when the reference visits it in index
// order, so it turns solid now (src/CAupdate.cpp:333-334, 722-725); a capture by itself or by a
// later cell leaves it active. A cell not past its liquidus time is not visited this step.
__device__ __forceinline__ void settle_body(const DomainView &d, int p, int layer, uint32_t b, uint32_t nb) {
    __shared__ unsigned long long sh[kBlock / 64];
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    const int cycle = ctl->cycle[p];
    const uint32_t n = ctl->ncand < d.list_cap ? ctl->ncand : d.list_cap;
    const uint32_t *cand = d.list[1 - p];
    const Grid gr = grid_of(d);
    unsigned long long nsol = 0;
    for (uint32_t i = b * kBlock + threadIdx.x; i < n; i += nb * kBlock) {
        const uint32_t r = cand[i];
        if ((d.cells[r].ct & kTypeMask) != kActive) continue;
        if (!(cycle > d.cells[r].cts && d.cells[r].lid <= layer)) continue;
        bool solid = true;
        for (uint32_t m = static_cast<uint32_t>(d.cells[r].lm >> 32); m; m &= m - 1)
            if (d.cells[r + gr.offset(__builtin_ctz(m))].win >= r) {
                solid = false;
                break;
            }
        if (solid) {
            d.cells[r].ct = kSolid;
            clear_active_bit(d, r);
            nsol++;
        }
    }
    const unsigned long long t = block_sum(nsol, sh);
    if (threadIdx.x == 0 && t) atomicAdd(&d.stats->solidified, t);
}

// Device-side end of a domain with autostop: at a check step (cycle a multiple of 1000) with no
// liquid left. Liquid cells only ever leave the liquid state by capture or nucleation, so the
// reference's superheated + undercooled count (src/CAupdate.cpp:738-783) is the initial liquid
// count minus both.
__device__ __forceinline__ void maybe_stop(const DomainView &d, StepCtl *ctl, int p) {
    const int cycle = ctl->cycle[p];
    if (d.autostop && cycle % 1000 == 0 &&
        d.liquid0 - static_cast<long long>(d.stats->captures) - static_cast<long long>(d.stats->activated) == 0) {
        ctl->done = 1;
        ctl->final_cycle = cycle;
        if (d.status) *d.status = cycle;
    }
}

// The next frontier is every set bit of `bits` in ascending cell index, so a wave's cells are
// neighbours in the reference's linear order. Three passes over the bitmap chunks that can hold an
// active cell: per-chunk counts, one exclusive scan of the chunk counts, and the write.
// Every active cell of a step lies in [rlo, rhi]: the write sets it to the first and last cell of
// the new frontier widened by the largest neighbour offset (captures stay within it), and
// nucleation widens it to each new nucleus.
constexpr uint32_t kChunkWords = 8 * kBlock;
constexpr uint32_t kChunkCells = 32 * kChunkWords;

__device__ __forceinline__ bool chunk_range(const DomainView &d, int p, uint32_t &c0, uint32_t &c1) {
    const uint32_t lo = d.ctl->rlo[p];
    uint32_t hi = d.ctl->rhi[p];
    if (lo > hi) return false;
    if (hi >= static_cast<uint64_t>(d.ncell)) hi = static_cast<uint32_t>(d.ncell - 1);
    c0 = lo / kChunkCells;
    c1 = hi / kChunkCells;
    return true;
}

__device__ __forceinline__ void bits_count_body(const DomainView &d, int p, uint32_t b, uint32_t nb) {
    __shared__ unsigned long long sh[kBlock / 64];
    if (d.ctl->done) return;
    uint32_t c0, c1;
    if (!chunk_range(d, p, c0, c1)) return;
    const uint32_t nwords = static_cast<uint32_t>((d.ncell + 31) / 32);
    for (uint32_t c = c0 + b; c <= c1; c += nb) {
        uint32_t cnt = 0;
#pragma unroll
        for (uint32_t k = 0; k < kChunkWords / kBlock; k++) {
            const uint32_t wi = c * kChunkWords + k * kBlock + threadIdx.x;
            if (wi < nwords) cnt += __builtin_popcount(d.bits[wi]);
        }
        const unsigned long long t = block_sum(cnt, sh);
        if (threadIdx.x == 0) d.chunk_sum[c] = static_cast<uint32_t>(t);
    }
}

__device__ __forceinline__ void bits_scan_body(const DomainView &d, int p) {
    __shared__ uint32_t part[kBlock];
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    uint32_t c0 = 1, c1 = 0;
    const bool any = chunk_range(d, p, c0, c1);
    const uint32_t nc = any ? c1 - c0 + 1 : 0;
    const uint32_t per = (nc + kBlock - 1) / kBlock;
    const uint32_t s0 = c0 + threadIdx.x * per;
    const uint32_t s1 = any ? (s0 + per <= c1 + 1 ? s0 + per : c1 + 1) : 0;
    uint32_t mine = 0;
    for (uint32_t c = s0; c < s1; c++) mine += d.chunk_sum[c];
    part[threadIdx.x] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (int t = 0; t < kBlock; t++) {
            const uint32_t v = part[t];
            part[t] = static_cast<uint32_t>(run);
            run += v;
        }
        ctl->count[1 - p] = static_cast<uint32_t>(run < d.list_cap ? run : d.list_cap);
        if (run > d.list_cap) atomicOr(&ctl->error, 1);
        if (run == 0) {
            ctl->rlo[1 - p] = 0xFFFFFFFFu;
            ctl->rhi[1 - p] = 0;
        }
        maybe_stop(d, ctl, p);
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (uint32_t c = s0; c < s1; c++) {
        const uint32_t v = d.chunk_sum[c];
        d.chunk_sum[c] = run;
        run += v;
    }
}

// Chunk c is written in kChunkWords / kBlock rounds of one bitmap word per thread (consecutive
// threads, consecutive words). A round's cells are staged in LDS at their offsets within the round
// and then stored to the list with consecutive lanes on consecutive entries.
__device__ __forceinline__ void bits_write_body(const DomainView &d, int p, uint32_t b, uint32_t nb) {
    __shared__ uint32_t wave_tot[kBlock / 64];
    __shared__ uint32_t stage[32 * kBlock];
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    uint32_t c0, c1;
    if (!chunk_range(d, p, c0, c1)) return;
    const uint32_t nwords = static_cast<uint32_t>((d.ncell + 31) / 32);
    const uint32_t total = ctl->count[1 - p];
    const uint32_t reach = static_cast<uint32_t>(d.X) * static_cast<uint32_t>(d.Y) + static_cast<uint32_t>(d.Y) + 1;
    const uint32_t last = static_cast<uint32_t>(d.ncell - 1);
    uint32_t *next = d.list[1 - p];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t c = c0 + b; c <= c1; c += nb) {
        uint32_t base = d.chunk_sum[c];
        for (uint32_t r = 0; r < kChunkWords / kBlock; r++) {
            const uint32_t wi = c * kChunkWords + r * kBlock + threadIdx.x;
            const uint32_t w = wi < nwords ? d.bits[wi] : 0u;
            const uint32_t cnt = __builtin_popcount(w);
            uint32_t wtot;
            const uint32_t incl = wave_inclusive_scan(cnt, wtot);
            if (lane == 63) wave_tot[wave] = wtot;
            __syncthreads();
            uint32_t o = 0, rtot = 0;
            for (int q = 0; q < kBlock / 64; q++) {
                if (q < wave) o += wave_tot[q];
                rtot += wave_tot[q];
            }
            o += incl - cnt;
            for (uint32_t m = w; m; m &= m - 1) stage[o++] = wi * 32u + __builtin_ctz(m);
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < rtot; k += kBlock) {
                const uint32_t pos = base + k;
                const uint32_t q = stage[k];
                if (pos < d.list_cap) next[pos] = q;
                if (pos == 0) ctl->rlo[1 - p] = q > reach ? q - reach : 0;
                if (pos + 1 == total) ctl->rhi[1 - p] = (last - q > reach) ? q + reach : last;
            }
            base += rtot;
            __syncthreads();
        }
    }
}

// Plain launches: one domain, passed by value.
__global__ __launch_bounds__(kBlock) void k_nucleate(DomainView d, int p, int layer) {
    nucleate_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_claim(DomainView d, int p, int layer) {
    claim_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_apply(DomainView d, int p, int nbc) {
    apply_body(d, p, blockIdx.x, static_cast<uint32_t>(nbc));
}
__global__ __launch_bounds__(kBlock) void k_settle(DomainView d, int p, int layer) {
    settle_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_bits_count(DomainView d, int p) {
    bits_count_body(d, p, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_bits_scan(DomainView d, int p) { bits_scan_body(d, p); }
__global__ __launch_bounds__(kBlock) void k_bits_write(DomainView d, int p) {
    bits_write_body(d, p, blockIdx.x, gridDim.x);
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
__global__ __launch_bounds__(kBlock) void k_apply_n(const DomainView *v, int p, int nbc) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) apply_body(d, p, blockIdx.x, static_cast<uint32_t>(nbc));
}
__global__ __launch_bounds__(kBlock) void k_settle_n(const DomainView *v, int p, int layer) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) settle_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_bits_count_n(const DomainView *v, int p) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) bits_count_body(d, p, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_bits_scan_n(const DomainView *v, int p) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) bits_scan_body(d, p);
}
__global__ __launch_bounds__(kBlock) void k_bits_write_n(const DomainView *v, int p) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) bits_write_body(d, p, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kBlock) void k_check(DomainView d, int cycle, int layer, CheckResult *out) {
    __shared__ unsigned long long sh[kBlock / 64];
    __shared__ unsigned long long shmin[kBlock / 64];
    unsigned long long c[5] = {0, 0, 0, 0, 0};
    unsigned long long next = ~0ull;
    const long long active_end = static_cast<long long>(d.X) * d.Y * d.nzA;
    for (long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; q < d.ncell;
         q += static_cast<long long>(gridDim.x) * kBlock) {
        const int t = d.cells[q].ct & kTypeMask;
        if (d.cells[q].lid == layer) {
            if (t == kLiquid) {
                const int crit = d.cells[q].cts;
                if (crit > cycle)
                    c[0]++;
                else
                    c[1]++;
                const unsigned long long v = static_cast<unsigned long long>(static_cast<long long>(crit));
                if (q < active_end && v < next) next = v;
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
}

__global__ __launch_bounds__(kBlock) void k_seed_init(DomainView d, const uint32_t *cells, uint32_t n) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t q = cells[i];
        const Coord c = decode(q, d.X, d.Y);
        d.cells[q].dl = 0.01f;
        d.centre[q] = Centre{(c.x + d.xoff) + 0.5, (c.y + d.yoff) + 0.5, c.z + 0.5, 0.0};
    }
}

// Initial cell records from the host's compact arrays: no octahedron, no claim, liquid mask not
// yet known; the centres start at zero.
__global__ __launch_bounds__(kBlock) void k_pack(DomainView d, const uint8_t *ct, const int8_t *lid,
                                                 const int32_t *gid, const int32_t *cts) {
    for (long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; q < d.ncell;
         q += static_cast<long long>(gridDim.x) * kBlock) {
        uint4 *c = reinterpret_cast<uint4 *>(d.cells + q);
        c[0] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u);
        c[1] = make_uint4(static_cast<uint32_t>(cts[q]), static_cast<uint32_t>(gid[q]), kClaimNone,
                          static_cast<uint32_t>(static_cast<uint8_t>(lid[q])) | (static_cast<uint32_t>(ct[q]) << 8));
        d.centre[q] = Centre{0.0, 0.0, 0.0, 0.0};
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
        case kFieldDiagonal: static_cast<float *>(out)[q] = c.dl; break;  // float
        case kFieldCritTimeStep: static_cast<int32_t *>(out)[q] = c.cts; break;
        case kFieldLayerId: static_cast<int32_t *>(out)[q] = c.lid; break;
        case kFieldCentre: {  // float x, y, z
            const Centre &e = d.centre[q];
            float *o = static_cast<float *>(out) + 3 * q;
            o[0] = static_cast<float>(e.x);
            o[1] = static_cast<float>(e.y);
            o[2] = static_cast<float>(e.z);
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

void launch_nucleation(const DomainView &d, int parity, int layer, hipStream_t s) {
    k_nucleate<<<grid_for(d.nnuclei, 1024), kBlock, 0, s>>>(d, parity, layer);
}
void launch_claim(const DomainView &d, int parity, int layer, int blocks, hipStream_t s) {
    k_claim<<<blocks, kBlock, 0, s>>>(d, parity, layer);
}
void launch_apply(const DomainView &d, int parity, int layer, int blocks, hipStream_t s) {
    const int cb = static_cast<int>(d.nchunks < static_cast<uint32_t>(blocks) ? d.nchunks : blocks);
    k_apply<<<apply_blocks(blocks), kBlock, 0, s>>>(d, parity, blocks);
    k_settle<<<blocks, kBlock, 0, s>>>(d, parity, layer);
    k_bits_count<<<cb, kBlock, 0, s>>>(d, parity);
    k_bits_scan<<<1, kBlock, 0, s>>>(d, parity);
    k_bits_write<<<cb, kBlock, 0, s>>>(d, parity);
}
void launch_step_timed(const DomainView &d, int parity, int layer, int blocks, hipStream_t s, hipEvent_t *ev) {
    const int cb = static_cast<int>(d.nchunks < static_cast<uint32_t>(blocks) ? d.nchunks : blocks);
    int e = 0;
    auto rec = [&]() { (void)hipEventRecord(ev[e++], s); };
    rec();
    k_nucleate<<<grid_for(d.nnuclei, 1024), kBlock, 0, s>>>(d, parity, layer);
    rec();
    rec();
    k_claim<<<blocks, kBlock, 0, s>>>(d, parity, layer);
    rec();
    rec();
    k_apply<<<apply_blocks(blocks), kBlock, 0, s>>>(d, parity, blocks);
    rec();
    rec();
    k_settle<<<blocks, kBlock, 0, s>>>(d, parity, layer);
    rec();
    rec();
    k_bits_count<<<cb, kBlock, 0, s>>>(d, parity);
    rec();
    rec();
    k_bits_scan<<<1, kBlock, 0, s>>>(d, parity);
    rec();
    rec();
    k_bits_write<<<cb, kBlock, 0, s>>>(d, parity);
    rec();
}
void launch_step_n(const DomainView *views, int slots, int parity, int layer, int nuc_blocks, int blocks,
                   hipStream_t s) {
    k_nucleate_n<<<dim3(nuc_blocks, slots), kBlock, 0, s>>>(views, parity, layer);
    k_claim_n<<<dim3(blocks, slots), kBlock, 0, s>>>(views, parity, layer);
    k_apply_n<<<dim3(apply_blocks(blocks), slots), kBlock, 0, s>>>(views, parity, blocks);
    k_settle_n<<<dim3(blocks, slots), kBlock, 0, s>>>(views, parity, layer);
    k_bits_count_n<<<dim3(blocks, slots), kBlock, 0, s>>>(views, parity);
    k_bits_scan_n<<<dim3(1, slots), kBlock, 0, s>>>(views, parity);
    k_bits_write_n<<<dim3(blocks, slots), kBlock, 0, s>>>(views, parity);
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
void launch_extract(const DomainView &d, int field, void *out, hipStream_t s) {
    k_extract<<<grid_for(d.ncell, 1 << 16), kBlock, 0, s>>>(d, field, out);
}
void launch_copy(const uint4 *src, uint4 *dst, size_t n, hipStream_t s) {
    k_copy<<<grid_for(static_cast<long long>(n), 1 << 16), kBlock, 0, s>>>(src, dst, n);
}
void launch_geometry_probe(const float *guv, const double *in, double *out, int n, hipStream_t s) {
    k_geometry_probe<<<grid_for(n, 1 << 20), kBlock, 0, s>>>(guv, in, out, n);
}

}  // namespace exaca
