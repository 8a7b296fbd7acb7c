// ca_kernels.hip -- the CA time step as memory-bound CDNA4 kernels over a compacted frontier.
//
// One step (reference: src/runCA.cpp:291-311) is three launches on one stream:
//   k_nucleate  activates the pre-sampled nuclei whose time has come (src/CAupdate.cpp:67-200)
//   k_claim     for every frontier cell past its liquidus time: growth of its octahedron, scan of
//               the 26 neighbours, and a claim on every liquid neighbour it reaches, written as
//               atomicMin of (generation << 32 | capturer index) into win[]
//   k_apply     the winner of each claimed cell builds the new decentred octahedron; a frontier
//               cell whose liquid neighbours are all taken by lower-indexed cells turns solid;
//               survivors and new cells are appended to the next frontier list
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
struct Grid {
    Divisor xy, y;
    int X, Y, nzA;
    int off[26];  // linear offsets of the 26 neighbours
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
__device__ __forceinline__ uint32_t nb_index(uint32_t p, int l, int X, int Y) {
    return static_cast<uint32_t>(static_cast<long long>(p) + (static_cast<long long>(kNbZ[l]) * X + kNbX[l]) * Y +
                                 kNbY[l]);
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
    const double ex = xp + kNbX[l] - cx, ey = yp + kNbY[l] - cy, ez = zp + kNbZ[l] - cz;
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
                                               uint32_t &begin, uint32_t &end, uint32_t *slice_out = nullptr) {
    const uint32_t n = count < cap ? count : cap;  // an overflowed list is flagged, never over-read
    const uint32_t logical = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    if (slice_out) *slice_out = logical;
    uint32_t chunk = (n + nb - 1) / nb;
    chunk = (chunk + kBlock - 1) / kBlock * kBlock;
    const unsigned long long b0 = static_cast<unsigned long long>(logical) * chunk;
    begin = b0 >= n ? n : static_cast<uint32_t>(b0);
    end = (b0 + chunk >= n) ? n : static_cast<uint32_t>(b0 + chunk);
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
        ctl->serial[1 - p] = ctl->serial[p] + 1;
        ctl->count[1 - p] = 0;
    }
    unsigned long long nn = 0, act = 0;
    for (int i = b * kBlock + threadIdx.x; i < d.nnuclei; i += nb * kBlock) {
        const uint32_t q = static_cast<uint32_t>(d.nloc[i]);
        if (!(cycle >= d.ntime[i] && d.ct[q] == kLiquid && d.lid[q] <= layer)) continue;
        const Coord c = decode(q, d.X, d.Y);
        if (c.x > 0 && c.x < d.X - 1 && c.y > 0 && c.y < d.Y - 1) nn++;
        d.dl[q] = 0.01f;
        d.cx[q] = (c.x + d.xoff) + 0.5;
        d.cy[q] = (c.y + d.yoff) + 0.5;
        d.cz[q] = c.z + 0.5;
        d.ct[q] = kActive | kSeedBit;
        if (d.bits) atomicOr(&d.bits[q >> 5], 1u << (q & 31));
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

// Pass A. All per-cell loads (own state, octahedron, the 26 neighbour types) are issued before any
// of them is used, so a frontier cell costs one memory round trip; a neighbour outside the slab
// loads the cell's own (active, never liquid) type instead and is masked out.
__device__ __forceinline__ void claim_body(const DomainView &d, int p, int layer, uint32_t b, uint32_t nb) {
    __shared__ unsigned long long sh[kBlock / 64];
    const StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    const int cycle = ctl->cycle[p];
    const unsigned long long stamp = static_cast<unsigned long long>(0xFFFFFFFFu - ctl->serial[p]) << 32;
    const int njumps = ctl->njumps;
    uint32_t begin, end;
    frontier_slice(ctl->count[p], d.list_cap, b, nb, begin, end);
    const uint32_t *list = d.list[p];
    const Grid gr = grid_of(d);
    unsigned long long front = 0;
    for (uint32_t i = begin + threadIdx.x; i < end; i += kBlock) {
        const uint32_t j = list[i];
        const int t = d.cts[j];
        const int8_t lay = d.lid[j];
        const float dl0 = d.dl[j];
        const uint8_t own = d.ct[j];
        const int h = d.gid[j];
        const double cx = d.cx[j], cy = d.cy[j], cz = d.cz[j];
        const Coord c = decode(j, gr);
        const uint32_t valid = valid_neighbours(c, gr.X, gr.Y, gr.nzA);
        uint32_t liquid = 0;
#pragma unroll
        for (int l = 0; l < 26; l++) {
            const uint32_t q = ((valid >> l) & 1u) ? j + gr.off[l] : j;
            liquid |= static_cast<uint32_t>(d.ct[q] == kLiquid) << l;
        }
        if (!(cycle > t && lay <= layer)) {
            d.mask[i] = kWaitMask;
            continue;
        }
        front++;
        int k = executed_since(t, cycle, ctl, njumps);
        k = k < 0 ? 0 : (k < d.ngrow - 1 ? k : d.ngrow - 1);
        const float dlj = static_cast<float>(static_cast<double>(dl0) + d.grow[k]);
        d.dl[j] = dlj;
        d.mask[i] = liquid;
        if (!liquid) continue;
        int o = orientation_of(h);
        if (o < 0) {
            atomicOr(&d.ctl->error, 2);
            o = 0;
        }
        const unsigned long long key = stamp | j;
        if (own & kSeedBit) {
            const float *tab = d.seed_cdl + 26 * o;
            for (uint32_t m = liquid; m; m &= m - 1) {
                const int l = __builtin_ctz(m);
                if (dlj >= tab[l]) atomicMin(&d.win[j + gr.off[l]], key);
            }
        }
        else {
            const Faces f = faces_of(d.guv + 9 * o);
            const double xp = (c.x + d.xoff) + 0.5, yp = (c.y + d.yoff) + 0.5, zp = c.z + 0.5;
            // A neighbour whose distance from the centre bounds its critical length above the
            // half-diagonal cannot be reached this step; only the others need the exact length.
            const double reach = d.reach[o];
            const double dl2 = static_cast<double>(dlj) * static_cast<double>(dlj);
            for (uint32_t m = liquid; m; m &= m - 1) {
                const int l = __builtin_ctz(m);
                const double ex = xp + kNbX[l] - cx, ey = yp + kNbY[l] - cy, ez = zp + kNbZ[l] - cz;
                if (dl2 < reach * (ex * ex + ey * ey + ez * ez)) continue;
                if (dlj >= critical_length(f, xp, yp, zp, cx, cy, cz, l)) atomicMin(&d.win[j + gr.off[l]], key);
            }
        }
    }
    const unsigned long long tot = block_sum(front, sh);
    if (threadIdx.x == 0 && tot) atomicAdd(&d.stats->frontier, tot);
}

__device__ __forceinline__ void maybe_stop(const DomainView &d, StepCtl *ctl, int p);

// Pass B. Per frontier entry the outcome is written back over its mask as (keep << 30 | captured
// neighbours); every workgroup stores the number of next-frontier entries its slice produces so
// that k_emit can lay the next frontier out in the same order as the current one.
__device__ __forceinline__ void apply_body(const DomainView &d, int p, uint32_t b, uint32_t nb) {
    __shared__ unsigned long long sh[kBlock / 64];
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    const unsigned long long stamp = static_cast<unsigned long long>(0xFFFFFFFFu - ctl->serial[p]) << 32;
    uint32_t begin, end, slice;
    frontier_slice(ctl->count[p], d.list_cap, b, nb, begin, end, &slice);
    const uint32_t *list = d.list[p];
    const Grid gr = grid_of(d);
    unsigned long long ncap = 0, nsol = 0, nwait = 0, produced = 0;
    for (uint32_t i = begin + threadIdx.x; i < end; i += kBlock) {
        const uint32_t j = list[i];
        const uint32_t m = d.mask[i];
        uint32_t captured = 0, keep = 0;
        if (m == kWaitMask) {
            keep = 1;
            nwait++;
        }
        else {
            // Claims on every liquid neighbour, loaded together (masked lanes reload win[j]).
#pragma unroll
            for (int l = 0; l < 26; l++) {
                const uint32_t on = (m >> l) & 1u;
                const unsigned long long w = d.win[on ? j + gr.off[l] : j];
                const uint32_t now = (w & 0xFFFFFFFF00000000ull) == stamp;
                const uint32_t wj = static_cast<uint32_t>(w);
                // claimed this step: liquid for j only if the winner comes later in index order;
                // unclaimed: still liquid
                keep |= on & (now ? static_cast<uint32_t>(wj >= j) : 1u);
                captured |= (on & now & static_cast<uint32_t>(wj == j)) << l;
            }
            if (!keep) {
                d.ct[j] = kSolid;
                if (d.bits) atomicAnd(&d.bits[j >> 5], ~(1u << (j & 31)));
                nsol++;
            }
        }
        if (captured) {
            const int h = d.gid[j];
            int o = orientation_of(h);
            if (o < 0) {
                atomicOr(&ctl->error, 2);
                o = 0;
            }
            const float *g = d.guv + 9 * o;
            const Coord c = decode(j, gr);
            const double cxold = static_cast<float>(d.cx[j]);
            const double cyold = static_cast<float>(d.cy[j]);
            const double czold = static_cast<float>(d.cz[j]);
            for (uint32_t r = captured; r; r &= r - 1) {
                const int l = __builtin_ctz(r);
                const uint32_t q = j + gr.off[l];
                const double xp = (c.x + d.xoff + kNbX[l]) + 0.5;
                const double yp = (c.y + d.yoff + kNbY[l]) + 0.5;
                const double zp = (c.z + kNbZ[l]) + 0.5;
                const Octahedron oc = new_octahedron(g, cxold, cyold, czold, xp, yp, zp);
                d.gid[q] = h;
                d.dl[q] = static_cast<float>(oc.dl);
                d.cx[q] = oc.cx;
                d.cy[q] = oc.cy;
                d.cz[q] = oc.cz;
                d.ct[q] = kActive;
                if (d.bits) atomicOr(&d.bits[q >> 5], 1u << (q & 31));
                ncap++;
            }
        }
        d.mask[i] = (keep << 30) | captured;
        produced += keep + __builtin_popcount(captured);
    }
    unsigned long long t = block_sum(produced, sh);
    if (threadIdx.x == 0) d.slice_out[slice] = static_cast<uint32_t>(t);
    t = block_sum(ncap, sh);
    if (threadIdx.x == 0 && t) atomicAdd(&d.stats->captures, t);
    t = block_sum(nsol, sh);
    if (threadIdx.x == 0 && t) atomicAdd(&d.stats->solidified, t);
    t = block_sum(nwait, sh);
    if (threadIdx.x == 0 && t) atomicAdd(&d.stats->waiting, t);
}

// Writes the next frontier: slice s's entries start after those of slices 0 .. s-1, and within a
// slice each entry's survivor comes first, then its captured neighbours in neighbour order. The
// front therefore keeps the spatial order of the current list (a column-ordered front stays
// column-ordered), which keeps the per-cell reads of the next step coalesced.
// With autostop set, the workgroup of slice 0 also ends the domain at a check step (cycle a
// multiple of 1000) that finds no liquid left: liquid cells only ever leave the liquid state by
// capture or nucleation, so the reference's superheated + undercooled count
// (src/CAupdate.cpp:738-783) is the initial liquid count minus both.
__device__ __forceinline__ void emit_body(const DomainView &d, int p, uint32_t b, uint32_t nb) {
    __shared__ unsigned long long sh[kBlock / 64];
    __shared__ uint32_t wave_tot[kBlock / 64];
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    uint32_t begin, end, slice;
    frontier_slice(ctl->count[p], d.list_cap, b, nb, begin, end, &slice);
    unsigned long long before = 0, all = 0;
    for (uint32_t s = threadIdx.x; s < nb; s += kBlock) {
        const uint32_t v = d.slice_out[s];
        all += v;
        if (s < slice) before += v;
    }
    before = block_sum(before, sh);
    all = block_sum(all, sh);
    __shared__ unsigned long long base_sh;
    if (threadIdx.x == 0) {
        base_sh = before;
        if (slice == 0) {
            ctl->count[1 - p] = static_cast<uint32_t>(all < d.list_cap ? all : d.list_cap);
            if (all > d.list_cap) atomicOr(&ctl->error, 1);
            maybe_stop(d, ctl, p);
        }
    }
    __syncthreads();
    unsigned long long base = base_sh;
    const uint32_t *list = d.list[p];
    uint32_t *next = d.list[1 - p];
    const Grid gr = grid_of(d);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t i0 = begin; i0 < end; i0 += kBlock) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t j = 0, m = 0;
        if (i < end) {
            j = list[i];
            m = d.mask[i];
        }
        const uint32_t keep = (m >> 30) & 1u, captured = m & ((1u << 26) - 1);
        const uint32_t mine = keep + __builtin_popcount(captured);
        uint32_t wtot;
        const uint32_t incl = wave_inclusive_scan(mine, wtot);
        if (lane == 63) wave_tot[wave] = wtot;
        __syncthreads();
        uint32_t wbase = 0, round = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            if (w < wave) wbase += wave_tot[w];
            round += wave_tot[w];
        }
        unsigned long long pos = base + wbase + incl - mine;
        if (pos + mine <= d.list_cap) {
            if (keep) next[pos++] = j;
            for (uint32_t r = captured; r; r &= r - 1) next[pos++] = j + gr.off[__builtin_ctz(r)];
        }
        base += round;
        __syncthreads();
    }
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

// Index-ordered frontier: the next list is every set bit of `bits` in ascending cell index, so a
// wave's cells are neighbours in the reference's linear order. Three passes over the bitmap:
// per-chunk counts, one exclusive scan of the chunk counts, and the write.
constexpr uint32_t kChunkWords = 8 * kBlock;

__device__ __forceinline__ void bits_count_body(const DomainView &d, uint32_t b, uint32_t nb) {
    __shared__ unsigned long long sh[kBlock / 64];
    if (d.ctl->done) return;
    const uint32_t nwords = static_cast<uint32_t>((d.ncell + 31) / 32);
    for (uint32_t c = b; c < d.nchunks; c += nb) {
        const uint32_t w0 = c * kChunkWords + threadIdx.x * 8;
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (w0 + k < nwords) cnt += __builtin_popcount(d.bits[w0 + k]);
        const unsigned long long t = block_sum(cnt, sh);
        if (threadIdx.x == 0) d.chunk_sum[c] = static_cast<uint32_t>(t);
    }
}

__device__ __forceinline__ void bits_scan_body(const DomainView &d, int p) {
    __shared__ uint32_t part[kBlock];
    StepCtl *ctl = d.ctl;
    if (ctl->done) return;
    const uint32_t per = (d.nchunks + kBlock - 1) / kBlock;
    const uint32_t c0 = threadIdx.x * per;
    uint32_t mine = 0;
    for (uint32_t c = c0; c < c0 + per && c < d.nchunks; c++) mine += d.chunk_sum[c];
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
        maybe_stop(d, ctl, p);
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (uint32_t c = c0; c < c0 + per && c < d.nchunks; c++) {
        const uint32_t v = d.chunk_sum[c];
        d.chunk_sum[c] = run;
        run += v;
    }
}

__device__ __forceinline__ void bits_write_body(const DomainView &d, int p, uint32_t b, uint32_t nb) {
    __shared__ uint32_t wave_tot[kBlock / 64];
    if (d.ctl->done) return;
    const uint32_t nwords = static_cast<uint32_t>((d.ncell + 31) / 32);
    uint32_t *next = d.list[1 - p];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t c = b; c < d.nchunks; c += nb) {
        const uint32_t w0 = c * kChunkWords + threadIdx.x * 8;
        uint32_t w[8], cnt = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            w[k] = (w0 + k < nwords) ? d.bits[w0 + k] : 0u;
            cnt += __builtin_popcount(w[k]);
        }
        uint32_t wtot;
        const uint32_t incl = wave_inclusive_scan(cnt, wtot);
        if (lane == 63) wave_tot[wave] = wtot;
        __syncthreads();
        uint32_t wbase = 0;
        for (int q = 0; q < wave; q++) wbase += wave_tot[q];
        unsigned long long pos = static_cast<unsigned long long>(d.chunk_sum[c]) + wbase + incl - cnt;
        if (pos + cnt <= d.list_cap) {
#pragma unroll
            for (int k = 0; k < 8; k++)
                for (uint32_t r = w[k]; r; r &= r - 1) next[pos++] = (w0 + k) * 32u + __builtin_ctz(r);
        }
        __syncthreads();
    }
}

// Plain launches: one domain, passed by value.
__global__ __launch_bounds__(kBlock) void k_nucleate(DomainView d, int p, int layer) {
    nucleate_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_claim(DomainView d, int p, int layer) {
    claim_body(d, p, layer, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_apply(DomainView d, int p) { apply_body(d, p, blockIdx.x, gridDim.x); }
__global__ __launch_bounds__(kBlock) void k_emit(DomainView d, int p) { emit_body(d, p, blockIdx.x, gridDim.x); }
__global__ __launch_bounds__(kBlock) void k_bits_count(DomainView d) { bits_count_body(d, blockIdx.x, gridDim.x); }
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
__global__ __launch_bounds__(kBlock) void k_apply_n(const DomainView *v, int p) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl) apply_body(d, p, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_emit_n(const DomainView *v, int p) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl && !d.index_order) emit_body(d, p, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_bits_count_n(const DomainView *v) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl && d.index_order) bits_count_body(d, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_bits_scan_n(const DomainView *v, int p) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl && d.index_order) bits_scan_body(d, p);
}
__global__ __launch_bounds__(kBlock) void k_bits_write_n(const DomainView *v, int p) {
    const DomainView &d = v[blockIdx.y];
    if (d.ctl && d.index_order) bits_write_body(d, p, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kBlock) void k_check(DomainView d, int cycle, int layer, CheckResult *out) {
    __shared__ unsigned long long sh[kBlock / 64];
    __shared__ unsigned long long shmin[kBlock / 64];
    unsigned long long c[5] = {0, 0, 0, 0, 0};
    unsigned long long next = ~0ull;
    const long long active_end = static_cast<long long>(d.X) * d.Y * d.nzA;
    for (long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; q < d.ncell;
         q += static_cast<long long>(gridDim.x) * kBlock) {
        const int t = d.ct[q] & kTypeMask;
        if (d.lid[q] == layer) {
            if (t == kLiquid) {
                const int crit = d.cts[q];
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
        d.dl[q] = 0.01f;
        d.cx[q] = (c.x + d.xoff) + 0.5;
        d.cy[q] = (c.y + d.yoff) + 0.5;
        d.cz[q] = c.z + 0.5;
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
    (void)layer;
    k_apply<<<blocks, kBlock, 0, s>>>(d, parity);
    if (d.index_order) {
        const int cb = static_cast<int>(d.nchunks < static_cast<uint32_t>(blocks) ? d.nchunks : blocks);
        k_bits_count<<<cb, kBlock, 0, s>>>(d);
        k_bits_scan<<<1, kBlock, 0, s>>>(d, parity);
        k_bits_write<<<cb, kBlock, 0, s>>>(d, parity);
    }
    else {
        k_emit<<<blocks, kBlock, 0, s>>>(d, parity);
    }
}
void launch_step_timed(const DomainView &d, int parity, int layer, int blocks, hipStream_t s, hipEvent_t *ev) {
    auto rec = [&](int i) { (void)hipEventRecord(ev[i], s); };
    rec(0);
    k_nucleate<<<grid_for(d.nnuclei, 1024), kBlock, 0, s>>>(d, parity, layer);
    rec(1);
    rec(2);
    k_claim<<<blocks, kBlock, 0, s>>>(d, parity, layer);
    rec(3);
    rec(4);
    k_apply<<<blocks, kBlock, 0, s>>>(d, parity);
    rec(5);
    const int cb = static_cast<int>(d.nchunks < static_cast<uint32_t>(blocks) ? d.nchunks : blocks);
    rec(6);
    if (d.index_order)
        k_bits_count<<<cb, kBlock, 0, s>>>(d);
    else
        k_emit<<<blocks, kBlock, 0, s>>>(d, parity);
    rec(7);
    rec(8);
    if (d.index_order) k_bits_scan<<<1, kBlock, 0, s>>>(d, parity);
    rec(9);
    rec(10);
    if (d.index_order) k_bits_write<<<cb, kBlock, 0, s>>>(d, parity);
    rec(11);
}
void launch_step_n(const DomainView *views, int slots, int parity, int layer, int nuc_blocks, int blocks,
                   hipStream_t s) {
    k_nucleate_n<<<dim3(nuc_blocks, slots), kBlock, 0, s>>>(views, parity, layer);
    k_claim_n<<<dim3(blocks, slots), kBlock, 0, s>>>(views, parity, layer);
    k_apply_n<<<dim3(blocks, slots), kBlock, 0, s>>>(views, parity);
    k_emit_n<<<dim3(blocks, slots), kBlock, 0, s>>>(views, parity);
    k_bits_count_n<<<dim3(blocks, slots), kBlock, 0, s>>>(views);
    k_bits_scan_n<<<dim3(1, slots), kBlock, 0, s>>>(views, parity);
    k_bits_write_n<<<dim3(blocks, slots), kBlock, 0, s>>>(views, parity);
}
void launch_check(const DomainView &d, int cycle, int layer, CheckResult *out, hipStream_t s) {
    k_check<<<grid_for(d.ncell, 4096), kBlock, 0, s>>>(d, cycle, layer, out);
}
void launch_seed_init(const DomainView &d, const uint32_t *cells, uint32_t n, hipStream_t s) {
    k_seed_init<<<grid_for(n, 1024), kBlock, 0, s>>>(d, cells, n);
}
void launch_copy(const uint4 *src, uint4 *dst, size_t n, hipStream_t s) {
    k_copy<<<grid_for(static_cast<long long>(n), 1 << 16), kBlock, 0, s>>>(src, dst, n);
}
void launch_geometry_probe(const float *guv, const double *in, double *out, int n, hipStream_t s) {
    k_geometry_probe<<<grid_for(n, 1 << 20), kBlock, 0, s>>>(guv, in, out, n);
}

}  // namespace exaca
