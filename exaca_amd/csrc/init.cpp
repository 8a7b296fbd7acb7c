// init.cpp -- host construction of the initial CA state for one rank (np = 1). Each function
// names the reference routine whose loop order and arithmetic it keeps; the random streams are
// libstdc++'s, drawn in the reference's order, so the state is identical cell for cell.
#include "init.hpp"

#include <algorithm>
#include <atomic>
#include <system_error>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>

#include "ca_common.hpp"

namespace exaca {
namespace {

inline long long cell(const HostDomain &h, int k, int i, int j) {
    return (static_cast<long long>(k) * h.X + i) * h.Y + j;
}

// Runs f(k) for k = k0 .. k1-1 of h's planes on up to 16 threads, one per 256 K cells or more
// (small domains run inline); f(k) may write only plane k's cells.
template <typename F>
void for_planes(const HostDomain &h, int k0, int k1, F f) {
    const int n = k1 - k0;
    if (n <= 0) return;
    const long long by_work = std::max(1LL, static_cast<long long>(h.X) * h.Y * n / (1 << 18));
    const int nt = static_cast<int>(std::min<long long>(
        {static_cast<long long>(n), by_work,
         static_cast<long long>(std::min(16u, std::max(1u, std::thread::hardware_concurrency())))}));
    std::atomic<int> next{k0};
    auto work = [&] {
        for (int k; (k = next.fetch_add(1)) < k1;) f(k);
    };
    std::vector<std::thread> th;
    try {
        for (int t = 1; t < nt; t++) th.emplace_back(work);
    } catch (const std::system_error &) {
        // fewer threads than asked for: the ones running (and this one) take every plane
    }
    work();
    for (auto &x : th) x.join();
}

// The 64-bit Mersenne Twister with the parameters the C++ standard fixes for std::mt19937_64
// ([rand.predef]; Matsumoto and Nishimura's MT19937-64), refilled 312 outputs at a time: the
// same sequence as std::mt19937_64 with the same seed, without its per-call bookkeeping.
class Mt64Block {
  public:
    static constexpr int kN = 312, kM = 156;
    explicit Mt64Block(uint64_t seed) {
        s_[0] = seed;
        for (int i = 1; i < kN; i++) s_[i] = 6364136223846793005ULL * (s_[i - 1] ^ (s_[i - 1] >> 62)) + i;
    }
    // the next 312 outputs
    const uint64_t *refill() {
        constexpr uint64_t kUpper = ~0ULL << 31, kLower = (1ULL << 31) - 1, kA = 0xB5026F5AA96619E9ULL;
        auto mix = [&](uint64_t hi, uint64_t lo, uint64_t far) {
            const uint64_t y = (hi & kUpper) | (lo & kLower);
            return far ^ (y >> 1) ^ (-(y & 1) & kA);  // branch-free, so the loops vectorise
        };
        for (int i = 0; i < kN - kM; i++) s_[i] = mix(s_[i], s_[i + 1], s_[i + kM]);
        for (int i = kN - kM; i < kN - 1; i++) s_[i] = mix(s_[i], s_[i + 1], s_[i + kM - kN]);
        s_[kN - 1] = mix(s_[kN - 1], s_[0], s_[kM - 1]);
        for (int i = 0; i < kN; i++) {
            uint64_t z = s_[i];
            z ^= (z >> 29) & 0x5555555555555555ULL;
            z ^= (z << 17) & 0x71D67FFFEDA60000ULL;
            z ^= (z << 37) & 0xFFF7EEE000000000ULL;
            out_[i] = z ^ (z >> 43);
        }
        return out_;
    }

  private:
    uint64_t s_[kN], out_[kN];
};

// libstdc++'s uniform_real_distribution<double>(0, 1) over a 64-bit engine draws
// generate_canonical with one engine call: double(x) * 2^-64, clamped below 1. That is
// nondecreasing in x, so `draw < v` holds exactly for x below the cut returned here (all x when
// `all` is set).
uint64_t draw_cut(double v, bool &all) {
    auto below = [v](uint64_t x) {
        double u = static_cast<double>(x) * 0x1p-64;
        if (u >= 1.0) u = std::nextafter(1.0, 0.0);
        return u < v;
    };
    all = below(~0ULL);
    if (all) return 0;
    uint64_t lo = 0, hi = ~0ULL;  // below(hi) is false; find the least such x
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (below(mid))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Indices (in draw order) of the draws below v among the first n draws of
// uniform_real_distribution<double>(0, 1) over mt19937_64(seed).
std::vector<long long> draws_below(uint64_t seed, double v, long long n) {
    bool all = false;
    const uint64_t cut = draw_cut(v, all);
    std::vector<long long> hits;
    Mt64Block g(seed);
    for (long long base = 0; base < n; base += Mt64Block::kN) {
        const uint64_t *x = g.refill();
        const int m = static_cast<int>(std::min<long long>(Mt64Block::kN, n - base));
        for (int i = 0; i < m; i++)
            if (all || x[i] < cut) hits.push_back(base + i);
    }
    return hits;
}

// FindItBounds-equivalent neighbour filter: a neighbour exists when it stays inside the rank's
// X-Y slab (src/CAfunctions.cpp:169-213).
inline bool in_slab(const HostDomain &h, int i, int j, int n) {
    int x = i + kNbX[n], y = j + kNbY[n];
    return x >= 0 && x < h.X && y >= 0 && y < h.Y;
}

// CritTimeStep of every cell of plane k in a directional-solidification domain.
int crit_directional(const Deck &d, int k) {
    return static_cast<int>(((k - 1) * d.G * d.deltax) / (d.R * d.deltat));
}

// TempInit_DirSolidification (src/CAinitialize.cpp:1025-1055).
void temperature_directional(const Deck &d, HostDomain &h) {
    h.zlow = 0;
    h.zhigh = h.nz - 1;
    h.nzA = h.nz;
    h.dU = static_cast<float>(d.R * d.deltat);
    for_planes(h, 0, h.nz, [&](int k) {
        const int crit = crit_directional(d, k);
        for (int i = 0; i < h.X; i++)
            for (int j = 0; j < h.Y; j++) {
                long long p = cell(h, k, i, j);
                h.lid[p] = 0;
                h.cts[p] = crit;
                int gx = i + h.xoff, gy = j + h.yoff;
                h.melted[p] = (gx > -1 && gx < h.nx && gy > -1 && gy < h.ny && k > 0 && k < h.nz - 1);
            }
    });
}

// TempInit_SpotMelt (src/CAinitialize.cpp:1058-1124). Cells outside every spot keep
// CritTimeStep 0: the reference leaves them unset in a fresh (zero) allocation.
void temperature_spots(const Deck &d, HostDomain &h) {
    h.zlow = 0;
    h.zhigh = d.spot_radius + 1;
    h.nzA = d.spot_radius + 2;
    h.dU = static_cast<float>(d.R * d.deltat);
    std::fill(h.lid.begin(), h.lid.end(), static_cast<int8_t>(-1));
    const int nspots = d.nspots_x * d.nspots_y;
    const float iso = static_cast<float>((d.R / d.G) * d.deltat / d.deltax);
    const int between = static_cast<int>(d.spot_radius / iso + (d.freezing_range / d.R) / d.deltat);
    for (int layer = 0; layer < d.num_layers; layer++)
        for (int n = 0; n < nspots; n++) {
            const int xs = 2 + d.spot_radius + (n % d.nspots_x) * d.spot_offset;
            const int ys = 2 + d.spot_radius + (n / d.nspots_x) * d.spot_offset;
            const int zs = d.spot_radius + 2 + d.layer_height * layer;
            for (int k = 0; k < h.nzA; k++) {
                const float dz = static_cast<float>(zs - (k + d.layer_height * layer));
                for (int i = 0; i < h.X; i++) {
                    const float dx = static_cast<float>(xs - (i + h.xoff));
                    for (int j = 0; j < h.Y; j++) {
                        const float dy = static_cast<float>(ys - (j + h.yoff));
                        const float dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                        if (dist <= d.spot_radius) {
                            long long p = cell(h, k + layer * d.layer_height, i, j);
                            h.cts[p] = 1 + static_cast<int>((static_cast<float>(d.spot_radius) - dist) / iso) +
                                       between * n;
                            h.lid[p] = static_cast<int8_t>(layer);
                            h.melted[p] = 1;
                        }
                    }
                }
            }
        }
}

// SubstrateInit_ConstrainedGrowth (src/CAinitialize.cpp:1460-1529), rank 0 of 1.
// The draws happen on plane 1 only, so the other planes (walls and liquid) run on threads, and
// the active cells to number are plane 1's.
void substrate_directional(const Deck &d, HostDomain &h) {
    std::mt19937_64 gen(d.seed);
    std::uniform_real_distribution<double> dis(0.0, 1.0);
    auto plane = [&](int k) {
        for (int i = 0; i < h.X; i++)
            for (int j = 0; j < h.Y; j++) {
                const int gx = i + h.xoff, gy = j + h.yoff;
                long long p = cell(h, k, i, j);
                if (gx == -1 || gx == h.nx || gy == -1 || gy == h.ny || k == 0 || k == h.nz - 1) {
                    h.ct[p] = kWall;
                    h.gid[p] = 0;
                }
                else if (k == 1 && i > 0 && i < h.X - 1 && j > 0 && j < h.Y - 1) {
                    h.ct[p] = dis(gen) < d.fract_active ? kActive : kLiquid;
                }
                else {
                    h.ct[p] = kLiquid;
                }
            }
    };
    for_planes(h, 0, h.nz, [&](int k) {
        if (k != 1) plane(k);
    });
    if (h.nz < 2) return;
    plane(1);
    int next = 1;
    for (long long p = cell(h, 1, 0, 0); p < cell(h, 2, 0, 0); p++)
        if (h.ct[p] == kActive) {
            h.gid[p] = next++;
            h.substrate_active++;
        }
}

// SubstrateInit_FromFile (src/CAinitialize.cpp:1532-1592).
void substrate_file(const Deck &d, HostDomain &h) {
    std::ifstream in(d.substrate_file);
    std::string s;
    int dims[3];
    for (int q = 0; q < 3; q++) {
        std::getline(in, s);
        dims[q] = std::stoi(s.substr(s.find('=') + 1), nullptr, 10);
    }
    const int nzS = dims[0], nyS = dims[1], nxS = dims[2];
    if (nzS < h.nz)
        throw std::runtime_error("Error: only " + std::to_string(nzS) + " layers of substrate data are present in the file " +
                                 d.substrate_file + " ; at least " + std::to_string(h.nz) +
                                 " layers of substrate data are required to simulate the specified solidification problem");
    for (int k = 0; k < nzS && k != h.nz; k++)
        for (int j = 0; j < nyS; j++)
            for (int i = 0; i < nxS; i++) {
                std::getline(in, s);
                if (i >= h.xoff && i < h.xoff + h.X && j >= h.yoff && j < h.yoff + h.Y) {
                    long long p = cell(h, k, i - h.xoff, j - h.yoff);
                    h.gid[p] = (h.cts[p] == 0) ? std::stoi(s, nullptr, 10) : 0;
                }
            }
}

// SubstrateInit_FromGrainSpacing (src/CAinitialize.cpp:1597-1732): baseplate seeds drawn from
// mt19937_64(1) over the global grid, then a Voronoi assignment of the substrate cells in the
// first layer's active window (nearest seed by float distance; the lowest seed index wins ties,
// as with the reference's strict < scan), then one new grain per powder-layer cell above it.
void substrate_spacing(const Deck &d, HostDomain &h) {
    std::mt19937_64 gen(1);
    std::uniform_real_distribution<double> dis(0.0, 1.0);
    const float sp = d.substrate_spacing;
    const double prob = (d.deltax * d.deltax * d.deltax) / (sp * sp * sp * std::pow(10, -18));
    std::vector<int> bx, by, bz;
    for (int k = 1; k <= h.nzA; k++)
        for (int i = 1; i < h.nx - 1; i++)
            for (int j = 1; j < h.ny - 1; j++)
                if (dis(gen) < prob) {
                    bx.push_back(i);
                    by.push_back(j);
                    bz.push_back(k);
                }
    const int nb = static_cast<int>(bx.size());
    h.baseplate_grains = nb;
    const float far = static_cast<float>(static_cast<long long>(h.X) * h.Y * h.nzA);
    // Each cell's nearest seed is independent of every other cell: planes are shared out
    // across host threads.
    auto plane = [&](int k) {
        for (int i = 0; i < h.X; i++)
            for (int j = 0; j < h.Y; j++) {
                const int gx = i + h.xoff, gy = j + h.yoff;
                long long p = cell(h, k, i, j);
                if (h.cts[p] != 0) {
                    h.gid[p] = 0;
                    continue;
                }
                float best = far;
                int closest = -1;
                for (int n = 0; n < nb; n++) {
                    const float ex = static_cast<float>(std::abs(bx[n] - gx));
                    const float ey = static_cast<float>(std::abs(by[n] - gy));
                    const float ez = static_cast<float>(std::abs(bz[n] - k));
                    const float dist = sqrtf(ex * ex + ey * ey + ez * ez);
                    if (dist < best) {
                        closest = n;
                        best = dist;
                    }
                }
                h.gid[p] = closest + 1;
            }
    };
    std::atomic<int> next_plane{1};
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nthreads; t++)
        pool.emplace_back([&] {
            for (int k = next_plane++; k < h.nzA; k = next_plane++) plane(k);
        });
    for (auto &th : pool) th.join();
    for (int k = h.nzA; k < h.nz; k++)
        for (int i = 1; i < h.X - 1; i++)
            for (int j = 1; j < h.Y - 1; j++) {
                long long p = cell(h, k, i, j);
                if (h.cts[p] != 0) h.gid[p] = 0;
            }
    int next = nb + 1;
    for (long long p = static_cast<long long>(h.X) * h.Y * h.nzA; p < h.ncell; p++)
        if (h.cts[p] == 0) h.gid[p] = next++;
}

// ActiveCellWallInit (src/CAinitialize.cpp:1736-1830) with ExtraWalls = false: walls on the
// ghost ring and the top/bottom planes, solid where no melting data exists, liquid elsewhere,
// then solid cells touching a melted non-wall neighbour become active.
void walls_and_interface(HostDomain &h) {
    for (int k = 0; k < h.nz; k++)
        for (int i = 0; i < h.X; i++)
            for (int j = 0; j < h.Y; j++) {
                const int gx = i + h.xoff, gy = j + h.yoff;
                if (gx == -1 || gx == h.nx || gy == -1 || gy == h.ny || k == 0 || k == h.nz - 1) {
                    long long p = cell(h, k, i, j);
                    h.ct[p] = kWall;
                    h.gid[p] = 0;
                }
            }
    for (int k = 1; k < h.nz - 1; k++)
        for (int i = 0; i < h.X; i++)
            for (int j = 0; j < h.Y; j++) {
                long long p = cell(h, k, i, j);
                h.ct[p] = (h.cts[p] == 0) ? kSolid : kLiquid;
            }
    // A cell's activation depends only on its neighbours' CritTimeStep and on whether they are
    // walls, neither of which this loop changes, so the scan order does not matter.
    for (int k = 1; k < h.nz - 1; k++)
        for (int i = 0; i < h.X; i++)
            for (int j = 0; j < h.Y; j++) {
                long long p = cell(h, k, i, j);
                if (h.cts[p] != 0) continue;
                for (int n = 0; n < 26; n++) {
                    if (!in_slab(h, i, j, n)) continue;
                    long long q = cell(h, k + kNbZ[n], i + kNbX[n], j + kNbY[n]);
                    if (h.cts[q] > 0 && h.ct[q] != kWall) {
                        h.ct[p] = kActive;
                        h.substrate_active++;
                        break;
                    }
                }
            }
}

// Draws of GrainInit's potential nuclei (src/CAinitialize.cpp:1834-2206): one per interior cell
// with melting data, in (z, y, x) order, from mt19937_64(seed). first[k] is the number of draws
// before plane k (k = 1 .. nz-1); the draws themselves need nothing else, so they can run while
// the substrate is built.
struct NucleusDraws {
    std::vector<long long> first, hits;
    std::vector<uint32_t> sites;  // the cells the hits mark, in index order (filled by grains)
};
std::vector<long long> draws_before_planes(const HostDomain &h) {
    std::vector<long long> n(h.nz + 1, 0);
    for_planes(h, 1, h.nz - 1, [&](int k) {
        long long c = 0;
        for (int j = 1; j < h.Y - 1; j++)
            for (int i = 1; i < h.X - 1; i++) c += h.cts[cell(h, k, i, j)] != 0;
        n[k + 1] = c;
    });
    n[1] = 0;
    for (int k = 2; k <= h.nz; k++) n[k] += n[k - 1];
    return n;
}
// The same counts for a directional-solidification domain before its fields exist: every cell of
// plane k has CritTimeStep crit_directional(k).
std::vector<long long> draws_before_planes_directional(const Deck &d, const HostDomain &h) {
    std::vector<long long> n(h.nz + 1, 0);
    const long long interior = static_cast<long long>(h.X - 2) * (h.Y - 2);
    for (int k = 1; k < h.nz - 1; k++) n[k + 1] = crit_directional(d, k) != 0 ? interior : 0;
    for (int k = 2; k <= h.nz; k++) n[k] += n[k - 1];
    return n;
}

// GrainInit (src/CAinitialize.cpp:1834-2206), first layer of rank 0 of 1: the cells whose draw is
// below the bulk nucleation probability are potential nuclei, numbered -1, -2, ... in index
// order; then liquid cells with no non-solid, non-wall neighbour turn solid, in index order.
void grains(const Deck &d, HostDomain &h, NucleusDraws &nd) {
    // each plane's draws are the range [first[k], first[k+1]): planes on threads
    std::vector<std::vector<uint32_t>> marked(h.nz);
    for_planes(h, 1, h.nz - 1, [&](int k) {
        auto it = std::lower_bound(nd.hits.begin(), nd.hits.end(), nd.first[k]);
        if (it == nd.hits.end() || *it >= nd.first[k + 1]) return;
        long long c = nd.first[k];
        for (int j = 1; j < h.Y - 1; j++)
            for (int i = 1; i < h.X - 1; i++) {
                long long p = cell(h, k, i, j);
                if (h.cts[p] == 0) continue;
                if (it != nd.hits.end() && *it == c) {
                    h.ct[p] = kTemporaryInit;
                    marked[k].push_back(static_cast<uint32_t>(p));
                    ++it;
                }
                c++;
            }
        std::sort(marked[k].begin(), marked[k].end());  // draw order is (y, x), index order (x, y)
    });
    for (int k = 1; k < h.nz - 1; k++) nd.sites.insert(nd.sites.end(), marked[k].begin(), marked[k].end());
    // Sites are numbered -1, -2, ... in index order and the active cells listed in index order:
    // per-plane counts first, then each plane from its offset.
    const long long plane_cells = static_cast<long long>(h.X) * h.Y;
    std::vector<long long> sites(h.nz + 1, 0);
    std::vector<std::vector<uint32_t>> act(h.nz);
    for_planes(h, 1, h.nz - 1, [&](int k) {
        long long n = 0;
        for (long long p = k * plane_cells; p < (k + 1) * plane_cells; p++) {
            if (h.ct[p] == kActive) {
                if (k <= h.zhigh) act[k].push_back(static_cast<uint32_t>(p));
            }
            else if (h.ct[p] == kTemporaryInit) {
                n++;
            }
        }
        sites[k + 1] = n;
    });
    for (int k = 1; k <= h.nz; k++) sites[k] += sites[k - 1];
    for_planes(h, 1, h.nz - 1, [&](int k) {
        int id = -1 - static_cast<int>(sites[k]);
        for (long long p = k * plane_cells; p < (k + 1) * plane_cells; p++)
            if (h.ct[p] == kTemporaryInit) h.gid[p] = id--;
    });
    for (int k = 1; k < h.nz - 1; k++) h.active.insert(h.active.end(), act[k].begin(), act[k].end());
    // The index order of this pass does not matter, so planes run on threads: a liquid cell
    // turns solid only when none of its neighbours is liquid, so no neighbour of a liquid cell
    // ever changes here, and no cell reads a type that another thread writes.
    for_planes(h, 1, h.nz - 1, [&](int k) {
        for (int i = 1; i < h.X - 1; i++)
            for (int j = 1; j < h.Y - 1; j++) {
                long long p = cell(h, k, i, j);
                if (h.ct[p] != kLiquid) continue;
                bool bordered = false;
                for (int n = 0; n < 26 && !bordered; n++) {
                    uint8_t t = h.ct[cell(h, k + kNbZ[n], i + kNbX[n], j + kNbY[n])];
                    bordered = (t != kSolid && t != kWall);
                }
                if (!bordered) h.ct[p] = kSolid;
            }
    });
}

// NucleiInit (src/CAinitialize.cpp:2213-2248, 2764-2767), rank 0: one normal draw per site in
// index order from a default-seeded std::default_random_engine. The sites (the cells GrainInit
// marked) are visited from their list, which is in index order.
void nuclei(const Deck &d, HostDomain &h, const std::vector<uint32_t> &sites) {
    std::default_random_engine generator(static_cast<std::default_random_engine::result_type>(d.seed % 2147483646u + 1u));
    std::normal_distribution<double> distribution(d.dTN, d.dTsigma);
    for (const uint32_t p : sites) {
        const double und = distribution(generator);
        const int t = static_cast<int>(h.cts[p] + std::round(und / h.dU));
        h.nloc.push_back(static_cast<int32_t>(p));
        h.ntime.push_back(std::max(h.cts[p], t));
        h.ct[p] = kLiquid;
    }
}

}  // namespace

std::vector<float> read_orientations(const std::string &path) {
    std::vector<float> guv(9 * kNumOrientations, 0.0f);
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("Could not locate/open Substrate file");
    std::string line;
    std::getline(in, line);
    for (int i = 0; i < kNumOrientations; i++) {
        if (!std::getline(in, line)) break;
        std::istringstream ss(line);
        std::string tok;
        int slot = 0;
        while (slot < 9 && std::getline(ss, tok, ',')) guv[9 * i + slot++] = static_cast<float>(std::atof(tok.c_str()));
    }
    return guv;
}

void seed_critical_lengths(const float *g, float out[26]) {
    for (int n = 0; n < 26; n++) {
        const double x0 = kNbX[n], y0 = kNbY[n], z0 = kNbZ[n];
        const double mag0 = std::pow(std::pow(x0, 2.0) + std::pow(y0, 2.0) + std::pow(z0, 2.0), 0.5);
        double diag[3][3];
        for (int a = 0; a < 3; a++) {
            const double ang = (g[3 * a] * x0 + g[3 * a + 1] * y0 + g[3 * a + 2] * z0) / mag0;
            for (int c = 0; c < 3; c++) diag[a][c] = (ang < 0) ? g[3 * a + c] : -g[3 * a + c];
        }
        const double u1[3] = {diag[1][0] - diag[0][0], diag[1][1] - diag[0][1], diag[1][2] - diag[0][2]};
        const double u2[3] = {diag[2][0] - diag[0][0], diag[2][1] - diag[0][1], diag[2][2] - diag[0][2]};
        const double uu[3] = {u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2],
                              u1[0] * u2[1] - u1[1] * u2[0]};
        const double nd = std::sqrt(uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2]);
        const double nrm[3] = {uu[0] / nd, uu[1] / nd, uu[2] / nd};
        const double t = (nrm[0] * x0 + nrm[1] * y0 + nrm[2] * z0) /
                         (nrm[0] * diag[0][0] + nrm[1] * diag[0][1] + nrm[2] * diag[0][2]);
        out[n] = static_cast<float>(std::pow(
            std::pow(t * diag[0][0], 2.0) + std::pow(t * diag[0][1], 2.0) + std::pow(t * diag[0][2], 2.0), 0.5));
    }
}

std::vector<double> growth_table(double A, double B, double C, double D, float dU, int max_entries) {
    // U_k = U_{k-1} + dU in float from U_0 = 0 (src/CAupdate.cpp:236-237; TempInit sets U = 0);
    // increment = min(0.045, max(0, A U^3 + B U^2 + C U + D)) with U clamped at 210 (:257-261).
    std::vector<double> t;
    float u = 0.0f;
    for (int k = 0; k < max_entries; k++) {
        double lu = u;
        lu = (lu < 210.0) ? lu : 210.0;
        double v = A * std::pow(lu, 3.0) + B * std::pow(lu, 2.0) + C * lu + D;
        v = (0.0 < v) ? v : 0.0;
        t.push_back((v < 0.045) ? v : 0.045);
        if (lu >= 210.0) break;
        float next = u + dU;
        if (next == u) break;
        u = next;
    }
    return t;
}

int growth_plateau(const std::vector<double> &grow) {
    int k = static_cast<int>(grow.size()) - 1;
    while (k > 0 && grow[k - 1] == grow.back()) k--;
    return k < 0 ? 0 : k;
}

HostDomain initialize_sparse(const Deck &d, bool verbose) {
    if (d.sim_type != "C" || d.nx < 2 || d.ny < 2 || d.nz < 3) return initialize(d, verbose);
    HostDomain h;
    h.nx = d.nx;
    h.ny = d.ny;
    h.nz = d.nz;
    h.X = d.nx + 2;
    h.Y = d.ny + 2;
    h.xoff = -1;
    h.yoff = -1;
    h.ncell = static_cast<long long>(h.X) * h.Y * h.nz;
    if (h.ncell >= (1LL << 32) - 1) throw std::runtime_error("domain exceeds 2^32 cells");
    h.sparse = true;
    h.zlow = 0;
    h.zhigh = h.nz - 1;
    h.nzA = h.nz;
    h.dU = static_cast<float>(d.R * d.deltat);
    h.plane_cts.resize(h.nz);
    for (int k = 0; k < h.nz; k++) h.plane_cts[k] = crit_directional(d, k);
    // GrainInit's draws (one per interior cell of a plane with melting data, (z, y, x) order) on
    // their own thread while the substrate's draws run here
    NucleusDraws nd;
    nd.first = draws_before_planes_directional(d, h);
    const double bulk = d.nmax * d.deltax * d.deltax * d.deltax;
    std::thread drawer([&] { nd.hits = draws_below(d.seed, bulk, nd.first[std::max(h.nz - 1, 1)]); });
    struct JoinOnExit {
        std::thread &t;
        ~JoinOnExit() {
            if (t.joinable()) t.join();
        }
    } join_drawer{drawer};
    // SubstrateInit_ConstrainedGrowth: one draw per interior cell of plane 1, in index order
    {
        std::mt19937_64 gen(d.seed);
        std::uniform_real_distribution<double> dis(0.0, 1.0);
        for (int i = 1; i < h.X - 1; i++)
            for (int j = 1; j < h.Y - 1; j++)
                if (dis(gen) < d.fract_active) h.active.push_back(static_cast<uint32_t>(cell(h, 1, i, j)));
        h.substrate_active = static_cast<int>(h.active.size());
    }
    h.guv = read_orientations(d.orientation_file);
    drawer.join();
    // the marked cells: draw c of plane k is interior cell (i, j) with c - first[k] = (j - 1) (X - 2) + i - 1
    const long long row = h.X - 2;
    std::vector<uint32_t> sites;
    sites.reserve(nd.hits.size());
    for (size_t a = 0; a < nd.hits.size();) {
        int k = static_cast<int>(std::upper_bound(nd.first.begin(), nd.first.end(), nd.hits[a]) - nd.first.begin()) - 1;
        const size_t b0 = sites.size();
        for (; a < nd.hits.size() && nd.hits[a] < nd.first[k + 1]; a++) {
            const long long c = nd.hits[a] - nd.first[k];
            sites.push_back(static_cast<uint32_t>(cell(h, k, static_cast<int>(c % row) + 1, static_cast<int>(c / row) + 1)));
        }
        std::sort(sites.begin() + b0, sites.end());  // draw order is (y, x), index order (x, y)
    }
    // NucleiInit: one normal draw per site in index order
    std::default_random_engine generator(static_cast<std::default_random_engine::result_type>(d.seed % 2147483646u + 1u));
    std::normal_distribution<double> distribution(d.dTN, d.dTsigma);
    const long long plane_cells = static_cast<long long>(h.X) * h.Y;
    for (const uint32_t p : sites) {
        const int cts = h.plane_cts[p / plane_cells];
        const double und = distribution(generator);
        const int t = static_cast<int>(cts + std::round(und / h.dU));
        h.nloc.push_back(static_cast<int32_t>(p));
        h.ntime.push_back(std::max(cts, t));
    }
    h.A = d.A * d.deltat / d.deltax;
    h.B = d.B * d.deltat / d.deltax;
    h.C = d.C * d.deltat / d.deltax;
    h.D = d.D;
    if (verbose) {
        std::cout << "Number of substrate active cells: " << h.substrate_active << std::endl;
        std::cout << "Number of potential nucleated grains: " << h.nloc.size() << std::endl;
    }
    return h;
}

HostDomain initialize(const Deck &d, bool verbose) {
    HostDomain h;
    h.nx = d.nx;
    h.ny = d.ny;
    h.nz = d.nz;
    h.X = d.nx + 2;
    h.Y = d.ny + 2;
    h.xoff = -1;
    h.yoff = -1;
    h.ncell = static_cast<long long>(h.X) * h.Y * h.nz;
    if (h.ncell >= (1LL << 32) - 1) throw std::runtime_error("domain exceeds 2^32 cells");
    // GrainInit's draws run on their own thread from the moment their count per plane is known:
    // at once for a directional-solidification domain, after the temperature field otherwise. They
    // read only CritTimeStep (final by then); the substrate writes only CellType and GrainID.
    NucleusDraws nd;
    const double bulk = d.nmax * d.deltax * d.deltax * d.deltax;
    std::thread drawer;
    struct JoinOnExit {
        std::thread &t;
        ~JoinOnExit() {
            if (t.joinable()) t.join();
        }
    } join_drawer{drawer};
    auto start_draws = [&] {
        drawer = std::thread([&] { nd.hits = draws_below(d.seed, bulk, nd.first[std::max(h.nz - 1, 1)]); });
    };
    if (d.sim_type == "C") {
        nd.first = draws_before_planes_directional(d, h);
        start_draws();
    }
    // every field starts at zero: filled plane by plane on threads (first touch included)
    h.ct.resize(h.ncell);
    h.gid.resize(h.ncell);
    h.cts.resize(h.ncell);
    h.lid.resize(h.ncell);
    h.melted.resize(h.ncell);
    for_planes(h, 0, h.nz, [&](int k) {
        const long long a = cell(h, k, 0, 0), b = cell(h, k + 1, 0, 0);
        std::fill(h.ct.begin() + a, h.ct.begin() + b, 0);
        std::fill(h.gid.begin() + a, h.gid.begin() + b, 0);
        std::fill(h.cts.begin() + a, h.cts.begin() + b, 0);
        std::fill(h.lid.begin() + a, h.lid.begin() + b, 0);
        std::fill(h.melted.begin() + a, h.melted.begin() + b, 0);
    });
    if (d.sim_type == "C") {
        temperature_directional(d, h);
    }
    else {
        temperature_spots(d, h);
        nd.first = draws_before_planes(h);
        start_draws();
    }
    h.guv = read_orientations(d.orientation_file);
    if (d.sim_type == "C") {
        substrate_directional(d, h);
    }
    else {
        if (d.use_substrate_file)
            substrate_file(d, h);
        else
            substrate_spacing(d, h);
        walls_and_interface(h);
    }
    drawer.join();
    grains(d, h, nd);
    nuclei(d, h, nd.sites);
    h.A = d.A * d.deltat / d.deltax;
    h.B = d.B * d.deltat / d.deltax;
    h.C = d.C * d.deltat / d.deltax;
    h.D = d.D;
    if (verbose) {
        std::cout << "Number of substrate active cells: " << h.substrate_active << std::endl;
        std::cout << "Number of potential nucleated grains: " << h.nloc.size() << std::endl;
    }
    return h;
}

}  // namespace exaca
