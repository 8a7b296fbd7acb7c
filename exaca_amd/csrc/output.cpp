// output.cpp -- vtk and log writers with the reference's layout (src/CAprint.cpp:285-571), np = 1.
// The whole-domain views of the reference hold the rank's interior (ghost ring dropped) for
// z = 1 .. nz-2 and zero on the bottom and top planes (src/CAprint.cpp:18-112).
#include "output.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <system_error>
#include <thread>

#include "ca_common.hpp"

namespace exaca {
namespace {

template <typename V>
int gathered(const HostDomain &h, const V &f, int k, int i, int j) {
    if (k < 1 || k > h.nz - 2) return 0;
    return static_cast<int>(f[(static_cast<long long>(k) * h.X + (i + 1)) * h.Y + (j + 1)]);
}

// "<v> " at p; returns the end. At most 12 bytes.
inline char *put_num(char *p, int v) {
    char tmp[12];
    int L = 0;
    unsigned u = v < 0 ? 0u - static_cast<unsigned>(v) : static_cast<unsigned>(v);
    do {
        tmp[L++] = static_cast<char>('0' + u % 10);
        u /= 10;
    } while (u);
    if (v < 0) *p++ = '-';
    while (L) *p++ = tmp[--L];
    *p++ = ' ';
    return p;
}

// Buffered writer: ASCII vtk of a large domain is dominated by formatting, so integers are
// formatted by hand into a large buffer.
struct Out {
    FILE *f;
    std::vector<char> buf;
    size_t n = 0;
    explicit Out(const std::string &path) : f(std::fopen(path.c_str(), "w")), buf(1 << 22) {
        if (!f) throw std::runtime_error("cannot open " + path + " for writing");
    }
    ~Out() {
        flush();
        std::fclose(f);
    }
    void flush() {
        if (n) std::fwrite(buf.data(), 1, n, f);
        n = 0;
    }
    void str(const std::string &s) {
        if (n + s.size() > buf.size()) flush();
        if (s.size() > buf.size()) {
            std::fwrite(s.data(), 1, s.size(), f);
            return;
        }
        std::copy(s.begin(), s.end(), buf.begin() + n);
        n += s.size();
    }
    void bytes(const char *p, size_t len) {
        flush();
        std::fwrite(p, 1, len, f);
    }
};

// Text buffers of the plane formatter, one per worker, kept for every field of a run so that
// their pages are touched once.
struct PlaneBuffers {
    int nt = static_cast<int>(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    size_t size = 0;
    std::vector<std::unique_ptr<char[]>> buf;  // left uninitialised: only written bytes are read
    char *get(int t, size_t need) {
        if (need > size) {
            buf.clear();
            for (int i = 0; i < nt; i++) buf.emplace_back(new char[need]);
            size = need;
        }
        return buf[t].get();
    }
};

// Writes planes k0 .. k1-1 of a field, one line per plane of `cols` numbers. Up to 16 workers
// format the planes of a batch into their own buffers at once (about 4 MB of text each, at least
// one plane), and the buffers go to the file in plane order. Each plane's numbers come from
// plane(k, put), put(v) appending "<v> ". Returns the sums over planes of the two counters plane()
// returns (the misorientation writer's melted and nucleated cells).
using Counts = std::array<long long, 2>;
template <typename F>
Counts write_planes(Out &o, PlaneBuffers &pb, int k0, int k1, long long cols, F plane) {
    const int nk = k1 - k0;
    if (nk <= 0) return {0, 0};
    const size_t per_plane = static_cast<size_t>(cols) * 12 + 1;
    const int nt = std::min(pb.nt, nk);
    const int pp = std::min(static_cast<int>(std::max<size_t>(1, (4u << 20) / per_plane)), (nk + nt - 1) / nt);
    for (int t = 0; t < nt; t++) pb.get(t, per_plane * pp);
    std::vector<size_t> len(nt);
    std::vector<Counts> acc(nt);
    Counts total{0, 0};
    for (int b0 = k0; b0 < k1; b0 += nt * pp) {
        auto work = [&](int t) {
            char *const p0 = pb.buf[t].get();
            char *p = p0;
            Counts a{0, 0};
            const int s = b0 + t * pp, e = std::min(k1, s + pp);
            for (int k = s; k < e; k++) {
                const Counts c = plane(k, [&p](int v) { p = put_num(p, v); });
                a[0] += c[0];
                a[1] += c[1];
                *p++ = '\n';
            }
            len[t] = static_cast<size_t>(p - p0);
            acc[t] = a;
        };
        std::vector<std::thread> th;
        int started = 1;
        try {
            for (; started < nt; started++) th.emplace_back(work, started);
        } catch (const std::system_error &) {
            // fewer threads than asked for: this thread formats the rest
        }
        work(0);
        for (int t = started; t < nt; t++) work(t);
        for (auto &x : th) x.join();
        for (int t = 0; t < nt; t++) {
            o.bytes(pb.buf[t].get(), len[t]);
            total[0] += acc[t][0];
            total[1] += acc[t][1];
        }
    }
    return total;
}

std::string header(int nx, int ny, int nz) {
    std::ostringstream o;
    o << "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET STRUCTURED_POINTS\n";
    o << "DIMENSIONS " << nx << " " << ny << " " << nz << "\nORIGIN 0 0 0\nSPACING 1 1 1\n";
    o << "POINT_DATA " << nx * ny * nz << "\n";
    return o.str();
}

int orientation_index(int gid) {
    int k = (std::abs(gid) - 1) % kNumOrientations;
    return k < 0 ? 0 : k;  // a GrainID of 0 never reaches here in a completed run
}

}  // namespace

float write_vtk_files(const Deck &d, const HostDomain &h, const std::vector<int32_t> &gid) {
    float fraction = 0.0f;
    PlaneBuffers pb;
    if (d.print_misorientation) {
        std::string fname = d.path_to_output + d.output_base + "_Misorientations.vtk";
        std::cout << "Printing Paraview file of grain misorientations" << std::endl;
        Out o(fname);
        o.str(header(h.nx - 2, h.ny - 2, h.nz - 2));
        o.str("SCALARS Angle_z int 1\nLOOKUP_TABLE default\n");
        std::vector<int> rounded(kNumOrientations);
        for (int n = 0; n < kNumOrientations; n++) {
            double amin = 62.7;
            for (int ll = 0; ll < 3; ll++) {
                // the reference's unqualified acos of a float is the C double function
                double a = std::abs((180 / M_PI) * std::acos(static_cast<double>(h.guv[9 * n + 3 * ll + 2])));
                if (a < amin) amin = a;
            }
            rounded[n] = static_cast<int>(std::round(amin));
        }
        const Counts mn = write_planes(o, pb, 1, h.nz - 1, static_cast<long long>(h.nx - 2) * (h.ny - 2),
                                              [&](int k, auto put) {
            long long m = 0, nuc = 0;
            for (int j = 1; j < h.ny - 1; j++)
                for (int i = 1; i < h.nx - 1; i++) {
                    if (gathered(h, h.melted, k, i, j) == 0) {
                        put(200);
                        continue;
                    }
                    m++;
                    int g = gathered(h, gid, k, i, j);
                    int a = rounded[orientation_index(g)];
                    if (g < 0) {
                        a += 100;
                        nuc++;
                    }
                    put(a);
                }
            return Counts{m, nuc};
        });
        const long long melted = mn[0], nucleated = mn[1];
        fraction = static_cast<float>(nucleated) / static_cast<float>(melted);
        std::cout << "Volume fraction of solidified portion of domain claimed by nucleated grains: " << fraction
                  << std::endl;
    }
    if (d.print_misorientation || d.print_full) {
        std::string fname = d.path_to_output + d.output_base + (d.print_full ? ".vtk" : "_debug.vtk");
        Out o(fname);
        o.str(header(h.nx, h.ny, h.nz));
        // Without full output the reference writes only LayerID from a view it never gathered
        // (empty); zeros are written in its place.
        o.str("SCALARS LayerID int 1\nLOOKUP_TABLE default\n");
        const long long cols = static_cast<long long>(h.nx) * h.ny;
        write_planes(o, pb, 0, h.nz, cols, [&](int k, auto put) {
            for (int j = 0; j < h.ny; j++)
                for (int i = 0; i < h.nx; i++) put(d.print_full ? gathered(h, h.lid, k, i, j) : 0);
            return Counts{0, 0};
        });
        if (d.print_full) {
            o.str("SCALARS GrainID int 1\nLOOKUP_TABLE default\n");
            write_planes(o, pb, 0, h.nz, cols, [&](int k, auto put) {
                for (int j = 0; j < h.ny; j++)
                    for (int i = 0; i < h.nx; i++) put(gathered(h, gid, k, i, j));
                return Counts{0, 0};
            });
            o.str("SCALARS Melted int 1\nLOOKUP_TABLE default\n");
            write_planes(o, pb, 0, h.nz, cols, [&](int k, auto put) {
                for (int j = 0; j < h.ny; j++)
                    for (int i = 0; i < h.nx; i++) put(gathered(h, h.melted, k, i, j) ? 1 : 0);
                return Counts{0, 0};
            });
        }
    }
    return fraction;
}

void write_log(const Deck &d, const HostDomain &h, int cycle, const RunTimes &t) {
    std::string fname = d.path_to_output + d.output_base + ".log";
    std::ofstream L(fname);
    L << "log file for a simulation run with input file " << d.input_file << "  run on 1"
      << " MPI ranks, output written at cycle " << cycle << std::endl;
    L << "This simulation took " << t.init + t.run + t.out << " seconds to run, with the init/run/output breakdown as "
      << t.init << "/" << t.run << "/" << t.out << std::endl;
    L << "This simulation was type: " << d.sim_type << std::endl;
    L << "Domain size in x: " << d.nx << std::endl;
    L << "Domain size in y: " << d.ny << std::endl;
    L << "Domain size in z: " << d.nz << std::endl;
    L << "Cell size: " << d.deltax << " microns" << std::endl;
    L << "Time step: " << d.deltat << " microseconds" << std::endl;
    L << "Nucleation density was " << d.nmax << " m^-3 , mean nucleation undercooling was " << d.dTN
      << " K, and standard deviation of nucleation undercooling was " << d.dTsigma << " K" << std::endl;
    L << "Interfacial response function parameters used were " << h.A << " , " << h.B << " , " << h.C << " , " << h.D
      << " , and the alloy freezing range was " << d.freezing_range << std::endl;
    if (d.sim_type == "C") {
        L << "The thermal gradient was " << d.G << " K/m, and the cooling rate " << d.R << " K/s" << std::endl;
        L << "The fraction of surface sites active was " << d.fract_active << std::endl;
    }
    else {
        L << "A total of " << d.nspots_x << " in X and " << d.nspots_y << " in Y were considered" << std::endl;
        L << "The spots were offset by " << d.spot_offset << " microns, and had radii of " << d.spot_radius
          << " microns" << std::endl;
        L << "This pattern was repeated for " << d.num_layers << " layers" << std::endl;
    }
    L << "The decomposition scheme used was: " << d.decomposition << std::endl;
    L << "Rank 0 contained " << h.X << " cells in x , " << h.Y << " cells in y; subdomain was offset by " << h.xoff
      << " in x , " << h.yoff << " in y" << std::endl;
    L << "Max/min rank time initializing data  = " << t.init << " / " << t.init << " s" << std::endl;
    L << "Max/min rank time in CA nucleation   = " << t.nucleation << " / " << t.nucleation << " s" << std::endl;
    L << "Max/min rank time in CA cell capture = " << t.capture << " / " << t.capture << " s" << std::endl;
    L << "Max/min rank time in CA ghosting     = " << t.ghost << " / " << t.ghost << " s" << std::endl;
    L << "Max/min rank time exporting data     = " << t.out << " / " << t.out << " s\n" << std::endl;
}

}  // namespace exaca
