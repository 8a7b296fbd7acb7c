// output.cpp -- vtk and log writers with the reference's layout (src/CAprint.cpp:285-571), np = 1.
// The whole-domain views of the reference hold the rank's interior (ghost ring dropped) for
// z = 1 .. nz-2 and zero on the bottom and top planes (src/CAprint.cpp:18-112).
#include "output.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "ca_common.hpp"

namespace exaca {
namespace {

template <typename T>
int gathered(const HostDomain &h, const std::vector<T> &f, int k, int i, int j) {
    if (k < 1 || k > h.nz - 2) return 0;
    return static_cast<int>(f[(static_cast<long long>(k) * h.X + (i + 1)) * h.Y + (j + 1)]);
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
    void num_space(int v) {  // "<v> "
        if (n + 16 > buf.size()) flush();
        char tmp[16];
        int L = 0;
        unsigned u = v < 0 ? 0u - static_cast<unsigned>(v) : static_cast<unsigned>(v);
        do {
            tmp[L++] = static_cast<char>('0' + u % 10);
            u /= 10;
        } while (u);
        if (v < 0) buf[n++] = '-';
        while (L) buf[n++] = tmp[--L];
        buf[n++] = ' ';
    }
    void nl() {
        if (n + 1 > buf.size()) flush();
        buf[n++] = '\n';
    }
};

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
        long long nucleated = 0, melted = 0;
        for (int k = 1; k < h.nz - 1; k++) {
            for (int j = 1; j < h.ny - 1; j++)
                for (int i = 1; i < h.nx - 1; i++) {
                    if (gathered(h, h.melted, k, i, j) == 0) {
                        o.num_space(200);
                        continue;
                    }
                    melted++;
                    int g = gathered(h, gid, k, i, j);
                    int a = rounded[orientation_index(g)];
                    if (g < 0) {
                        a += 100;
                        nucleated++;
                    }
                    o.num_space(a);
                }
            o.nl();
        }
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
        for (int k = 0; k < h.nz; k++) {
            for (int j = 0; j < h.ny; j++)
                for (int i = 0; i < h.nx; i++) o.num_space(d.print_full ? gathered(h, h.lid, k, i, j) : 0);
            o.nl();
        }
        if (d.print_full) {
            o.str("SCALARS GrainID int 1\nLOOKUP_TABLE default\n");
            for (int k = 0; k < h.nz; k++) {
                for (int j = 0; j < h.ny; j++)
                    for (int i = 0; i < h.nx; i++) o.num_space(gathered(h, gid, k, i, j));
                o.nl();
            }
            o.str("SCALARS Melted int 1\nLOOKUP_TABLE default\n");
            for (int k = 0; k < h.nz; k++) {
                for (int j = 0; j < h.ny; j++)
                    for (int i = 0; i < h.nx; i++) o.num_space(gathered(h, h.melted, k, i, j) ? 1 : 0);
                o.nl();
            }
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
