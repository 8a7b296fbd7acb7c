// deck.cpp -- input-deck reader following InputReadFromFile (src/CAinitialize.cpp:232-576).
#include "deck.hpp"

#include <cmath>
#include <cstdlib>
#include <dlfcn.h>
#include <fstream>
#include <stdexcept>

namespace exaca {
namespace {

// skipLines: everything up to and including the "*****" line is a free-form header
// (src/CAinitialize.cpp:24-30).
void skip_header(std::ifstream &in) {
    std::string line;
    while (std::getline(in, line))
        if (line == "*****") break;
}

std::string strip_spaces(const std::string &s) {
    std::string out;
    for (char c : s)
        if (!(c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f')) out.push_back(c);
    return out;
}

// getKey (src/CAinitialize.cpp:36-62). The reference's deprecated-key loop ends as soon as one
// deprecated name is absent from the key, so every key is returned as read.
std::string next_key(std::ifstream &in, std::string &line, std::size_t &colon) {
    std::getline(in, line);
    colon = line.find(':');
    std::string key = line.substr(0, colon);
    if (colon == std::string::npos)
        throw std::runtime_error("Input \"" + key + "\" must be separated from value by \":\".");
    return key;
}

std::string value_for(std::ifstream &in, const std::string &want) {
    std::string line;
    std::size_t colon;
    std::string key = next_key(in, line, colon);
    if (key.find(want) == std::string::npos)
        throw std::runtime_error("Required input not present: \"" + want + "\" not found in the input file");
    return strip_spaces(line.substr(colon + 1));
}

std::string value_for_either(std::ifstream &in, const std::string &k1, const std::string &k2, int &which) {
    std::string line;
    std::size_t colon;
    std::string key = next_key(in, line, colon);
    if (key.find(k1) != std::string::npos)
        which = 1;
    else if (key.find(k2) != std::string::npos)
        which = 2;
    else
        throw std::runtime_error("Required input not present: Neither " + k1 + " nor " + k2 +
                                 " was found in the input file");
    return strip_spaces(line.substr(colon + 1));
}

bool yes_no(std::ifstream &in, const std::string &want) {
    std::string v = value_for(in, want);
    if (v == "N") return false;
    if (v == "Y") return true;
    throw std::runtime_error("Input \"" + want + "\" must be \"Y\" or \"N\".");
}

int to_int(const std::string &v) { return std::stoi(v, nullptr, 10); }

void require_file(const std::string &path, const std::string &type) {
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Could not locate/open " + type + " file");
}

}  // namespace

std::string data_dir() {
    if (const char *env = std::getenv("EXACA_DATA_DIR")) return env;
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&data_dir), &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        std::size_t slash = so.rfind('/');
        return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/data";
    }
    return "data";
}

Deck read_deck(const std::string &path, const std::string &dd) {
    Deck d;
    d.input_file = path;
    require_file(path, "Input");
    std::ifstream in(path);
    skip_header(in);
    d.sim_type = value_for(in, "Problem type");
    d.decomposition = to_int(value_for(in, "Decomposition strategy"));
    d.material = value_for(in, "Material");
    {
        std::string mpath = dd + "/Materials/" + d.material;
        require_file(mpath, "Materials");
        std::ifstream mat(mpath);
        skip_header(mat);
        d.A = std::atof(value_for(mat, "A").c_str());
        d.B = std::atof(value_for(mat, "B").c_str());
        d.C = std::atof(value_for(mat, "C").c_str());
        d.D = std::atof(value_for(mat, "D").c_str());
        d.freezing_range = std::atof(value_for(mat, "Alloy freezing range").c_str());
    }
    d.deltax = std::atof(value_for(in, "Cell size").c_str()) * std::pow(10, -6);
    d.nread = std::atof(value_for(in, "Heterogeneous nucleation density").c_str());
    d.nmax = d.nread * std::pow(10, 12);
    d.dTN = std::atof(value_for(in, "Mean nucleation undercooling").c_str());
    d.dTsigma = std::atof(value_for(in, "Standard deviation of nucleation undercooling").c_str());
    d.path_to_output = value_for(in, "Path to output");
    d.output_base = value_for(in, "Output file base name");
    d.orientation_file = dd + "/Substrate/" + value_for(in, "File of grain orientations");
    require_file(d.orientation_file, "Substrate");

    if (d.sim_type == "C") {
        d.num_layers = 1;
        d.G = std::atof(value_for(in, "Thermal gradient").c_str());
        d.R = std::atof(value_for(in, "Cooling rate").c_str());
        d.nx = to_int(value_for(in, "Domain size in x")) + 2;
        d.ny = to_int(value_for(in, "Domain size in y")) + 2;
        d.nz = to_int(value_for(in, "Domain size in z")) + 2;
        d.layer_height = d.nz;
        d.nratio = to_int(value_for(in, "\"N\""));
        d.deltat = d.deltax / (d.nratio * (d.R / d.G));
        d.fract_active = std::atof(value_for(in, "Fraction of surface sites active").c_str());
    }
    else if (d.sim_type == "S") {
        d.G = std::atof(value_for(in, "Thermal gradient").c_str());
        d.R = std::atof(value_for(in, "Cooling rate").c_str());
        d.nspots_x = to_int(value_for(in, "Number of spots in x"));
        d.nspots_y = to_int(value_for(in, "Number of spots in y"));
        float offset_um = static_cast<float>(std::atof(value_for(in, "Offset between spot centers").c_str()));
        d.spot_offset = static_cast<int>(offset_um * 1.0 * std::pow(10, -6) / d.deltax);
        float radius_um = static_cast<float>(std::atof(value_for(in, "Radii of spots").c_str()));
        d.spot_radius = static_cast<int>(radius_um * 1.0 * std::pow(10, -6) / d.deltax);
        d.num_layers = to_int(value_for(in, "Number of layers"));
        d.layer_height = to_int(value_for(in, "Offset between layers"));
        d.nratio = to_int(value_for(in, "\"N\""));
        d.deltat = d.deltax / (d.nratio * (d.R / d.G));
        int which = 0;
        std::string v = value_for_either(in, "Path to substrate file", "Substrate grain spacing", which);
        if (which == 1) {
            d.substrate_file = v + "/" + value_for(in, "Substrate file name");
            d.use_substrate_file = true;
            require_file(d.substrate_file, "Substrate");
        }
        else {
            d.substrate_spacing = static_cast<float>(std::atof(v.c_str()));
            d.use_substrate_file = false;
        }
        d.nz = d.spot_radius + 3 + (d.num_layers - 1) * d.layer_height;
        d.nx = 4 + 2 * d.spot_radius + d.spot_offset * (d.nspots_x - 1);
        d.ny = 4 + 2 * d.spot_radius + d.spot_offset * (d.nspots_y - 1);
    }
    else if (d.sim_type == "R") {
        throw std::runtime_error("problem type R (temperature data from files) is not provided by this build");
    }
    else {
        throw std::runtime_error("unknown problem type \"" + d.sim_type + "\"");
    }

    int which = 0;
    std::string v = value_for_either(in, "Print file of grain orientations", "Output data printing options", which);
    if (which == 1) {
        bool old[6] = {v == "Y", false, false, false, false, false};
        old[1] = yes_no(in, "Print csv file of grain id values");
        old[2] = yes_no(in, "Print csv file of ExaConstit-formatted grain id values");
        old[3] = yes_no(in, "Print Paraview vtk file of grain misorientation values");
        old[4] = yes_no(in, "Print file of grain area values");
        old[5] = yes_no(in, "Print file of weighted grain area value");
        d.print_debug = 0;
        d.print_misorientation = old[3];
        d.print_full = (old[0] + old[1] + old[2] + old[4] + old[5]) > 0;
    }
    else {
        d.print_misorientation = yes_no(in, "Paraview vtk file of grain misorientation");
        d.print_full = yes_no(in, "Paraview vtk file of all ExaCA data");
        bool a = yes_no(in, "Print data for main Kokkos views");
        bool b = yes_no(in, "Print data for all main data structures");
        d.print_debug = b ? 2 : (a ? 1 : 0);
    }
    if (d.num_layers < 1 || d.num_layers > 255)
        throw std::runtime_error("number of layers must be 1 .. 255 (" + std::to_string(d.num_layers) + ")");
    return d;
}

}  // namespace exaca
