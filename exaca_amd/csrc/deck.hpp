// deck.hpp -- the reference's input-deck format (src/CAinitialize.cpp:24-576).
#pragma once
#include <cstdint>
#include <string>

namespace exaca {

struct Deck {
    std::string input_file;
    std::string sim_type;  // "C" (directional solidification) or "S" (spot-melt array)
    int decomposition = 1;
    std::string material;
    double A = 0, B = 0, C = 0, D = 0, freezing_range = 0;  // raw (unnormalised) response function
    double deltax = 0, nread = 0, nmax = 0, dTN = 0, dTsigma = 0;
    std::string path_to_output, output_base, orientation_file;
    double G = 0, R = 0, deltat = 0, fract_active = 0;
    int nx = 0, ny = 0, nz = 0, nratio = 0;
    int nspots_x = 0, nspots_y = 0, spot_offset = 0, spot_radius = 0, num_layers = 1, layer_height = 0;
    float substrate_spacing = 0;
    bool use_substrate_file = false;
    std::string substrate_file;
    bool print_misorientation = false, print_full = false;
    int print_debug = 0;
    // Random-stream seed of the initial state. 0 reproduces the reference (mt19937_64(0) for the
    // substrate and the nuclei sites, a default-seeded default_random_engine for the nucleation
    // undercooling); ensembles give each instance its own.
    uint64_t seed = 0;
};

// Directory holding Materials/ and Substrate/ (the reference's ExaCA_DATA_INSTALL /
// ExaCA_DATA_SOURCE, src/CAconfig.hpp.cmakein:9-10). EXACA_DATA_DIR overrides the built-in one.
std::string data_dir();

// Parses a deck exactly as InputReadFromFile does (ordered "key: value" lines after a "*****"
// line, keys matched by substring, whitespace stripped from values). Throws std::runtime_error
// with the reference's messages; problem type R and multilayer decks are rejected.
Deck read_deck(const std::string &path, const std::string &data_directory);

}  // namespace exaca
