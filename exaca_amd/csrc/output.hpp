// output.hpp -- end-of-run files in the reference's formats (src/CAprint.cpp:285-571).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "deck.hpp"
#include "init.hpp"

namespace exaca {

struct RunTimes {
    double init = 0, run = 0, out = 0, nucleation = 0, capture = 0, ghost = 0;
};

// <path><base>_Misorientations.vtk and <path><base>.vtk (or _debug.vtk), from final GrainID and
// the initial LayerID / Melted fields. Returns the nucleated volume fraction that the reference
// prints to stdout.
float write_vtk_files(const Deck &d, const HostDomain &h, const std::vector<int32_t> &grain_id);

// <path><base>.log (src/CAprint.cpp:463-547).
void write_log(const Deck &d, const HostDomain &h, int cycle, const RunTimes &t);

}  // namespace exaca
