// main.cpp -- command-line entry with the reference executable's contract (src/main.cpp:15-47):
// one argument, the input deck; the whole program runs on the current GPU.
#include <iostream>

#include "../../include/exaca_ca.h"

int main(int argc, char *argv[]) {
    std::cout << "Number of MPI ranks = 1" << std::endl;
    if (argc < 2) {
        std::cerr << "Error: Must provide path to input file on the command line." << std::endl;
        return 1;
    }
    int rc = exaca_run_program(0, 1, argv[1]);
    if (rc != 0) {
        std::cerr << "exaca: " << exaca_last_error() << std::endl;
        return rc;
    }
    return 0;
}
