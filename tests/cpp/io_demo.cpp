// Initial condition from a MATLAB level-4 or CSV file and the field written
// back as a MATLAB file (test infrastructure, no GPU work): sets up a uniform
// mesh centred on the origin with densities s0, s1, ..., runs
// load_initial_conditions_from_matlab (or _from_csv for a .csv input), then
// Microenvironment::write_to_matlab.
//   io_demo nx ny nz ns dx in.mat|in.csv out.mat
// Exit status 0 on success, 1 when the initial condition is rejected.
#include <cstdlib>
#include <iostream>
#include <string>

#include "BioFVM_amd.h"

using namespace BioFVM_amd;

int main(int argc, char** argv) {
    if (argc != 8) {
        std::cerr << "usage: io_demo nx ny nz ns dx in.mat out.mat" << std::endl;
        return 2;
    }
    const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]),
              ns = std::atoi(argv[4]);
    const double dx = std::atof(argv[5]);
    Microenvironment M;
    M.set_density(0, "s0", "dimensionless");
    for (int d = 1; d < ns; ++d) M.add_density("s" + std::to_string(d), "dimensionless", 1.0, 0.0);
    M.resize_space_uniform(-0.5 * dx * nx, 0.5 * dx * nx, -0.5 * dx * ny, 0.5 * dx * ny,
                           -0.5 * dx * nz, 0.5 * dx * nz, dx);
    const std::string in = argv[6];
    const bool csv = in.size() > 4 && in.compare(in.size() - 4, 4, ".csv") == 0;
    if (!(csv ? load_initial_conditions_from_csv(M, in) : load_initial_conditions_from_matlab(M, in))) return 1;
    M.write_to_matlab(argv[7]);
    return 0;
}
