// Reference-style driver of the C++ host side (test infrastructure).
//
// Reads a case file written by tests/test_gpu_adapter.py, sets the
// microenvironment up the way initialize_microenvironment does (densities,
// uniform mesh, initial field, Dirichlet faces, activation), registers
// agents, and runs the main-loop pair of the reference (main.cpp:244 then
// main.cpp:259): simulate_diffusion_decay(dt) then the agents' secretion and
// uptake.  Writes the final field (reference layout) and the total mass of
// every substrate after every step, then the gradient field of the final
// field (compute_all_gradient_vectors, [voxel][ns][3]), then the field after
// one call of simulate_bulk_sources_and_sinks with the rate callbacks below.
//
// Case file (little endian): int32 {nx, ny, nz, ns, steps, n_agents,
// precision}; float64 {dx, dt, D[ns], decay[ns], active[ns], registered[6],
// face_values[6*ns], field[nx*ny*nz*ns], then per agent: position[3],
// volume, secretion[ns], uptake[ns], export[ns], saturation[ns]}; then
// optionally int32 {n_nodes, n_single, solver (0 BLOCKING, 1 AVX256D),
// track_internalized} and float64 {per node: voxel, values[ns]; per single
// update: voxel, substrate, value}: Dirichlet nodes added after the faces
// (add_dirichlet_node, then update_dirichlet_node for one substrate).  With
// tracking on, the agents' internalized_substrates [n_agents][ns] follow the
// other outputs.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "BioFVM_amd.h"

using namespace BioFVM_amd;

// bulk rate callbacks: a function of the voxel index and the substrate only
static void supply(Microenvironment* m, int n, double* w) {
    for (int d = 0; d < m->number_of_densities(); ++d) w[d] = n % 3 == 0 ? 0.5 * (d + 1) : 0.0;
}
static void target(Microenvironment* m, int, double* w) {
    for (int d = 0; d < m->number_of_densities(); ++d) w[d] = 30.0 + d;
}
static void uptake(Microenvironment* m, int n, double* w) {
    for (int d = 0; d < m->number_of_densities(); ++d) w[d] = n % 5 == 0 ? 0.2 : 0.0;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: adapter_demo case.bin out.bin" << std::endl;
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    int32_t hdr[7];
    in.read(reinterpret_cast<char*>(hdr), sizeof hdr);
    const int nx = hdr[0], ny = hdr[1], nz = hdr[2], ns = hdr[3], steps = hdr[4], na = hdr[5];
    auto rd = [&](size_t n) {
        std::vector<double> v(n);
        in.read(reinterpret_cast<char*>(v.data()), n * sizeof(double));
        return v;
    };
    const double dx = rd(1)[0], dt = rd(1)[0];
    std::vector<double> D = rd(ns), decay = rd(ns), active = rd(ns), reg = rd(6), fv = rd(6 * (size_t)ns);
    std::vector<double> field = rd((size_t)nx * ny * nz * ns);
    if (!in) {
        std::cerr << "short case file" << std::endl;
        return 2;
    }

    Microenvironment M;
    M.precision = hdr[6];
    // as initialize_microenvironment: density 0 is set, the others added
    M.set_density(0, "oxygen", "mmHg");
    M.diffusion_coefficients[0] = D[0];
    M.decay_rates[0] = decay[0];
    for (int d = 1; d < ns; ++d) M.add_density("s" + std::to_string(d), "dimensionless", D[d], decay[d]);
    M.resize_space_uniform(-0.5 * dx * nx, 0.5 * dx * nx, -0.5 * dx * ny, 0.5 * dx * ny,
                           -0.5 * dx * nz, 0.5 * dx * nz, dx);
    std::vector<double>& u = M.density_field();
    if (u.size() != field.size()) {
        std::cerr << "size mismatch " << u.size() << " vs " << field.size() << std::endl;
        return 2;
    }
    u = field;
    for (int f = 0; f < 6; ++f)
        if (reg[f] != 0.0) M.add_dirichlet_face(f, std::vector<double>(fv.begin() + f * ns, fv.begin() + (f + 1) * ns));
    for (int d = 0; d < ns; ++d) M.set_substrate_dirichlet_activation(d, active[d] != 0.0);

    std::vector<Basic_Agent> agents(na);
    std::vector<Basic_Agent*> list;
    for (int a = 0; a < na; ++a) {
        Basic_Agent& ag = agents[a];
        ag.register_microenvironment(&M);
        std::vector<double> p = rd(3);
        ag.position = p;
        ag.volume = rd(1)[0];
        ag.secretion_rates = rd(ns);
        ag.uptake_rates = rd(ns);
        ag.net_export_rates = rd(ns);
        ag.saturation_densities = rd(ns);
        ag.update_voxel_index();
        list.push_back(&ag);
    }
    int32_t extra[4] = {0, 0, 0, 0};
    in.read(reinterpret_cast<char*>(extra), sizeof extra);
    if (!in) extra[0] = extra[1] = extra[2] = extra[3] = 0;
    in.clear();
    for (int q = 0; q < extra[0]; ++q) {
        const std::vector<double> nd = rd(1 + (size_t)ns);
        M.add_dirichlet_node((int)nd[0], std::vector<double>(nd.begin() + 1, nd.end()));
    }
    for (int q = 0; q < extra[1]; ++q) {
        const std::vector<double> up = rd(3);
        M.update_dirichlet_node((int)up[0], (int)up[1], up[2]);
    }
    if (extra[2] == 1) M.diffusion_decay_solver = diffusion_decay_solver__constant_coefficients_LOD_3D_AVX256D_GPU;
    M.track_internalized_substrates_in_each_agent = extra[3] != 0;

    std::vector<double> masses;
    for (int s = 0; s < steps; ++s) {
        M.simulate_diffusion_decay(dt);
        M.simulate_cell_sources_and_sinks(list, dt);
        for (int d = 0; d < ns; ++d) masses.push_back(M.total_mass(d));
    }
    const std::vector<double> out = M.density_field();
    M.compute_all_gradient_vectors();
    std::vector<double> grad((size_t)M.number_of_voxels() * ns * 3);
    for (int n = 0; n < M.number_of_voxels(); ++n) {
        const double* g = M.gradient_vector(n);
        std::copy(g, g + 3 * ns, grad.begin() + (size_t)n * ns * 3);
    }
    M.bulk_supply_rate_function = supply;
    M.bulk_supply_target_densities_function = target;
    M.bulk_uptake_rate_function = uptake;
    M.simulate_bulk_sources_and_sinks(dt);
    const std::vector<double>& bulk = M.density_field();
    std::ofstream o(argv[2], std::ios::binary);
    o.write(reinterpret_cast<const char*>(out.data()), out.size() * sizeof(double));
    o.write(reinterpret_cast<const char*>(masses.data()), masses.size() * sizeof(double));
    o.write(reinterpret_cast<const char*>(grad.data()), grad.size() * sizeof(double));
    o.write(reinterpret_cast<const char*>(bulk.data()), bulk.size() * sizeof(double));
    if (extra[3])
        for (const Basic_Agent& ag : agents)
            o.write(reinterpret_cast<const char*>(ag.internalized_substrates.data()), ns * sizeof(double));
    std::cout << "adapter_demo: " << steps << " steps of " << nx << "x" << ny << "x" << nz << "x" << ns
              << " with " << na << " agents done" << std::endl;
    return 0;
}
