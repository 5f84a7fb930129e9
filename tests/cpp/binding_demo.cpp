// Driver of the reference-tree binding (physicell-x_amd/host/BioFVM_reference_binding.h),
// test infrastructure.
//
// The binding is a set of templates over the reference's Microenvironment and
// agent types.  The reference itself cannot be built here (it needs MPI), so
// this driver instantiates the templates with RefMicro / RefAgent below: small
// classes with the members and the semantics of the reference's public
// interface that the binding uses (p_density_vectors, mesh coordinates,
// voxel_index, number_of_densities, diffusion_coefficients, decay_rates,
// get_substrate_dirichlet_activation, apply_dirichlet_boundaries_conditions
// over a per-voxel value store, the Basic_Agent source/sink constants and
// internalized_substrates).  They are not the reference's classes.
//
// The run: the solver through the MPI-signature function pointer type
// (Microenvironment&, double, mpi_Environment&, mpi_Cartesian&) at world size
// 1, the agents through the container-level call with internalised tracking,
// and between steps the Dirichlet state changes the reference allows
// (update_dirichlet_node on one substrate of a face voxel, a substrate's
// activation switched off) to show that every call picks them up.
//
// Case file (little endian): int32 {nx, ny, nz, ns, steps, n_agents,
// n_nodes, change_step}; float64 {dx, dt, D[ns], decay[ns], active[ns],
// field[nx*ny*nz*ns], per node: voxel, values[ns]; the change: voxel,
// substrate, value, substrate switched off; per agent: voxel, active,
// c1[ns], c2[ns], e1[ns], e2[ns]}.  Output: the field, then the agents'
// internalized totals [n_agents][ns].
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "BioFVM_reference_binding.h"

struct RefMesh {
    std::vector<double> x_coordinates, y_coordinates, z_coordinates;
    double dx = 1.0;
};

class RefMicro {
  public:
    std::vector<double> density;
    std::vector<double>* p_density_vectors = &density;
    RefMesh mesh;
    std::vector<double> diffusion_coefficients, decay_rates;
    std::vector<std::vector<double>> dirichlet_value_vectors;  // per voxel, default 1.0
    std::vector<bool> dirichlet_activation_vector;
    int nx = 0, ny = 0, nz = 0, ns = 0;

    unsigned int number_of_densities() { return (unsigned)ns; }
    int voxel_index(int i, int j, int k) { return (i * ny + j) * nz + k; }
    bool get_substrate_dirichlet_activation(int d) { return dirichlet_activation_vector[d]; }
    void set_substrate_dirichlet_activation(int d, bool v) { dirichlet_activation_vector[d] = v; }
    void add_dirichlet_node(int v, const std::vector<double>& value) { dirichlet_value_vectors[v] = value; }
    void update_dirichlet_node(int v, int d, double value) { dirichlet_value_vectors[v][d] = value; }
    // the face pass: every face voxel takes its value vector for the active
    // substrates (x faces only on the first / last rank)
    void apply_dirichlet_boundaries_conditions(int rank, int size) {
        auto put = [&](int i, int j, int k) {
            const int v = voxel_index(i, j, k);
            for (int d = 0; d < ns; ++d)
                if (dirichlet_activation_vector[d]) density[(size_t)v * ns + d] = dirichlet_value_vectors[v][d];
        };
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                if (rank == 0) put(0, j, k);
                if (rank == size - 1) put(nx - 1, j, k);
            }
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) {
                put(i, 0, k);
                put(i, ny - 1, k);
            }
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) {
                put(i, j, 0);
                put(i, j, nz - 1);
            }
    }
};

struct RefAgent {
    int current_voxel_index = 0;
    bool is_active = true;
    std::vector<double> cell_source_sink_solver_temp1, cell_source_sink_solver_temp2;
    std::vector<double> cell_source_sink_solver_temp_export1, cell_source_sink_solver_temp_export2;
    std::vector<double> internal;
    std::vector<double>* internalized_substrates = &internal;
};

struct World {
    int size = 1, rank = 0;
};
struct Cart {};

// the solver with the reference's MPI function-pointer signature
static void solver_mpi(RefMicro& M, double dt, World& w, Cart& c) { bfa_binding::lod_3d_mi355x_mpi(M, dt, w, c); }

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: binding_demo case.bin out.bin" << std::endl;
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    int32_t h[8];
    in.read(reinterpret_cast<char*>(h), sizeof h);
    const int nx = h[0], ny = h[1], nz = h[2], ns = h[3], steps = h[4], na = h[5], nn = h[6], change = h[7];
    auto rd = [&](size_t n) {
        std::vector<double> v(n);
        in.read(reinterpret_cast<char*>(v.data()), n * sizeof(double));
        return v;
    };
    const double dx = rd(1)[0], dt = rd(1)[0];
    RefMicro M;
    M.nx = nx;
    M.ny = ny;
    M.nz = nz;
    M.ns = ns;
    M.mesh.dx = dx;
    for (int i = 0; i < nx; ++i) M.mesh.x_coordinates.push_back((i + 0.5) * dx);
    for (int j = 0; j < ny; ++j) M.mesh.y_coordinates.push_back((j + 0.5) * dx);
    for (int k = 0; k < nz; ++k) M.mesh.z_coordinates.push_back((k + 0.5) * dx);
    M.diffusion_coefficients = rd(ns);
    M.decay_rates = rd(ns);
    for (double a : rd(ns)) M.dirichlet_activation_vector.push_back(a != 0.0);
    M.density = rd((size_t)nx * ny * nz * ns);
    M.dirichlet_value_vectors.assign((size_t)nx * ny * nz, std::vector<double>(ns, 1.0));
    for (int q = 0; q < nn; ++q) {
        const std::vector<double> nd = rd(1 + (size_t)ns);
        M.add_dirichlet_node((int)nd[0], std::vector<double>(nd.begin() + 1, nd.end()));
    }
    const std::vector<double> chg = rd(4);
    std::vector<RefAgent> agents(na);
    std::vector<RefAgent*> list;
    for (auto& a : agents) {
        const std::vector<double> hd = rd(2);
        a.current_voxel_index = (int)hd[0];
        a.is_active = hd[1] != 0.0;
        a.cell_source_sink_solver_temp1 = rd(ns);
        a.cell_source_sink_solver_temp2 = rd(ns);
        a.cell_source_sink_solver_temp_export1 = rd(ns);
        a.cell_source_sink_solver_temp_export2 = rd(ns);
        a.internal.assign(ns, 0.0);
        list.push_back(&a);
    }
    if (!in) {
        std::cerr << "short case file" << std::endl;
        return 2;
    }
    void (*diffusion_decay_solver_mpi)(RefMicro&, double, World&, Cart&) = solver_mpi;
    World world;
    Cart cart;
    for (int s = 0; s < steps; ++s) {
        if (s == change) {
            M.update_dirichlet_node((int)chg[0], (int)chg[1], chg[2]);
            M.set_substrate_dirichlet_activation((int)chg[3], false);
        }
        diffusion_decay_solver_mpi(M, dt, world, cart);
        bfa_binding::secretion_uptake_mi355x(M, list, true);
    }
    // at world size 2 the solver refuses (prints) and leaves the field alone
    World two;
    two.size = 2;
    const std::vector<double> before = M.density;
    diffusion_decay_solver_mpi(M, dt, two, cart);
    if (M.density != before) {
        std::cerr << "the solver changed the field at world size 2" << std::endl;
        return 3;
    }
    std::ofstream o(argv[2], std::ios::binary);
    o.write(reinterpret_cast<const char*>(M.density.data()), M.density.size() * sizeof(double));
    for (auto& a : agents) o.write(reinterpret_cast<const char*>(a.internal.data()), ns * sizeof(double));
    std::cout << "binding_demo: " << steps << " steps done" << std::endl;
    return 0;
}
