// BioFVM_amd.cpp -- C++ host side over the C ABI (see BioFVM_amd.h).
#include "BioFVM_amd.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

namespace BioFVM_amd {

// ---- mesh --------------------------------------------------------------------------

void Cartesian_Mesh::resize(double x_start, double x_end, double y_start, double y_end,
                            double z_start, double z_end, double dx_new, double dy_new,
                            double dz_new) {
    dx = dx_new;
    dy = dy_new;
    dz = dz_new;
    const double eps = 1e-16;  // BioFVM/BioFVM_mesh.cpp:1467-1473
    x_nodes = (int)std::ceil(eps + (x_end - x_start) / dx);
    y_nodes = (int)std::ceil(eps + (y_end - y_start) / dy);
    z_nodes = (int)std::ceil(eps + (z_end - z_start) / dz);
    const double tol = 1e-16;
    uniform_mesh = !(std::fabs(dx - dy) > tol || std::fabs(dy - dz) > tol || std::fabs(dx - dz) > tol);
    bounding_box[0] = x_start;
    bounding_box[1] = y_start;
    bounding_box[2] = z_start;
    bounding_box[3] = x_end;
    bounding_box[4] = y_end;
    bounding_box[5] = z_end;
    dV = dx * dy * dz;
}

bool Cartesian_Mesh::is_position_valid(double x, double y, double z) const {
    return !(x < bounding_box[0] || x > bounding_box[3] || y < bounding_box[1] ||
             y > bounding_box[4] || z < bounding_box[2] || z > bounding_box[5]);
}

int Cartesian_Mesh::nearest_voxel_index(const std::vector<double>& position) const {
    // floor of the offset from the box corner, truncated into the grid
    // (BioFVM/BioFVM_mesh.cpp:1825-1843)
    unsigned int i = (unsigned int)std::floor((position[0] - bounding_box[0]) / dx);
    unsigned int j = (unsigned int)std::floor((position[1] - bounding_box[1]) / dy);
    unsigned int k = (unsigned int)std::floor((position[2] - bounding_box[2]) / dz);
    if (i >= (unsigned)x_nodes) i = x_nodes - 1;
    if (j >= (unsigned)y_nodes) j = y_nodes - 1;
    if (k >= (unsigned)z_nodes) k = z_nodes - 1;
    return (int)((i * (unsigned)y_nodes + j) * (unsigned)z_nodes + k);
}

// ---- agents --------------------------------------------------------------------------

void Basic_Agent::register_microenvironment(Microenvironment* m) {
    microenvironment = m;
    const size_t ns = (size_t)m->number_of_densities();
    secretion_rates.resize(ns, 0.0);
    uptake_rates.resize(ns, 0.0);
    saturation_densities.resize(ns, 0.0);
    net_export_rates.resize(ns, 0.0);
    volume_is_changed = true;
}

void Basic_Agent::update_voxel_index() {
    const Cartesian_Mesh& mesh = microenvironment->mesh;
    if (!mesh.is_position_valid(position[0], position[1], position[2])) {
        current_voxel_index = -1;
        is_active = false;
        return;
    }
    current_voxel_index = mesh.nearest_voxel_index(position);
}

void Basic_Agent::set_internal_uptake_constants(double dt) {
    // dp/dt = S (T - p) - U p, implicit in p; the net export is explicit.
    const double W = microenvironment->mesh.dV;
    const double f = dt * volume / W;
    const size_t ns = secretion_rates.size();
    cell_source_sink_solver_temp1.assign(ns, 0.0);
    cell_source_sink_solver_temp2.assign(ns, 1.0);
    cell_source_sink_solver_temp_export1.assign(ns, 0.0);
    cell_source_sink_solver_temp_export2.assign(ns, 0.0);
    for (size_t d = 0; d < ns; ++d) {
        double t1 = secretion_rates[d];
        t1 *= saturation_densities[d];
        t1 *= f;
        cell_source_sink_solver_temp1[d] = t1;
        double t2 = std::fma(f, secretion_rates[d], 1.0);
        cell_source_sink_solver_temp2[d] = std::fma(f, uptake_rates[d], t2);
        const double e1 = net_export_rates[d] * dt;
        cell_source_sink_solver_temp_export1[d] = e1;
        cell_source_sink_solver_temp_export2[d] = e1 / W;
    }
    volume_is_changed = false;
}

void Basic_Agent::simulate_secretion_and_uptake(Microenvironment* m, double dt) {
    if (!is_active) return;
    if (volume_is_changed) set_internal_uptake_constants(dt);
    double* rho = m->density_vector(current_voxel_index);
    const int ns = m->number_of_densities();
    const bool track = m->track_internalized_substrates_in_each_agent;
    if (track) {
        // BioFVM/BioFVM_basic_agent.cpp:520-532: what the update below takes
        // out of the voxel, W*((1 - c2)*rho + c1)/c2, goes into the agent
        internalized_substrates.resize(ns, 0.0);
        const double W = m->mesh.dV;
        for (int d = 0; d < ns; ++d) {
            double t = 1.0;
            t -= cell_source_sink_solver_temp2[d];
            t *= rho[d];
            t += cell_source_sink_solver_temp1[d];
            t /= cell_source_sink_solver_temp2[d];
            t *= W;
            internalized_substrates[d] -= t;
        }
    }
    for (int d = 0; d < ns; ++d) rho[d] += cell_source_sink_solver_temp1[d];
    for (int d = 0; d < ns; ++d) rho[d] /= cell_source_sink_solver_temp2[d];
    for (int d = 0; d < ns; ++d) rho[d] += cell_source_sink_solver_temp_export2[d];
    if (track)
        for (int d = 0; d < ns; ++d) internalized_substrates[d] -= cell_source_sink_solver_temp_export1[d];
}

// ---- microenvironment -------------------------------------------------------------------

Microenvironment::Microenvironment() {
    set_density(0, "unnamed", "none");
    mesh.resize(-0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 1, 1, 1);
    host_field.assign((size_t)number_of_voxels() * number_of_densities(), 0.0);
    diffusion_decay_solver = diffusion_decay_solver__constant_coefficients_LOD_3D_GPU;
}

Microenvironment::~Microenvironment() {
    if (env) bfa_destroy(env);
}

int Microenvironment::find_density_index(const std::string& nm) const {
    for (int d = 0; d < (int)density_names.size(); ++d)
        if (density_names[d] == nm) return d;
    return -1;
}

void Microenvironment::set_density(int index, const std::string& nm, const std::string& units) {
    if ((int)density_names.size() <= index) {
        density_names.resize(index + 1, "unnamed");
        density_units.resize(index + 1, "none");
        diffusion_coefficients.resize(index + 1, 0.0);
        decay_rates.resize(index + 1, 0.0);
        dirichlet_activation_vector.resize(index + 1, false);
        face_values.assign(6 * density_names.size(), 1.0);
        dirichlet_planes.clear();
        host_field.assign((size_t)number_of_voxels() * density_names.size(), 0.0);
        if (env) {
            bfa_destroy(env);
            env = nullptr;
        }
    }
    density_names[index] = nm;
    density_units[index] = units;
    state_dirty = true;
}

void Microenvironment::add_density(const std::string& nm, const std::string& units, double D,
                                   double decay) {
    const int i = number_of_densities();
    set_density(i, nm, units);
    diffusion_coefficients[i] = D;
    decay_rates[i] = decay;
}

void Microenvironment::resize_space_uniform(double x_start, double x_end, double y_start,
                                            double y_end, double z_start, double z_end,
                                            double dx_new) {
    mesh.resize(x_start, x_end, y_start, y_end, z_start, z_end, dx_new, dx_new, dx_new);
    host_field.assign((size_t)number_of_voxels() * number_of_densities(), 0.0);
    dirichlet_planes.clear();
    dirichlet_flags.clear();
    if (env) {
        bfa_destroy(env);
        env = nullptr;
    }
    host_newer = true;
    device_newer = false;
    state_dirty = true;
    gradients_valid = false;
}

int Microenvironment::ensure_device() {
    if (env) return BFA_OK;
    int rc = bfa_create(&env, mesh.x_nodes, mesh.y_nodes, mesh.z_nodes, number_of_densities(),
                        mesh.dx, precision);
    if (rc) {
        last_error = bfa_last_error(nullptr);
        env = nullptr;
        return rc;
    }
    host_newer = true;
    state_dirty = true;
    agents_sig.clear();
    tracking_on_device = false;
    return BFA_OK;
}

bfa_env* Microenvironment::device() { return ensure_device() ? nullptr : env; }

int Microenvironment::push_state() {
    int rc = ensure_device();
    if (rc) return rc;
    if (state_dirty) {
        const int ns = number_of_densities();
        std::vector<uint8_t> act(ns);
        for (int d = 0; d < ns; ++d) act[d] = dirichlet_activation_vector[d] ? 1 : 0;
        if ((rc = bfa_set_coefficients(env, diffusion_coefficients.data(), decay_rates.data())) ||
            (rc = bfa_set_dirichlet(env, act.data(), face_registered, face_values.data())) ||
            (rc = bfa_set_dirichlet_planes(env, dirichlet_planes.empty() ? nullptr : dirichlet_planes.data()))) {
            last_error = bfa_last_error(env);
            return rc;
        }
        state_dirty = false;
    }
    if (host_newer) {
        if ((rc = bfa_upload(env, host_field.data(), BFA_LAYOUT_REF))) {
            last_error = bfa_last_error(env);
            return rc;
        }
        host_newer = false;
        device_newer = false;
    }
    return BFA_OK;
}

void Microenvironment::download_if_needed() {
    if (!device_newer || !env) return;
    if (bfa_download(env, host_field.data(), BFA_LAYOUT_REF)) {
        std::cout << "Error: density download failed: " << bfa_last_error(env) << std::endl;
        return;
    }
    device_newer = false;
}

double* Microenvironment::density_vector(int n) {
    download_if_needed();
    host_newer = true;
    gradients_valid = false;
    return host_field.data() + (size_t)n * number_of_densities();
}

std::vector<double>& Microenvironment::density_field() {
    download_if_needed();
    host_newer = true;
    gradients_valid = false;
    return host_field;
}

void Microenvironment::add_dirichlet_face(int face, const std::vector<double>& value) {
    if (face < 0 || face > 5 || (int)value.size() != number_of_densities()) {
        std::cout << "Error: add_dirichlet_face needs a face in 0..5 and one value per density"
                  << std::endl;
        return;
    }
    face_registered[face] = 1;
    std::copy(value.begin(), value.end(), face_values.begin() + (size_t)face * number_of_densities());
    const int nx = mesh.x_nodes, ny = mesh.y_nodes, nz = mesh.z_nodes;
    const bool track_nodes = !dirichlet_flags.empty() || !dirichlet_planes.empty();
    if (track_nodes) {
        // values written per voxel already: this face's voxels take the new vector
        for (int a = 0; a < (face < 2 ? ny : nx); ++a)
            for (int b = 0; b < (face < 4 ? nz : ny); ++b) {
                const int i = face < 2 ? (face == 0 ? 0 : nx - 1) : a;
                const int j = face < 2 ? a : (face < 4 ? (face == 2 ? 0 : ny - 1) : b);
                const int k = face < 4 ? b : (face == 4 ? 0 : nz - 1);
                const int v = mesh.voxel_index(i, j, k);
                if (!dirichlet_planes.empty())
                    for (int d = 0; d < number_of_densities(); ++d) write_dirichlet_value(v, d, value[d]);
                if (!dirichlet_flags.empty()) dirichlet_flags[v] = 1;
            }
    }
    state_dirty = true;
}

// Face planes of the C ABI (bfa_set_dirichlet_planes) from the registered
// faces, replayed in registration order xmin..zmax over the default 1.0.
void Microenvironment::ensure_planes() {
    if (!dirichlet_planes.empty()) return;
    const int nx = mesh.x_nodes, ny = mesh.y_nodes, nz = mesh.z_nodes, ns = number_of_densities();
    dirichlet_planes.assign((size_t)2 * ((size_t)ny * nz + (size_t)nx * nz + (size_t)nx * ny) * ns, 1.0);
    for (int f = 0; f < 6; ++f) {
        if (!face_registered[f]) continue;
        for (int a = 0; a < (f < 2 ? ny : nx); ++a)
            for (int b = 0; b < (f < 4 ? nz : ny); ++b) {
                const int i = f < 2 ? (f == 0 ? 0 : nx - 1) : a;
                const int j = f < 2 ? a : (f < 4 ? (f == 2 ? 0 : ny - 1) : b);
                const int k = f < 4 ? b : (f == 4 ? 0 : nz - 1);
                for (int d = 0; d < ns; ++d)
                    write_dirichlet_value(mesh.voxel_index(i, j, k), d, face_values[(size_t)f * ns + d]);
            }
    }
}

// One voxel's value into every face plane holding it (none for interior
// voxels: the LOD solvers never read them).
void Microenvironment::write_dirichlet_value(int v, int d, double value) {
    const int nx = mesh.x_nodes, ny = mesh.y_nodes, nz = mesh.z_nodes, ns = number_of_densities();
    const int k = v % nz, j = (v / nz) % ny, i = v / (nz * ny);
    const size_t pyz = (size_t)ny * nz * ns, pxz = (size_t)nx * nz * ns, pxy = (size_t)nx * ny * ns;
    double* P = dirichlet_planes.data();
    if (i == 0) P[((size_t)j * nz + k) * ns + d] = value;
    if (i == nx - 1) P[pyz + ((size_t)j * nz + k) * ns + d] = value;
    if (j == 0) P[2 * pyz + ((size_t)i * nz + k) * ns + d] = value;
    if (j == ny - 1) P[2 * pyz + pxz + ((size_t)i * nz + k) * ns + d] = value;
    if (k == 0) P[2 * (pyz + pxz) + ((size_t)i * ny + j) * ns + d] = value;
    if (k == nz - 1) P[2 * (pyz + pxz) + pxy + ((size_t)i * ny + j) * ns + d] = value;
}

void Microenvironment::ensure_flags() {
    if (!dirichlet_flags.empty()) return;
    const int nx = mesh.x_nodes, ny = mesh.y_nodes, nz = mesh.z_nodes;
    dirichlet_flags.assign(number_of_voxels(), 0);
    for (int v = 0; v < number_of_voxels(); ++v) {
        const int k = v % nz, j = (v / nz) % ny, i = v / (nz * ny);
        dirichlet_flags[v] = (face_registered[0] && i == 0) || (face_registered[1] && i == nx - 1) ||
                             (face_registered[2] && j == 0) || (face_registered[3] && j == ny - 1) ||
                             (face_registered[4] && k == 0) || (face_registered[5] && k == nz - 1);
    }
}

void Microenvironment::add_dirichlet_node(int voxel_index, const std::vector<double>& value) {
    // BioFVM/BioFVM_microenvironment.cpp:218-226
    update_dirichlet_node(voxel_index, value);
}

void Microenvironment::update_dirichlet_node(int voxel_index, const std::vector<double>& new_value) {
    // BioFVM/BioFVM_microenvironment.cpp:228-235
    if (voxel_index < 0 || voxel_index >= number_of_voxels() || (int)new_value.size() != number_of_densities()) {
        std::cout << "Error: Dirichlet node needs a voxel index in range and one value per density" << std::endl;
        return;
    }
    ensure_planes();
    ensure_flags();
    for (int d = 0; d < number_of_densities(); ++d) write_dirichlet_value(voxel_index, d, new_value[d]);
    dirichlet_flags[voxel_index] = 1;
    state_dirty = true;
}

void Microenvironment::update_dirichlet_node(int voxel_index, int substrate_index, double new_value) {
    // BioFVM/BioFVM_microenvironment.cpp:237-245 (its per-voxel activation flag
    // is not read by the LOD solvers)
    if (voxel_index < 0 || voxel_index >= number_of_voxels() || substrate_index < 0 ||
        substrate_index >= number_of_densities()) {
        std::cout << "Error: Dirichlet node needs a voxel and a density index in range" << std::endl;
        return;
    }
    ensure_planes();
    ensure_flags();
    write_dirichlet_value(voxel_index, substrate_index, new_value);
    dirichlet_flags[voxel_index] = 1;
    state_dirty = true;
}

void Microenvironment::remove_dirichlet_node(int voxel_index) {
    // BioFVM/BioFVM_microenvironment.cpp:247-252: clears is_Dirichlet only; the
    // face pass of the LOD solvers still writes the voxel's value vector
    if (voxel_index < 0 || voxel_index >= number_of_voxels()) return;
    ensure_flags();
    dirichlet_flags[voxel_index] = 0;
}

bool Microenvironment::is_dirichlet_node(int voxel_index) const {
    if (voxel_index < 0 || voxel_index >= number_of_voxels()) return false;
    if (!dirichlet_flags.empty()) return dirichlet_flags[voxel_index] != 0;
    const int nx = mesh.x_nodes, ny = mesh.y_nodes, nz = mesh.z_nodes;
    const int k = voxel_index % nz, j = (voxel_index / nz) % ny, i = voxel_index / (nz * ny);
    return (face_registered[0] && i == 0) || (face_registered[1] && i == nx - 1) ||
           (face_registered[2] && j == 0) || (face_registered[3] && j == ny - 1) ||
           (face_registered[4] && k == 0) || (face_registered[5] && k == nz - 1);
}

void Microenvironment::set_substrate_dirichlet_activation(int substrate_index, bool new_value) {
    dirichlet_activation_vector[substrate_index] = new_value;
    state_dirty = true;
}

void Microenvironment::apply_dirichlet_boundaries_conditions() {
    if (push_state() || bfa_apply_dirichlet(env)) {
        std::cout << "Error: apply_dirichlet_boundaries_conditions: "
                  << (env ? bfa_last_error(env) : last_error.c_str()) << std::endl;
        return;
    }
    device_newer = true;
    gradients_valid = false;
}

void Microenvironment::simulate_diffusion_decay(double dt) {
    // BioFVM/BioFVM_microenvironment.cpp:1132-1143
    if (diffusion_decay_solver) {
        diffusion_decay_solver(*this, dt);
    } else {
        std::cout << "Warning: diffusion-reaction-source/sink solver not set for Microenvironment object at "
                  << this << ". Nothing happened!" << std::endl;
    }
}

void Microenvironment::simulate_cell_sources_and_sinks(std::vector<Basic_Agent*>& agents, double dt) {
    if (push_state()) {
        std::cout << "Error: simulate_cell_sources_and_sinks: " << last_error << std::endl;
        return;
    }
    const int ns = number_of_densities();
    const int n = (int)agents.size();
    bool changed = agents_sig.size() != agents.size();
    for (int a = 0; a < n && !changed; ++a) changed = agents_sig[a] != agents[a];
    std::vector<int32_t> vox(n);
    std::vector<uint8_t> act(n);
    std::vector<double> c1((size_t)n * ns), c2((size_t)n * ns), e2((size_t)n * ns);
    for (int a = 0; a < n; ++a) {
        Basic_Agent* ag = agents[a];
        if (ag->volume_is_changed) {
            ag->set_internal_uptake_constants(dt);
            changed = true;
        }
        vox[a] = ag->current_voxel_index;
        act[a] = ag->is_active ? 1 : 0;
        for (int d = 0; d < ns; ++d) {
            c1[(size_t)a * ns + d] = ag->cell_source_sink_solver_temp1[d];
            c2[(size_t)a * ns + d] = ag->cell_source_sink_solver_temp2[d];
            e2[(size_t)a * ns + d] = ag->cell_source_sink_solver_temp_export2[d];
        }
    }
    // agents move and change: compare with what the device holds
    changed = changed || vox != sig_voxel || act != sig_active || c1 != sig_c1 || c2 != sig_c2 ||
              e2 != sig_e2;
    if (changed) {
        if (bfa_set_agents(env, n, vox.data(), act.data(), c1.data(), c2.data(), e2.data())) {
            std::cout << "Error: set_agents: " << bfa_last_error(env) << std::endl;
            return;
        }
        agents_sig.assign(agents.begin(), agents.end());
        sig_voxel.swap(vox);
        sig_active.swap(act);
        sig_c1.swap(c1);
        sig_c2.swap(c2);
        sig_e2.swap(e2);
        tracking_on_device = false;  // the agents' own totals are sent again below
    }
    const bool track = track_internalized_substrates_in_each_agent && n > 0;
    if (track && !tracking_on_device) {
        // the agents hold the authoritative totals between calls
        std::vector<double> tot((size_t)n * ns), e1((size_t)n * ns);
        for (int a = 0; a < n; ++a) {
            Basic_Agent* ag = agents[a];
            ag->internalized_substrates.resize(ns, 0.0);
            for (int d = 0; d < ns; ++d) {
                tot[(size_t)a * ns + d] = ag->internalized_substrates[d];
                e1[(size_t)a * ns + d] = ag->cell_source_sink_solver_temp_export1[d];
            }
        }
        if (bfa_track_internalized(env, 1, tot.data(), e1.data())) {
            std::cout << "Error: internalized tracking: " << bfa_last_error(env) << std::endl;
            return;
        }
        tracking_on_device = true;
    } else if (!track && tracking_on_device) {
        (void)bfa_track_internalized(env, 0, nullptr, nullptr);
        tracking_on_device = false;
    }
    if (bfa_secretion_uptake(env)) {
        std::cout << "Error: secretion/uptake: " << bfa_last_error(env) << std::endl;
        return;
    }
    if (track) {
        std::vector<double> tot((size_t)n * ns);
        if (bfa_get_internalized(env, tot.data())) {
            std::cout << "Error: internalized tracking: " << bfa_last_error(env) << std::endl;
            return;
        }
        for (int a = 0; a < n; ++a)
            std::copy(tot.begin() + (size_t)a * ns, tot.begin() + (size_t)(a + 1) * ns,
                      agents[a]->internalized_substrates.begin());
    }
    device_newer = true;
    gradients_valid = false;
}

void zero_function(Microenvironment* m, int, double* write_destination) {
    for (int d = 0; d < m->number_of_densities(); ++d) write_destination[d] = 0.0;
}

void one_function(Microenvironment* m, int, double* write_destination) {
    for (int d = 0; d < m->number_of_densities(); ++d) write_destination[d] = 1.0;
}

void Microenvironment::simulate_bulk_sources_and_sinks(double dt) {
    const int ns = number_of_densities();
    const size_t total = (size_t)number_of_voxels() * ns;
    std::vector<double> S(total), T(total), U(total);
    for (int n = 0; n < number_of_voxels(); ++n) {
        bulk_supply_rate_function(this, n, &S[(size_t)n * ns]);
        bulk_supply_target_densities_function(this, n, &T[(size_t)n * ns]);
        bulk_uptake_rate_function(this, n, &U[(size_t)n * ns]);
    }
    if (push_state() || bfa_set_bulk_rates(env, S.data(), T.data(), U.data(), BFA_LAYOUT_REF) ||
        bfa_bulk_sources_and_sinks(env, dt)) {
        std::cout << "Error: simulate_bulk_sources_and_sinks: " << (env ? bfa_last_error(env) : last_error.c_str())
                  << std::endl;
        return;
    }
    device_newer = true;
    gradients_valid = false;
}

void Microenvironment::compute_all_gradient_vectors() {
    // BioFVM/BioFVM_microenvironment.cpp:1454-1615, on the device
    if (push_state()) {
        std::cout << "Error: compute_all_gradient_vectors: " << last_error << std::endl;
        return;
    }
    gradient_field.resize((size_t)number_of_voxels() * number_of_densities() * 3);
    if (bfa_gradients(env, gradient_field.data(), BFA_LAYOUT_REF)) {
        std::cout << "Error: compute_all_gradient_vectors: " << bfa_last_error(env) << std::endl;
        return;
    }
    gradients_valid = true;
}

double* Microenvironment::gradient_vector(int n) {
    if (!gradients_valid) compute_all_gradient_vectors();
    if (!gradients_valid) return nullptr;
    return gradient_field.data() + (size_t)n * number_of_densities() * 3;
}

double* Microenvironment::nearest_gradient_vector(const std::vector<double>& position) {
    const int n = mesh.nearest_voxel_index(position);
    return n < 0 ? nullptr : gradient_vector(n);
}

double Microenvironment::total_mass(int substrate) {
    if (push_state()) return NAN;
    std::vector<double> m(number_of_densities());
    if (bfa_total_mass(env, m.data())) return NAN;
    return m[substrate];
}

void Microenvironment::write_to_matlab(const std::string& filename) {
    download_if_needed();
    const int ns = number_of_densities();
    const unsigned rows = 4 + ns, cols = (unsigned)number_of_voxels();
    std::vector<double> data((size_t)rows * cols);
    for (int i = 0; i < mesh.x_nodes; ++i)
        for (int j = 0; j < mesh.y_nodes; ++j)
            for (int k = 0; k < mesh.z_nodes; ++k) {
                const size_t n = (size_t)mesh.voxel_index(i, j, k);
                double* col = data.data() + n * rows;
                col[0] = mesh.x_coordinate(i);
                col[1] = mesh.y_coordinate(j);
                col[2] = mesh.z_coordinate(k);
                col[3] = mesh.dV;
                for (int d = 0; d < ns; ++d) col[4 + d] = host_field[n * ns + d];
            }
    write_matlab4(filename, rows, cols, "multiscale_microenvironment", data.data());
}

bool write_matlab4(const std::string& filename, unsigned rows, unsigned cols,
                   const std::string& variable_name, const double* column_major) {
    // header: type (0 = little-endian, double, full), rows, cols, imaginary
    // flag, name length (without a terminating NUL, as the reference writes
    // it), then the name and the data (BioFVM/BioFVM_matlab.cpp:536-580)
    FILE* fp = std::fopen(filename.c_str(), "wb");
    if (!fp) {
        std::cout << "Error: could not open file " << filename << "!" << std::endl;
        return false;
    }
    const uint32_t hdr[5] = {0u, rows, cols, 0u, (uint32_t)variable_name.size()};
    bool ok = std::fwrite(hdr, sizeof hdr, 1, fp) == 1 &&
              std::fwrite(variable_name.data(), 1, variable_name.size(), fp) == variable_name.size() &&
              std::fwrite(column_major, sizeof(double), (size_t)rows * cols, fp) == (size_t)rows * cols;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) std::cout << "Error: could not write " << filename << "!" << std::endl;
    return ok;
}

std::vector<std::vector<double>> read_matlab(const std::string& filename) {
    // BioFVM/BioFVM_matlab.cpp:71-270
    std::vector<std::vector<double>> out;
    FILE* fp = std::fopen(filename.c_str(), "rb");
    if (!fp) {
        std::cout << "Error: could not open file " << filename << "!" << std::endl;
        return out;
    }
    uint32_t hdr[5];
    if (std::fread(hdr, sizeof hdr, 1, fp) != 1) {
        std::fclose(fp);
        std::cout << "Error reading file " << filename << ": short header" << std::endl;
        return out;
    }
    const uint32_t type = hdr[0], rows = hdr[1], cols = hdr[2], imag = hdr[3], namelen = hdr[4];
    const uint32_t numeric = type / 1000, reserved = (type / 100) % 10, format = (type / 10) % 10,
                   matrix = type % 10;
    if (numeric != 0 || reserved != 0 || format > 5 || matrix != 0) {
        std::fclose(fp);
        std::cout << "Error reading file " << filename << ": I can't read this format yet!" << std::endl;
        return out;
    }
    if (imag != 0) {
        std::fclose(fp);
        std::cout << "Error: I can't read imaginary matrices yet!" << std::endl;
        return out;
    }
    std::vector<char> name(namelen);
    static const size_t width[6] = {8, 4, 4, 2, 2, 1};
    std::vector<unsigned char> raw((size_t)rows * cols * width[format]);
    if ((namelen && std::fread(name.data(), 1, namelen, fp) != namelen) ||
        (!raw.empty() && std::fread(raw.data(), 1, raw.size(), fp) != raw.size())) {
        std::fclose(fp);
        std::cout << "Error reading file " << filename << ": short data" << std::endl;
        return out;
    }
    std::fclose(fp);
    out.assign(rows, std::vector<double>(cols, 0.0));
    for (size_t n = 0; n < (size_t)rows * cols; ++n) {
        const unsigned char* p = raw.data() + n * width[format];
        double v = 0.0;
        switch (format) {
            case 0: { double x; std::memcpy(&x, p, 8); v = x; break; }
            case 1: { float x; std::memcpy(&x, p, 4); v = x; break; }
            case 2: { int32_t x; std::memcpy(&x, p, 4); v = x; break; }
            case 3: { int16_t x; std::memcpy(&x, p, 2); v = x; break; }
            case 4: { uint16_t x; std::memcpy(&x, p, 2); v = x; break; }
            default: v = *p; break;
        }
        out[n % rows][n / rows] = v;  // column-major
    }
    return out;
}

bool load_initial_conditions_from_matlab(Microenvironment& M, const std::string& filename) {
    std::vector<std::vector<double>> mat = read_matlab(filename);
    const int ns = M.number_of_densities();
    if ((int)mat.size() != M.number_of_voxels()) {
        std::cout << "ERROR : Wrong number of voxels supplied in the .mat file specifying BioFVM initial conditions."
                  << std::endl
                  << "\tExpected: " << M.number_of_voxels() << std::endl
                  << "\tFound: " << mat.size() << std::endl;
        return false;
    }
    if (mat.empty() || (int)mat[0].size() != ns + 3) {
        std::cout << "ERROR : Wrong number of density values supplied in the .mat file specifying BioFVM initial conditions."
                  << std::endl;
        return false;
    }
    std::vector<int> vox(mat.size());
    std::vector<char> seen(M.number_of_voxels(), 0);
    for (size_t r = 0; r < mat.size(); ++r) {
        vox[r] = M.mesh.nearest_voxel_index({mat[r][0], mat[r][1], mat[r][2]});
        if (vox[r] < 0 || seen[vox[r]]) {
            std::cout << "ERROR : the matlab-supplied initial conditions for BioFVM repeat the same voxel. "
                         "Fix the .mat file and try again." << std::endl;
            return false;
        }
        seen[vox[r]] = 1;
    }
    std::vector<double>& u = M.density_field();
    for (size_t r = 0; r < mat.size(); ++r)
        for (int d = 0; d < ns; ++d) u[(size_t)vox[r] * ns + d] = mat[r][3 + d];
    return true;
}

// The numbers of one CSV line: skip to the next character that can start a
// number (a digit, '.', '-', 'e' or 'E'), read it with strtod, repeat.
static std::vector<double> csv_numbers(const std::string& line) {
    std::vector<double> out;
    const char* s = line.c_str();
    const size_t n = line.size();
    size_t i = 0;
    while (i < n) {
        while (i < n && !std::isdigit((unsigned char)s[i]) && s[i] != '.' && s[i] != '-' && s[i] != 'e' &&
               s[i] != 'E')
            ++i;
        if (i >= n) break;
        char* end = nullptr;
        const double v = std::strtod(s + i, &end);
        if (end == s + i) {  // not a number after all (e.g. a lone 'e'): skip the character
            ++i;
            continue;
        }
        out.push_back(v);
        i = (size_t)(end - s);
    }
    return out;
}

bool load_initial_conditions_from_csv(Microenvironment& M, const std::string& filename) {
    std::ifstream file(filename);
    if (!file) {
        std::cout << "ERROR: " << filename << " not found during cell loading. Quitting." << std::endl;
        return false;
    }
    const int ns = M.number_of_densities();
    std::vector<std::string> lines;
    for (std::string line; std::getline(file, line);) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        bool blank = true;
        for (char c : line) blank = blank && std::isspace((unsigned char)c);
        if (!blank) lines.push_back(line);
    }
    std::vector<int> subs;
    size_t first = 0;
    const bool header = !lines.empty() && (lines[0][0] == 'x' || lines[0][0] == 'X');
    if (header) {
        const std::string& h = lines[0];
        if (h.size() < 5 || (h[2] != 'Y' && h[2] != 'y') || (h[4] != 'Z' && h[4] != 'z')) {
            std::cout << "ERROR: Header row starts with x but then not y,z? What is this? Exiting now." << std::endl;
            return false;
        }
        std::stringstream stream(h);
        std::vector<std::string> names;
        for (std::string field; std::getline(stream, field, ',');) names.push_back(field);
        for (size_t c = 3; c < names.size(); ++c) {
            const int d = M.find_density_index(names[c]);
            if (d < 0) {
                std::cout << "ERROR: Substrate " << names[c] << " not found in the BioFVM microenvironment. Exiting now."
                          << std::endl;
                return false;
            }
            subs.push_back(d);
        }
        first = 1;
    }
    std::cout << "Loading substrate initial conditions from CSV file " << filename << " ... " << std::endl;
    std::vector<std::vector<double>> rows;
    bool warned = false;
    for (size_t r = first; r < lines.size(); ++r) {
        std::vector<double> data = csv_numbers(lines[r]);
        if (!header && subs.empty())
            for (int c = 0; c + 3 < (int)data.size() && c < ns; ++c) subs.push_back(c);
        if (!warned && !header && (int)data.size() != ns + 3) {
            std::cout << "WARNING: Wrong number of density values supplied in the .csv file specifying BioFVM initial "
                         "conditions." << std::endl
                      << "\tExpected: " << ns << std::endl
                      << "\tFound: " << (int)data.size() - 3 << std::endl;
            warned = true;
        }
        if (data.size() < 3 + subs.size()) {
            std::cout << "ERROR : row " << r + 1 << " of " << filename << " has " << data.size()
                      << " values; expected x, y, z and " << subs.size() << " densities." << std::endl;
            return false;
        }
        rows.push_back(std::move(data));
    }
    std::vector<int> vox(rows.size());
    std::vector<char> seen(M.number_of_voxels(), 0);
    for (size_t r = 0; r < rows.size(); ++r) {
        vox[r] = M.mesh.nearest_voxel_index({rows[r][0], rows[r][1], rows[r][2]});
        if (vox[r] < 0 || seen[vox[r]]) {
            std::cout << "ERROR : the csv-supplied initial conditions for BioFVM repeat the same voxel. Fix the .csv file "
                         "and try again." << std::endl;
            return false;
        }
        seen[vox[r]] = 1;
    }
    if ((int)rows.size() != M.number_of_voxels()) {
        std::cout << "ERROR : Wrong number of voxels supplied in the .csv file specifying BioFVM initial conditions."
                  << std::endl
                  << "\tExpected: " << M.number_of_voxels() << std::endl
                  << "\tFound: " << rows.size() << std::endl;
        return false;
    }
    std::vector<double>& u = M.density_field();
    for (size_t r = 0; r < rows.size(); ++r)
        for (size_t c = 0; c < subs.size(); ++c) u[(size_t)vox[r] * ns + subs[c]] = rows[r][3 + c];
    return true;
}

int Microenvironment::step_with(int variant, double dt) {
    // the reference solvers need a uniform Cartesian mesh (BioFVM/BioFVM_solvers.cpp:116-121)
    if (!mesh.uniform_mesh) {
        std::cout << "Error: This algorithm is written for uniform Cartesian meshes. Try: other solvers!"
                  << std::endl
                  << std::endl;
        return BFA_ERR_ARG;
    }
    if (push_state()) {
        std::cout << "Error: diffusion solver setup failed: " << last_error << std::endl;
        return BFA_ERR_STATE;
    }
    int rc = bfa_set_solver(env, variant);
    if (rc == BFA_OK) rc = bfa_diffusion_decay(env, dt);
    if (rc) {
        std::cout << "Error: diffusion solver failed: " << bfa_last_error(env) << std::endl;
        return rc;
    }
    device_newer = true;
    gradients_valid = false;
    return BFA_OK;
}

void diffusion_decay_solver__constant_coefficients_LOD_3D_GPU(Microenvironment& M, double dt) {
    (void)M.step_with(BFA_LOD_3D_BLOCKING, dt);
}

void diffusion_decay_solver__constant_coefficients_LOD_3D_AVX256D_GPU(Microenvironment& M, double dt) {
    (void)M.step_with(BFA_LOD_3D_AVX256D, dt);
}

}  // namespace BioFVM_amd
