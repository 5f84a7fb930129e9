// BioFVM_reference_binding.h -- the drop-in binding for the reference tree.
//
// Header-only templates over the reference's own types, written against the
// public interface of BioFVM::Microenvironment (BioFVM/BioFVM_microenvironment.h)
// and Basic_Agent (BioFVM/BioFVM_basic_agent.h), so a maintainer includes this
// header in one translation unit of the reference and assigns
//
//   microenvironment.diffusion_decay_solver_mpi = lod_3d_mi355x_mpi;   // (Microenvironment&, double,
//                                                                        //  mpi_Environment&, mpi_Cartesian&)
//   microenvironment.diffusion_decay_solver     = lod_3d_mi355x;       // (Microenvironment&, double)
//
// after instantiating the two thin wrappers shown in INTEGRATION.md.  Nothing
// here needs a friend declaration: the Dirichlet values of the boundary voxels
// (the private dirichlet_value_vectors, BioFVM_microenvironment.cpp:218-245)
// are read through the reference's own public face pass,
// apply_dirichlet_boundaries_conditions(0, 1) (:348-446), which writes each
// face voxel's value vector into the density field for the active
// substrates; the binding keeps the field's face values, lets that pass run,
// reads the faces and puts the field back.  Every call re-reads the
// coefficients, the activation flags and those face values and sends only
// what changed, so later update_dirichlet_node, set_substrate_dirichlet_activation
// or coefficient changes take effect on the next step.
//
// The field lives in the reference's host memory (p_density_vectors, reference
// layout); each step uploads it, runs the device step and downloads it.  The
// device-resident form (no per-step transfers) is the C++ mirror in
// BioFVM_amd.h.
#pragma once

#include <cstdint>
#include <iostream>
#include <map>
#include <vector>

#include "biofvm_amd.h"

namespace bfa_binding {

struct State {
    bfa_env* env = nullptr;
    int nx = 0, ny = 0, nz = 0, ns = 0;
    std::vector<double> D, decay, planes;
    std::vector<uint8_t> active;
    bool have_planes = false;
    ~State() {
        if (env) bfa_destroy(env);
    }
};

template <class Micro>
State& state_of(Micro& M) {
    static std::map<const void*, State> states;
    return states[&M];
}

inline void report(const char* what, bfa_env* env) {
    std::cout << "Error: " << what << ": " << bfa_last_error(env) << std::endl;
}

// Face values of every boundary voxel for the active substrates (0 for the
// others, which the face pass neither writes nor reads), in the plane layout
// of bfa_set_dirichlet_planes, read through the reference's own face pass;
// the field is left as it was.
template <class Micro>
void gather_face_values(Micro& M, int nx, int ny, int nz, int ns, const std::vector<uint8_t>& active,
                        std::vector<double>& planes) {
    std::vector<double>& rho = *M.p_density_vectors;
    const size_t pyz = (size_t)ny * nz * ns, pxz = (size_t)nx * nz * ns, pxy = (size_t)nx * ny * ns;
    planes.assign(2 * (pyz + pxz + pxy), 0.0);
    // the six planes' voxels: x faces [j][k], y faces [i][k], z faces [i][j]
    auto for_face_voxels = [&](auto&& fn) {
        for (int f = 0; f < 6; ++f) {
            const int A = f < 2 ? ny : nx, B = f < 4 ? nz : ny;
            const size_t off = f < 2 ? f * pyz : (f < 4 ? 2 * pyz + (f - 2) * pxz : 2 * (pyz + pxz) + (f - 4) * pxy);
            for (int a = 0; a < A; ++a)
                for (int b = 0; b < B; ++b) {
                    const int i = f < 2 ? (f == 0 ? 0 : nx - 1) : a;
                    const int j = f < 2 ? a : (f < 4 ? (f == 2 ? 0 : ny - 1) : b);
                    const int k = f < 4 ? b : (f == 4 ? 0 : nz - 1);
                    fn((size_t)M.voxel_index(i, j, k) * ns, off + ((size_t)a * B + b) * ns);
                }
        }
    };
    std::vector<double> keep(planes.size());
    for_face_voxels([&](size_t v, size_t q) {
        for (int d = 0; d < ns; ++d) keep[q + d] = rho[v + d];
    });
    M.apply_dirichlet_boundaries_conditions(0, 1);
    for_face_voxels([&](size_t v, size_t q) {
        for (int d = 0; d < ns; ++d) planes[q + d] = active[d] ? rho[v + d] : 0.0;
    });
    for_face_voxels([&](size_t v, size_t q) {
        for (int d = 0; d < ns; ++d) rho[v + d] = keep[q + d];
    });
}

// Creates the device microenvironment on first use and sends the solver
// state that changed since the last call.  Returns false after printing the
// error, as the reference's solvers print and return.
template <class Micro>
bool sync_state(Micro& M, State& S) {
    const int nx = (int)M.mesh.x_coordinates.size(), ny = (int)M.mesh.y_coordinates.size(),
              nz = (int)M.mesh.z_coordinates.size(), ns = (int)M.number_of_densities();
    if (S.env && (S.nx != nx || S.ny != ny || S.nz != nz || S.ns != ns)) {
        bfa_destroy(S.env);
        S = State();
    }
    if (!S.env) {
        if (bfa_create(&S.env, nx, ny, nz, ns, M.mesh.dx, BFA_FP64) != BFA_OK) {
            report("bfa_create", nullptr);
            S.env = nullptr;
            return false;
        }
        S.nx = nx;
        S.ny = ny;
        S.nz = nz;
        S.ns = ns;
    }
    std::vector<double> D(M.diffusion_coefficients.begin(), M.diffusion_coefficients.end());
    std::vector<double> decay(M.decay_rates.begin(), M.decay_rates.end());
    if (D != S.D || decay != S.decay) {
        if (bfa_set_coefficients(S.env, D.data(), decay.data()) != BFA_OK) {
            report("bfa_set_coefficients", S.env);
            return false;
        }
        S.D.swap(D);
        S.decay.swap(decay);
    }
    std::vector<uint8_t> active(ns);
    for (int d = 0; d < ns; ++d) active[d] = M.get_substrate_dirichlet_activation(d) ? 1 : 0;
    std::vector<double> planes;
    gather_face_values(M, nx, ny, nz, ns, active, planes);
    if (active != S.active) {  // the face values come from the planes below
        const uint8_t reg[6] = {0, 0, 0, 0, 0, 0};
        std::vector<double> fv(6 * (size_t)ns, 1.0);
        if (bfa_set_dirichlet(S.env, active.data(), reg, fv.data()) != BFA_OK) {
            report("bfa_set_dirichlet", S.env);
            return false;
        }
        S.active = active;
    }
    if (!S.have_planes || planes != S.planes) {
        if (bfa_set_dirichlet_planes(S.env, planes.data()) != BFA_OK) {
            report("bfa_set_dirichlet_planes", S.env);
            return false;
        }
        S.planes.swap(planes);
        S.have_planes = true;
    }
    return true;
}

// The serial-signature solver: one simulate_diffusion_decay(dt) of
// LOD_3D_BLOCKING on the device, the field uploaded and downloaded around it.
template <class Micro>
void lod_3d_mi355x(Micro& M, double dt) {
    State& S = state_of(M);
    if (!sync_state(M, S)) return;
    std::vector<double>& rho = *M.p_density_vectors;
    if (bfa_upload(S.env, rho.data(), BFA_LAYOUT_REF) != BFA_OK || bfa_set_solver(S.env, BFA_LOD_3D_BLOCKING) != BFA_OK ||
        bfa_diffusion_decay(S.env, dt) != BFA_OK || bfa_download(S.env, rho.data(), BFA_LAYOUT_REF) != BFA_OK) {
        report("MI355X diffusion solver", S.env);
        return;
    }
}

// The MPI-signature solver (diffusion_decay_solver_mpi,
// BioFVM/BioFVM_microenvironment.h:206): the whole domain is resident on one
// GPU, so it runs at world size 1 only and prints and returns otherwise.
template <class Micro, class World, class Cart>
void lod_3d_mi355x_mpi(Micro& M, double dt, World& world, Cart& cart) {
    (void)cart;
    if (world.size != 1) {
        std::cout << "Error: the MI355X solver keeps the whole domain on one GPU; run with one MPI rank"
                  << std::endl;
        return;
    }
    lod_3d_mi355x(M, dt);
}

// The agents' secretion and uptake for a whole container in one device call
// (replaces the per-cell loop of Cell_Container::update_all_cells,
// core/PhysiCell_cell_container.cpp:427-435, through Secretion::advance).
// Agents is any range of pointers to objects with the Basic_Agent members
// used below.  With track (default_microenvironment_options.
// track_internalized_substrates_in_each_agent) the agents'
// internalized_substrates are updated too.  The field makes the same host
// round trip as the solver.
template <class Micro, class Agents>
void secretion_uptake_mi355x(Micro& M, Agents& agents, bool track) {
    State& S = state_of(M);
    if (!sync_state(M, S)) return;
    const int ns = S.ns;
    std::vector<int32_t> vox;
    std::vector<uint8_t> act;
    std::vector<double> c1, c2, e1, e2, tot;
    for (auto* a : agents) {
        vox.push_back(a->current_voxel_index);
        act.push_back(a->is_active ? 1 : 0);
        for (int d = 0; d < ns; ++d) {
            c1.push_back(a->cell_source_sink_solver_temp1[d]);
            c2.push_back(a->cell_source_sink_solver_temp2[d]);
            e1.push_back(a->cell_source_sink_solver_temp_export1[d]);
            e2.push_back(a->cell_source_sink_solver_temp_export2[d]);
            if (track) tot.push_back((*a->internalized_substrates)[d]);
        }
    }
    std::vector<double>& rho = *M.p_density_vectors;
    const int n = (int)vox.size();
    bool ok = bfa_upload(S.env, rho.data(), BFA_LAYOUT_REF) == BFA_OK &&
              bfa_set_agents(S.env, n, vox.data(), act.data(), c1.data(), c2.data(), e2.data()) == BFA_OK;
    if (ok && track && n > 0) ok = bfa_track_internalized(S.env, 1, tot.data(), e1.data()) == BFA_OK;
    if (ok) ok = bfa_secretion_uptake(S.env) == BFA_OK && bfa_download(S.env, rho.data(), BFA_LAYOUT_REF) == BFA_OK;
    if (ok && track && n > 0) ok = bfa_get_internalized(S.env, tot.data()) == BFA_OK;
    if (!ok) {
        report("MI355X secretion/uptake", S.env);
        return;
    }
    if (track && n > 0) {
        int a = 0;
        for (auto* ag : agents) {
            for (int d = 0; d < ns; ++d) (*ag->internalized_substrates)[d] = tot[(size_t)a * ns + d];
            ++a;
        }
        (void)bfa_track_internalized(S.env, 0, nullptr, nullptr);
    }
}

}  // namespace bfa_binding
