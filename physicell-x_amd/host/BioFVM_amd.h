// BioFVM_amd.h -- C++ host side of the MI355X diffusion-decay path.
//
// Mirrors the part of the reference's BioFVM interface that the LOD
// diffusion-decay path uses, with the same names, argument meaning and error
// behaviour, and runs the compute through the C ABI of include/biofvm_amd.h:
//
//   Cartesian_Mesh        BioFVM/BioFVM_mesh.h:215-301 (uniform mesh, world size 1)
//   Microenvironment      BioFVM/BioFVM_microenvironment.h:71-373
//     diffusion_decay_solver function pointer          (.h:199)
//     simulate_diffusion_decay(dt)                     (.cpp:1132-1143)
//     simulate_cell_sources_and_sinks(agents, dt)      (.cpp:1370-1384)
//     apply_dirichlet_boundaries_conditions()          (.cpp:348-446)
//     add_dirichlet_node / set_substrate_dirichlet_activation (.cpp:218-273)
//   Basic_Agent           BioFVM/BioFVM_basic_agent.h
//     set_internal_uptake_constants(dt)                (.cpp:224-274)
//     simulate_secretion_and_uptake(M, dt)             (.cpp:509-559)
//
// The density field lives on the GPU between steps.  density_vector() and
// the other host accessors download it when the device copy is newer, and
// mark it to be uploaded before the next device step when the caller writes
// through them.  Solver errors are printed and the call returns, as the
// reference does (BioFVM/BioFVM_solvers.cpp:116-121).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "biofvm_amd.h"

namespace BioFVM_amd {

class Microenvironment;

// Rate callbacks writing 0 / 1 for every density of a voxel
// (BioFVM/BioFVM_microenvironment.cpp:99-110).
void zero_function(Microenvironment* m, int voxel_index, double* write_destination);
void one_function(Microenvironment* m, int voxel_index, double* write_destination);

struct Cartesian_Mesh {
    // {x_min, y_min, z_min, x_max, y_max, z_max}, as bounding_box in the reference
    double bounding_box[6] = {0, 0, 0, 1, 1, 1};
    double dx = 1, dy = 1, dz = 1, dV = 1;
    int x_nodes = 1, y_nodes = 1, z_nodes = 1;
    bool uniform_mesh = true;

    // Cartesian_Mesh::resize (BioFVM/BioFVM_mesh.cpp:1436-1662) at world size 1
    void resize(double x_start, double x_end, double y_start, double y_end, double z_start,
                double z_end, double dx_new, double dy_new, double dz_new);
    int voxel_index(int i, int j, int k) const { return (i * y_nodes + j) * z_nodes + k; }
    // voxel centres start + (i + 1/2) d, as resize computes them
    // (BioFVM/BioFVM_mesh.cpp:1324-1329; the reference build contracts the
    // product and sum into one fused multiply-add)
    double x_coordinate(int i) const { return std::fma(i + 0.5, dx, bounding_box[0]); }
    double y_coordinate(int j) const { return std::fma(j + 0.5, dy, bounding_box[1]); }
    double z_coordinate(int k) const { return std::fma(k + 0.5, dz, bounding_box[2]); }
    int number_of_voxels() const { return x_nodes * y_nodes * z_nodes; }
    bool is_position_valid(double x, double y, double z) const;
    int nearest_voxel_index(const std::vector<double>& position) const;
};

class Basic_Agent {
  public:
    std::vector<double> position = {0, 0, 0};
    double volume = 2494.0;
    bool is_active = true;
    int current_voxel_index = -1;
    std::vector<double> secretion_rates, uptake_rates, saturation_densities, net_export_rates;
    // substrate taken up minus substrate released, per density, when the
    // microenvironment tracks it (track_internalized_substrates_in_each_agent;
    // the reference keeps it behind Basic_Agent::internalized_substrates)
    std::vector<double> internalized_substrates;
    // constants of the implicit source/sink update (BioFVM/BioFVM_basic_agent.cpp:241-265)
    std::vector<double> cell_source_sink_solver_temp1, cell_source_sink_solver_temp2;
    std::vector<double> cell_source_sink_solver_temp_export1, cell_source_sink_solver_temp_export2;
    bool volume_is_changed = true;

    void register_microenvironment(Microenvironment* m);
    void update_voxel_index();
    void set_internal_uptake_constants(double dt);
    // One agent's update on the host copy of the field (the reference's
    // per-agent call).  Microenvironment::simulate_cell_sources_and_sinks
    // does all agents at once on the GPU instead.
    void simulate_secretion_and_uptake(Microenvironment* m, double dt);

  private:
    Microenvironment* microenvironment = nullptr;
};

class Microenvironment {
  public:
    Cartesian_Mesh mesh;
    std::vector<std::string> density_names, density_units;
    std::vector<double> diffusion_coefficients, decay_rates;
    std::vector<bool> dirichlet_activation_vector;
    std::string name = "unnamed";
    int precision = BFA_FP64;

    void (*diffusion_decay_solver)(Microenvironment&, double) = nullptr;

    Microenvironment();
    ~Microenvironment();
    Microenvironment(const Microenvironment&) = delete;
    Microenvironment& operator=(const Microenvironment&) = delete;

    // densities and space
    void set_density(int index, const std::string& name, const std::string& units);
    void add_density(const std::string& name, const std::string& units, double D, double decay);
    int number_of_densities() const { return (int)density_names.size(); }
    // index of the density of that name, -1 if none (BioFVM/BioFVM_microenvironment.cpp:1043)
    int find_density_index(const std::string& name) const;
    int number_of_voxels() const { return mesh.number_of_voxels(); }
    void resize_space_uniform(double x_start, double x_end, double y_start, double y_end,
                              double z_start, double z_end, double dx);

    // field access in the reference layout ((i*ny + j)*nz + k)*ns + d
    double* density_vector(int n);
    double* density_vector(int i, int j, int k) { return density_vector(mesh.voxel_index(i, j, k)); }
    std::vector<double>& density_field();  // whole field (host copy, synchronised)

    // Dirichlet state (BioFVM/BioFVM_microenvironment.cpp:218-273,632).  Every
    // voxel holds one value vector (dirichlet_value_vectors, default 1.0),
    // written by the last call that covered it; the LOD solvers read it on the
    // six faces only, so this keeps the face voxels' values.  add_dirichlet_face
    // is add_dirichlet_node over every voxel of a face (xmin..zmax = 0..5), as
    // initialize_microenvironment registers faces (:2649-2786).
    void add_dirichlet_face(int face, const std::vector<double>& value);
    void add_dirichlet_node(int voxel_index, const std::vector<double>& value);
    void update_dirichlet_node(int voxel_index, const std::vector<double>& new_value);
    void update_dirichlet_node(int voxel_index, int substrate_index, double new_value);
    void remove_dirichlet_node(int voxel_index);
    bool is_dirichlet_node(int voxel_index) const;
    void set_substrate_dirichlet_activation(int substrate_index, bool new_value);
    void apply_dirichlet_boundaries_conditions();

    // default_microenvironment_options.track_internalized_substrates_in_each_agent
    // (BioFVM/BioFVM_basic_agent.cpp:520-532,553-556): agents' internalized_substrates
    // follow every simulate_cell_sources_and_sinks
    bool track_internalized_substrates_in_each_agent = false;

    // bulk sources and sinks (BioFVM/BioFVM_microenvironment.cpp:1327-1368):
    // the three rate callbacks (zero_function by default, :167-169) are
    // evaluated for every voxel on each call, as in the reference, and the
    // implicit update runs on the device
    void (*bulk_supply_rate_function)(Microenvironment*, int, double*) = zero_function;
    void (*bulk_supply_target_densities_function)(Microenvironment*, int, double*) = zero_function;
    void (*bulk_uptake_rate_function)(Microenvironment*, int, double*) = zero_function;
    void simulate_bulk_sources_and_sinks(double dt);

    // gradients (BioFVM/BioFVM_microenvironment.cpp:1430-1615): computed on the
    // device from the resident field; gradient_vector(n) points at the [ns][3]
    // block of voxel n and recomputes the field's gradients when it changed
    void compute_all_gradient_vectors();
    double* gradient_vector(int n);
    double* nearest_gradient_vector(const std::vector<double>& position);

    // stepping
    void simulate_diffusion_decay(double dt);
    void simulate_cell_sources_and_sinks(std::vector<Basic_Agent*>& agents, double dt);
    double total_mass(int substrate);

    // MATLAB level-4 output of every voxel: a (4 + ns) x nvox matrix named
    // multiscale_microenvironment with columns [x, y, z, dV, rho_0..rho_ns-1]
    // (BioFVM/BioFVM_microenvironment.cpp:1222-1250)
    void write_to_matlab(const std::string& filename);

    // device plumbing
    bfa_env* device();
    void download_if_needed();
    void mark_host_modified() { host_newer = true; }
    void invalidate_agents() { agents_sig.clear(); }
    std::string last_error;

  private:
    friend void diffusion_decay_solver__constant_coefficients_LOD_3D_GPU(Microenvironment&, double);
    friend void diffusion_decay_solver__constant_coefficients_LOD_3D_AVX256D_GPU(Microenvironment&, double);
    int ensure_device();
    int step_with(int variant, double dt);
    void ensure_planes();
    void ensure_flags();
    void write_dirichlet_value(int voxel_index, int substrate_index, double value);
    int push_state();
    std::vector<double> host_field;
    std::vector<double> gradient_field;  // [voxel][ns][3]
    bool gradients_valid = false;
    bfa_env* env = nullptr;
    bool host_newer = true;    // host copy was written since the last upload
    bool device_newer = false; // device copy was stepped since the last download
    bool state_dirty = true;   // coefficients / Dirichlet changed
    unsigned char face_registered[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> face_values;  // [6][ns]
    // per-voxel face values in the layout of bfa_set_dirichlet_planes, built
    // from the faces on the first per-node call (empty: the faces describe
    // every boundary voxel)
    std::vector<double> dirichlet_planes;
    std::vector<uint8_t> dirichlet_flags;  // is_Dirichlet per voxel, once any is set
    bool tracking_on_device = false;
    // agent state last sent to the device
    std::vector<const Basic_Agent*> agents_sig;
    std::vector<int32_t> sig_voxel;
    std::vector<uint8_t> sig_active;
    std::vector<double> sig_c1, sig_c2, sig_e2;
};

// MATLAB level-4 files (BioFVM/BioFVM_matlab.cpp): read_matlab returns the
// matrix as rows (output[row][col]; data types double, float, int32, int16,
// uint16, uint8; little-endian full real matrices only, as the reference);
// write_matlab4 writes a rows x cols double matrix from column-major data.
// On error they print a message and return an empty matrix / false.
std::vector<std::vector<double>> read_matlab(const std::string& filename);
bool write_matlab4(const std::string& filename, unsigned rows, unsigned cols,
                   const std::string& variable_name, const double* column_major);

// Initial condition from a .mat file holding one row per voxel,
// [x, y, z, rho_0, ..., rho_ns-1] (BioFVM/BioFVM_microenvironment.cpp:2341-2387).
// The reference exits on a malformed file; this prints the same message and
// returns false, leaving the field as it was.
bool load_initial_conditions_from_matlab(Microenvironment& M, const std::string& filename);

// Initial condition from a CSV file with one row per voxel, [x, y, z,
// rho_...] (BioFVM/BioFVM_microenvironment.cpp:2391-2480,2806-2840).  A first
// row "x,y,z,name,..." names the substrates of the following columns (each
// must exist); without it the columns after x, y, z are substrates 0, 1, ...
// Numbers are read as the reference's csv_to_vector reads them
// (BioFVM/BioFVM_vector.cpp:360-377).  Every voxel must be given exactly once.
// The reference exits on a malformed file; this prints the same messages and
// returns false, leaving the field as it was.  (The reference's own no-header
// path skips its column counter and rereads a closed stream, so it always
// stops with the voxel-count error; this reads the columns as its comments
// describe.)
bool load_initial_conditions_from_csv(Microenvironment& M, const std::string& filename);

// The drop-in solvers: assign one to Microenvironment::diffusion_decay_solver
// (the reference assigns diffusion_decay_solver__constant_coefficients_LOD_3D
// or _LOD_3D_BLOCKING at BioFVM/BioFVM_microenvironment.cpp:1943,2506, and the
// constructor defaults to _LOD_3D_AVX256D, :143).  _GPU reproduces
// LOD_3D_BLOCKING (Dirichlet after the z sweep too), _AVX256D_GPU the AVX256D
// solver (no Dirichlet pass after the z sweep).
void diffusion_decay_solver__constant_coefficients_LOD_3D_GPU(Microenvironment& M, double dt);
void diffusion_decay_solver__constant_coefficients_LOD_3D_AVX256D_GPU(Microenvironment& M, double dt);

}  // namespace BioFVM_amd
