// C++ facade with the interface of the reference's ElastodynamicsCQBEMMatrixCollection
// (include/elastodynamics_cqbem_matrix_collection.hpp:10-57): same constructor, bool init(config),
// and the members the time-stepping object reads (G_matrices, H_matrices negated, B_matrices,
// max_time_history).  The precompute runs on the GPU through include/cqbem.h.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cqbem.h"
#include "cqbem_dense.hpp"
#include "cqbem_wavelet.hpp"

namespace cqbem {

// Minimal JSON value (objects, arrays, numbers, strings, booleans, null) for scene files.
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    bool integer = false;  // written without fraction or exponent (then inum holds it exactly)
    int64_t inum = 0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;  // keys in byte order, as the reference's JSON library keeps them

    static Json parse(const std::string &text);  // throws std::runtime_error on malformed input
    bool contains(const std::string &k) const { return kind == Object && obj.count(k) != 0; }
    const Json &operator[](const std::string &k) const;
    double number() const;
    const std::string &string() const;
    bool boolean() const;
    // Serialisation in the reference's JSON library format: compact (indent < 0) or pretty-printed.
    std::string dump(int indent = -1) const;
};

// Key of the reference's matrix cache for one object group (matrix_collection.cpp:41-57): the group's
// JSON without simulation.compression_ratio and simulation.update_local_frame, compact dump, std::hash,
// upper-case hex.  `stripped` receives that JSON.
std::string cache_key(const Json &group, Json *stripped = nullptr);

// Row-major dense double matrix (the layout of the reference's MatrixXs, type_declaration.hpp:16-32).
struct Dense {
    int64_t rows = 0, cols = 0;
    std::vector<double> data;
    Dense() = default;
    Dense(int64_t r, int64_t c) : rows(r), cols(c), data((size_t)(r * c), 0.0) {}
    double &operator()(int64_t r, int64_t c) { return data[(size_t)(r * cols + c)]; }
    double operator()(int64_t r, int64_t c) const { return data[(size_t)(r * cols + c)]; }
};

// The reference's .mat format (include/matrix_io.hpp:13-61): int64 rows, int64 cols, then the entries in
// row-major order (MatrixXs is RowMajor, type_declaration.hpp:18,32).
bool save_matrix(const Dense &m, const std::string &path);
bool load_matrix(Dense &m, const std::string &path);

enum MultistepMethod { BDF1, BDF2 };

class ElastodynamicsCQBEMMatrixCollection {
  public:
    ElastodynamicsCQBEMMatrixCollection(double dt, bool enable_cuda, unsigned int cuda_thread_per_block);
    ~ElastodynamicsCQBEMMatrixCollection();

    // Reads mesh (OBJ file, scale, enable_traction_discontinuity with sharp_edge_threshold_degree),
    // material (rho, mu, nu) and simulation (dt, gaussian_quadrature_order, quadrature_subdivision,
    // multistep_method, use_galerkin, coeffs_folder_path: the reference's on-disk G/H cache) like the
    // reference; optional simulation.cq_length / simulation.cq_radius override the mesh-derived L and R.
    bool init(const Json &config);
    // Same, with an in-memory mesh instead of config["mesh"].
    bool init(const Json &config, const std::vector<double> &V, const std::vector<int32_t> &F);

    const std::string &last_error() const { return error_; }

    // --- members consumed by ElastodynamicsCQBEMObject (reference hpp:40-51, physics_base hpp:29-50) ---
    std::vector<Dense> G_matrices, H_matrices;  // L x (3 nv x 3 nv); H negated as matrix_collection.cpp:140
    std::vector<Dense> B_matrices;              // L x (3 nv x 6) = [B_trans | B_euler]
    size_t max_time_history = 0;
    int64_t num_vertices = 0;
    std::vector<double> V;  // centred vertices, nv x 3
    std::vector<int32_t> F;
    std::vector<int32_t> vertex_map_inverse;  // split vertex -> input vertex (empty without the split)
    double cm[3] = {0, 0, 0};           // centroid of the centred mesh: zero (physics_base_matrix_collection.cpp:45)
    double original_cm[3] = {0, 0, 0};  // volume centroid of the mesh as loaded (:43)
    double mass = 0.0;
    double I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // inertia tensor about the centroid, row-major
    double c1 = 0, c2 = 0, rho = 0, mu = 0, nu = 0, R = 0;
    bool use_galerkin = false;
    bool update_local_frame = true;
    // simulation.compression_ratio (elastics_bem_matrix_collection.cpp:22-24): the time stepping then uses
    // wavelet-compressed G_n, H_n, u_p and H_0^-1 keeping (int)(9 nv^2 ratio) coefficients each
    double compression_ratio = 0.0;
    bool use_compressed_matrices = false;
    MultistepMethod multistep_method = BDF2;
    double last_assembly_ms = 0.0, last_factor_ms = 0.0;
    // last_factor_ms by part: [0] geometry, inertia and acceleration matrices, [1] G', [2] the least-squares
    // products, [3] the LU of H_0 (GPU), [4] its solves for u_p, H_inv_B_trans and H_inv_B_euler
    double post_process_phase_ms[5] = {0, 0, 0, 0, 0};
    bool loaded_from_cache = false;  // G and H came from simulation.coeffs_folder_path
    // compressed matrices read from .cpmat files of the cache folder by resident() (the rest were compressed
    // on the GPU and written there)
    mutable int compressed_loaded = 0;
    std::string cache_folder;  // elastodynamicsCQBEM_<key> under simulation.coeffs_folder_path, or empty

    // --- after the assembly (matrix_collection.cpp:174-215, physics_base_matrix_collection.cpp:40-50) ---
    Dense translational_acc_mat, rotational_acc_mat;  // 3 x 3 nv
    DeviceLU H_inv;                                   // LU of H_0 (un-negated, with C), on the GPU
    ThinLeastSquares B_trans_inv, B_euler_inv;        // of B_0's two 3 nv x 3 halves
    Dense B_trans_inv_H, B_euler_inv_H, B_trans_inv_G_prime, B_euler_inv_G_prime;  // 3 x 3 nv
    Dense B_euler_inv_B_trans;                                                    // 3 x 3
    Dense H_inv_B_trans, H_inv_B_euler;                                           // 3 nv x 3
    Dense u_p;       // H_0^-1 G' (3 nv x 3 nv), G' = G_0 (+ B_trans T + B_euler R with the local frame update)
    Dense u_p_diag;  // nv x 9: the 3 x 3 diagonal blocks of u_p
    std::vector<double> rotation_removal_weights;  // nv

    // Device copies for time stepping on the GPU, on the device of the collection's context: G_n and H_n
    // (un-negated, as cqbem_assemble writes them, for cqbem_history), B_n, u_p, and H_0^-1 (row-major,
    // from the LU factors: one step's solve becomes a product).  With use_compressed_matrices, G_n, H_n
    // (n >= 1, H negated as the object reads it), u_p and H_0^-1 are held wavelet-compressed instead
    // (matrix_collection.cpp:97-134,229-271) and the dense G, H, u_p, H_inv stay null.  Made on the first
    // call, freed with the collection; throws std::runtime_error when that fails.
    struct Resident {
        cqbem_context *ctx = nullptr;
        double *G = nullptr, *H = nullptr, *B = nullptr, *u_p = nullptr, *H_inv = nullptr;
        std::vector<DeviceCsr> Gc, Hc;  // [n], n = 1 .. L-1 (entry 0 unused)
        DeviceCsr u_p_c, H_inv_c;
        int64_t vlen = 0;  // length of a compressed vector, 2 np - 2
    };
    const Resident &resident() const;

  private:
    bool init_mesh(const Json &config, std::vector<double> v, std::vector<int32_t> f, std::vector<int32_t> vmap);
    bool fail(const std::string &msg);
    bool run(const Json &config);
    bool post_process();

    mutable std::mutex resident_mu_;
    mutable Resident resident_;
    mutable bool resident_ready_ = false;
    void free_resident();
    // G_n, H_n (un-negated) and B_n as cqbem_assemble wrote them on the device during init, handed to resident()
    // so that the time stepping does not upload them again (dense matrices), or compressed from there.  Null when
    // init assembled through host memory (the device buffers could not be allocated) or read G and H from the
    // cache folder.
    mutable double *assembled_G_ = nullptr, *assembled_H_ = nullptr, *assembled_B_ = nullptr;
    void free_assembled() const;

    double dt_;
    bool enable_cuda_;
    unsigned int cuda_thread_per_block_;
    std::string error_;
};

// Traction-discontinuity vertex split (mesh_split.cpp); returns a cqbem status.
int split_traction_discontinuity(const double *V, int nv, const int32_t *F, int nf, double threshold_deg,
                                 std::vector<double> &V_out, std::vector<int32_t> &F_out, std::vector<int32_t> &vmap_out,
                                 std::string &err);

// Subdivision and decimation (mesh_refine.cpp); return a cqbem status.  V and F are replaced by the result.
int subdivide_mesh(std::vector<double> &V, std::vector<int32_t> &F, int scheme, int levels, std::string &err);
int decimate_mesh(std::vector<double> &V, std::vector<int32_t> &F, int64_t max_faces, std::string &err);

// The reference's compressed-matrix files (include/wavelets.hpp:451-504, SparseMatrix row-major, int32 indices):
// int N; int64 rows, cols, nnz, outerSize, innerSize; nnz doubles; outerSize row offsets (the last, nnz, is not
// stored); nnz column indices.  save_cpmat copies the device CSR to the host; load_cpmat checks the layout
// against an n x n compression and uploads it (false when the file is missing or does not match).
bool save_cpmat(const DeviceCsr &m, const std::string &path);
bool load_cpmat(DeviceCsr &m, int64_t n, const std::string &path);

// Loads an OBJ (v / f lines, polygons fanned into triangles).
bool load_obj(const std::string &path, std::vector<double> &V, std::vector<int32_t> &F, std::string &err);

// The mesh an elastodynamicsCQBEM object is assembled on, from the "mesh" block of its config
// (PhysicsBaseMatrixCollection::load_mesh, physics_base_matrix_collection.cpp:59-365, with vertex
// reordering off): OBJ loaded, unused vertices dropped, scaled, normals flipped on request, and
// split along sharp edges when enable_traction_discontinuity is set (vmap then maps every vertex
// to its original; empty otherwise).  subdivision_level > 0 and decimation_scale < 1 are rejected.
bool prepare_mesh(const Json &mesh, std::vector<double> &V, std::vector<int32_t> &F, std::vector<int32_t> &vmap,
                  std::string &err);

}  // namespace cqbem
