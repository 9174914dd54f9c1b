// C++ facade with the interface of the reference's ElastodynamicsCQBEMMatrixCollection
// (include/elastodynamics_cqbem_matrix_collection.hpp:10-57): same constructor, bool init(config),
// and the members the time-stepping object reads (G_matrices, H_matrices negated, B_matrices,
// max_time_history).  The precompute runs on the GPU through include/cqbem.h.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cqbem.h"

namespace cqbem {

// Minimal JSON value (objects, arrays, numbers, strings, booleans, null) for scene files.
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    static Json parse(const std::string &text);  // throws std::runtime_error on malformed input
    bool contains(const std::string &k) const { return kind == Object && obj.count(k) != 0; }
    const Json &operator[](const std::string &k) const;
    double number() const;
    const std::string &string() const;
    bool boolean() const;
};

// Row-major dense double matrix (the layout of the reference's MatrixXs, type_declaration.hpp:16-32).
struct Dense {
    int64_t rows = 0, cols = 0;
    std::vector<double> data;
    Dense() = default;
    Dense(int64_t r, int64_t c) : rows(r), cols(c), data((size_t)(r * c), 0.0) {}
    double &operator()(int64_t r, int64_t c) { return data[(size_t)(r * cols + c)]; }
    double operator()(int64_t r, int64_t c) const { return data[(size_t)(r * cols + c)]; }
};

enum MultistepMethod { BDF1, BDF2 };

class ElastodynamicsCQBEMMatrixCollection {
  public:
    ElastodynamicsCQBEMMatrixCollection(double dt, bool enable_cuda, unsigned int cuda_thread_per_block);
    ~ElastodynamicsCQBEMMatrixCollection();

    // Reads mesh (OBJ file, scale), material (rho, mu, nu) and simulation (dt, gaussian_quadrature_order,
    // quadrature_subdivision, multistep_method, use_galerkin) like the reference; optional
    // simulation.cq_length / simulation.cq_radius override the mesh-derived L and R.
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
    double cm[3] = {0, 0, 0};
    double mass = 0.0;
    double c1 = 0, c2 = 0, rho = 0, mu = 0, nu = 0, R = 0;
    bool use_galerkin = false;
    MultistepMethod multistep_method = BDF2;
    double last_assembly_ms = 0.0;

  private:
    bool fail(const std::string &msg);
    bool run(const Json &config);

    double dt_;
    bool enable_cuda_;
    unsigned int cuda_thread_per_block_;
    std::string error_;
};

// Loads an OBJ (v / f lines, polygons fanned into triangles).
bool load_obj(const std::string &path, std::vector<double> &V, std::vector<int32_t> &F, std::string &err);

}  // namespace cqbem
