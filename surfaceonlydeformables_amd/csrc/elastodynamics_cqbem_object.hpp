// C++ counterpart of the reference's ElastodynamicsCQBEMObject (include/elastodynamics_cqbem_object.hpp,
// src/elastodynamics_cqbem_object.cpp) with the PhysicsBaseObject state it steps
// (src/physics_base_object.cpp:9-95): the CQ time step of one deformable body on the matrices of an
// ElastodynamicsCQBEMMatrixCollection, uncompressed matrices.
//
// Where the work runs: the history sum b (2 (L-1) dense 3nv x 3nv products, the step's HBM-bound part)
// runs on the GPU on the collection's resident matrices, with the u / p / acceleration histories kept
// in device memory; the H_0 solve is a product with the resident H_0^-1 (made once from the LU factors:
// one right-hand side's triangular solves are latency-bound); u_p p, when there are constraints, is a
// device product.  The 3-vectors, the frame update and the drift elimination stay on the host.
//
// Contacts are not part of this class.  A caller with a constraint solver does what the reference's
// Simulator does (simulator.cpp:160-181): estimate_next_state(), update_V_estimate_matrices(), then
// sets p (nodal tractions, 3 nv) and has_constraints = true, and calls confirm_next_state().
#pragma once
#include <cstdint>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "elastodynamics_cqbem_matrix_collection.hpp"

namespace cqbem {

class ElastodynamicsCQBEMObject {
  public:
    ElastodynamicsCQBEMObject(std::shared_ptr<const ElastodynamicsCQBEMMatrixCollection> matrix_collection, double dt,
                              const double gravitational_constant[3]);
    ~ElastodynamicsCQBEMObject();
    ElastodynamicsCQBEMObject(const ElastodynamicsCQBEMObject &) = delete;
    ElastodynamicsCQBEMObject &operator=(const ElastodynamicsCQBEMObject &) = delete;

    // Object keys of a scene (physics_base_object.cpp:16-45): translational_vel, rotational_vel,
    // rotation (row-major 3 x 3), translation, use_gravity, is_fixed.
    bool init(const Json &config);
    // The unconstrained prediction: b, u_unconstrained, V_estimate_global (object.cpp:25-82).
    bool estimate_next_state();
    // V_u (the estimated rotation per vertex), rot_trans_acc, rot_rot_acc and V_const (object.cpp:84-124).
    void update_V_estimate_matrices();
    // The step's result from p: u, frame update, drift elimination, histories (object.cpp:126-238).
    bool confirm_next_state();

    const std::string &last_error() const { return error_; }

    // --- state (physics_base_object.hpp:68-90, elastodynamics_cqbem_object.hpp) ---
    double dt;
    double gravitational_constant[3];
    bool use_gravity = true, is_fixed = false, has_constraints = false;
    double rotation[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};  // row-major
    double translation[3] = {0, 0, 0}, translational_vel[3] = {0, 0, 0}, rotational_vel[3] = {0, 0, 0};
    std::vector<double> V_local, V_global, V_estimate_global;  // nv x 3 row-major
    std::vector<double> b, p, u_unconstrained;                 // 3 nv
    // update_V_estimate_matrices: V_u is block-diagonal with rotation_estimate in every block
    double rotation_estimate[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::vector<double> rot_trans_acc, rot_rot_acc;  // 3 nv x 3 row-major
    std::vector<double> V_const;                     // 3 nv
    // histories, oldest first, at most max_time_history of each (host mirrors of the device copies)
    std::deque<std::vector<double>> hist_u_vectors, hist_p_vectors;
    std::deque<std::vector<double>> hist_acc_vectors;  // 6 each: translational, rotational acceleration
    // accumulated wall time (ms) of the step's parts: [0] history sum, [1] H_0 solve, [2] u_p p,
    // [3] history upload, [4] everything else (host)
    double phase_ms[5] = {0, 0, 0, 0, 0};

  private:
    bool fail(const std::string &msg);
    void push_history(const std::vector<double> &u, const std::vector<double> &pv, const double acc[6]);
    // y = M x through the resident matrix: dense (cqbem_matvec), or compressed: reconstruct(M_c decompose(x))
    // (CompressedMatrixVectorProduct, wavelets.hpp:326-370).  Throws on errors.
    void product(const double *dense, const DeviceCsr *compressed, const std::vector<double> &x, std::vector<double> &y);

    std::shared_ptr<const ElastodynamicsCQBEMMatrixCollection> mc_;
    std::string error_;
    // device histories: rows [start_, start_ + count_) of 2 (L - 1)-row buffers, oldest first
    // (compressed vectors of 2 np - 2 entries with compressed matrices: w_ is the row width)
    double *d_u_ = nullptr, *d_p_ = nullptr, *d_acc_ = nullptr, *d_b_ = nullptr, *d_x_ = nullptr, *d_y_ = nullptr;
    CsrRef *d_refs_ = nullptr;
    int64_t cap_ = 0, start_ = 0, count_ = 0, w_ = 0;
};

// Rodrigues' formula for the rotation by |w| about w / |w| (physics_base_object.cpp:85-95); identity for w = 0.
void rotation_from_vector(const double w[3], double R[9]);

}  // namespace cqbem
