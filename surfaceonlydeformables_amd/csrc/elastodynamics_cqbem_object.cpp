// CQ time step of one deformable body (src/elastodynamics_cqbem_object.cpp, src/physics_base_object.cpp:9-95).
// See elastodynamics_cqbem_object.hpp.
#include "elastodynamics_cqbem_object.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>

namespace cqbem {

namespace {

// 3 x 3 row-major helpers
void matmul3(const double *A, const double *B, double *C) {
    double T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
    std::copy(T, T + 9, C);
}
void matvec3(const double *A, const double *x, double *y) {
    double t[3];
    for (int i = 0; i < 3; i++) t[i] = A[3 * i] * x[0] + A[3 * i + 1] * x[1] + A[3 * i + 2] * x[2];
    std::copy(t, t + 3, y);
}
void matTvec3(const double *A, const double *x, double *y) {  // A^T x
    double t[3];
    for (int i = 0; i < 3; i++) t[i] = A[i] * x[0] + A[3 + i] * x[1] + A[6 + i] * x[2];
    std::copy(t, t + 3, y);
}
// y = M x for a row-major 3 x n matrix
void mul3n(const Dense &M, const double *x, double *y) {
    for (int a = 0; a < 3; a++) {
        double s = 0;
        for (int64_t c = 0; c < M.cols; c++) s += M(a, c) * x[c];
        y[a] = s;
    }
}
double det3(const double *m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}
void inv3(const double *m, double *r) {
    const double d = det3(m);
    r[0] = (m[4] * m[8] - m[5] * m[7]) / d, r[1] = (m[2] * m[7] - m[1] * m[8]) / d, r[2] = (m[1] * m[5] - m[2] * m[4]) / d;
    r[3] = (m[5] * m[6] - m[3] * m[8]) / d, r[4] = (m[0] * m[8] - m[2] * m[6]) / d, r[5] = (m[2] * m[3] - m[0] * m[5]) / d;
    r[6] = (m[3] * m[7] - m[4] * m[6]) / d, r[7] = (m[1] * m[6] - m[0] * m[7]) / d, r[8] = (m[0] * m[4] - m[1] * m[3]) / d;
}

// Orthogonal factor of the polar decomposition A = Q S (A nonsingular), by Newton's iteration
// Q <- (Q + Q^-T) / 2 (Higham 1986, with determinant scaling for the first steps).  For A = U S V^T
// it is U V^T: the same matrix the reference forms from a JacobiSVD (object.cpp:176-177, R = V U^T
// of F = U S V^T, i.e. the polar factor of F^T).
void polar_factor(const double *A, double *Q) {
    std::copy(A, A + 9, Q);
    for (int it = 0; it < 100; it++) {
        double Y[9];
        inv3(Q, Y);
        const double g = it < 6 ? std::cbrt(std::fabs(1.0 / det3(Q))) : 1.0;  // |det(gQ)| = 1 before averaging
        double N[9], diff = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                N[3 * i + j] = 0.5 * (g * Q[3 * i + j] + Y[3 * j + i] / g);
                diff = std::max(diff, std::fabs(N[3 * i + j] - Q[3 * i + j]));
            }
        std::copy(N, N + 9, Q);
        if (diff < 1e-15 && it >= 6) break;
    }
}

// igl::centroid: volume centroid of a closed mesh from the signed tetrahedra (origin, v0, v1, v2).
void volume_centroid(const std::vector<double> &V, const std::vector<int32_t> &F, double c[3]) {
    double vol = 0, s[3] = {0, 0, 0};
    for (size_t t = 0; t < F.size(); t += 3) {
        const double *a = &V[3 * F[t]], *b = &V[3 * F[t + 1]], *d = &V[3 * F[t + 2]];
        const double tv = (a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                           a[2] * (b[0] * d[1] - b[1] * d[0])) / 6.0;
        vol += tv;
        for (int k = 0; k < 3; k++) s[k] += tv * (a[k] + b[k] + d[k]) / 4.0;
    }
    for (int k = 0; k < 3; k++) c[k] = s[k] / vol;
}

void normalize_rows(std::vector<double> &X) {
    for (size_t i = 0; i < X.size(); i += 3) {
        const double z = X[i] * X[i] + X[i + 1] * X[i + 1] + X[i + 2] * X[i + 2];
        if (z > 0) {
            const double r = std::sqrt(z);
            X[i] /= r, X[i + 1] /= r, X[i + 2] /= r;
        }
    }
}

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

}  // namespace

void rotation_from_vector(const double w[3], double R[9]) {
    const double z = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double t = std::sqrt(z);
    double n[3] = {w[0], w[1], w[2]};
    if (z > 0) n[0] /= t, n[1] /= t, n[2] /= t;
    const double K[9] = {0, -n[2], n[1], n[2], 0, -n[0], -n[1], n[0], 0};
    double K2[9];
    matmul3(K, K, K2);
    const double s = std::sin(t), c1 = 1. - std::cos(t);
    for (int e = 0; e < 9; e++) R[e] = (e % 4 == 0 ? 1.0 : 0.0) + K[e] * s + K2[e] * c1;
}

ElastodynamicsCQBEMObject::ElastodynamicsCQBEMObject(std::shared_ptr<const ElastodynamicsCQBEMMatrixCollection> mc, double dt_,
                                                     const double g[3])
    : dt(dt_), mc_(std::move(mc)) {
    std::copy(g, g + 3, gravitational_constant);
    V_local = mc_->V;
    V_global = V_estimate_global = V_local;
}

ElastodynamicsCQBEMObject::~ElastodynamicsCQBEMObject() {
    for (double *q : {d_u_, d_p_, d_acc_, d_b_, d_x_, d_y_})
        if (q) (void)hipFree(q);
    if (d_refs_) (void)hipFree(d_refs_);
}

void ElastodynamicsCQBEMObject::product(const double *dense, const DeviceCsr *compressed, const std::vector<double> &x,
                                        std::vector<double> &y) {
    const ElastodynamicsCQBEMMatrixCollection &mc = *mc_;
    const auto &res = mc.resident();
    const int64_t n3 = 3 * mc.num_vertices;
    y.assign(n3, 0.0);
    if (!compressed) {
        if (hipMemcpy(d_x_, x.data(), sizeof(double) * n3, hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy x failed");
        if (cqbem_matvec(res.ctx, n3, dense, d_x_, d_y_) != CQBEM_OK) throw std::runtime_error(cqbem_last_error(res.ctx));
        if (hipMemcpy(y.data(), d_y_, sizeof(double) * n3, hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy y failed");
        return;
    }
    const std::vector<double> xc = wavelet_decompose_vec(x.data(), n3);
    std::vector<double> yc(xc.size());
    if (hipMemcpy(d_x_, xc.data(), sizeof(double) * xc.size(), hipMemcpyHostToDevice) != hipSuccess ||
        launch_csr_matvec(compressed->rowptr, compressed->col, compressed->val, d_x_, compressed->rows, d_y_, nullptr) !=
            hipSuccess ||
        hipMemcpy(yc.data(), d_y_, sizeof(double) * yc.size(), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("compressed product failed");
    y = wavelet_reconstruct(yc.data(), n3);
}

bool ElastodynamicsCQBEMObject::fail(const std::string &msg) {
    error_ = msg;
    return false;
}

bool ElastodynamicsCQBEMObject::init(const Json &config) {
    try {
        auto vec3 = [&](const char *key, double *out) {
            if (config.contains(key))
                for (int i = 0; i < 3; i++) out[i] = config[key].arr.at(i).number();
        };
        vec3("translational_vel", translational_vel);
        vec3("rotational_vel", rotational_vel);
        if (config.contains("rotation"))
            for (int e = 0; e < 9; e++) rotation[e] = config["rotation"].arr.at(e).number();
        if (config.contains("translation")) {
            vec3("translation", translation);
            double rc[3];
            matvec3(rotation, mc_->original_cm, rc);
            for (int k = 0; k < 3; k++) translation[k] += rc[k];
        }
        if (config.contains("use_gravity")) use_gravity = config["use_gravity"].boolean();
        if (config.contains("is_fixed")) is_fixed = config["is_fixed"].boolean();
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    // V_global = ((V_local - cm) R^T) + (translation + cm), cm = 0 for the centred mesh
    const int64_t nv = mc_->num_vertices;
    V_global.resize(3 * nv);
    for (int64_t v = 0; v < nv; v++) {
        double x[3];
        for (int k = 0; k < 3; k++) x[k] = V_local[3 * v + k] - mc_->cm[k];
        matvec3(rotation, x, &V_global[3 * v]);
        for (int k = 0; k < 3; k++) V_global[3 * v + k] += translation[k] + mc_->cm[k];
    }
    V_estimate_global = V_global;
    return true;
}

bool ElastodynamicsCQBEMObject::estimate_next_state() {
    const auto t_start = Clock::now();
    double t_parts = 0;
    has_constraints = false;
    const ElastodynamicsCQBEMMatrixCollection &mc = *mc_;
    const int64_t nv = mc.num_vertices, n3 = 3 * nv;
    const int32_t L = (int32_t)mc.max_time_history + 1;
    b.assign(n3, 0.0);
    try {
        const auto &res = mc.resident();
        const bool cmp = mc.use_compressed_matrices;
        if (!d_b_) {
            cap_ = 2 * std::max<int64_t>(1, L - 1);
            w_ = cmp ? res.vlen : n3;
            const int64_t vmax = std::max(n3, w_);
            if (hipMalloc(&d_u_, sizeof(double) * cap_ * w_) != hipSuccess ||
                hipMalloc(&d_p_, sizeof(double) * cap_ * w_) != hipSuccess ||
                hipMalloc(&d_acc_, sizeof(double) * cap_ * 6) != hipSuccess || hipMalloc(&d_b_, sizeof(double) * n3) != hipSuccess ||
                hipMalloc(&d_x_, sizeof(double) * vmax) != hipSuccess || hipMalloc(&d_y_, sizeof(double) * vmax) != hipSuccess ||
                hipMalloc(&d_refs_, sizeof(CsrRef) * 2 * std::max(1, L - 1)) != hipSuccess)
                return fail("hipMalloc of the step buffers failed");
        }
        // b = sum_{i=1..m} (B_i acc + H_i u + G_i p) over the newest m = min(count, L - 1) histories, H_i as the
        // object reads it (negated): cqbem_history subtracts the assembled H_i (object.cpp:42-49).  With
        // compressed matrices B stays dense and the H, G terms are summed in the compressed space, then
        // mapped back once (object.cpp:31-41).
        if (count_ > 0) {
            const auto t0 = Clock::now();
            if (cqbem_history(res.ctx, n3, L, cmp ? nullptr : res.G, cmp ? nullptr : res.H, res.B,
                              cmp ? nullptr : d_p_ + start_ * w_, cmp ? nullptr : d_u_ + start_ * w_, d_acc_ + start_ * 6,
                              (int32_t)count_, d_b_) != CQBEM_OK)
                return fail(cqbem_last_error(res.ctx));
            if (hipMemcpy(b.data(), d_b_, sizeof(double) * n3, hipMemcpyDeviceToHost) != hipSuccess)
                return fail("hipMemcpy b failed");
            if (cmp) {
                const int64_t m = std::min<int64_t>(count_, L - 1);
                std::vector<CsrRef> refs;
                for (int64_t i = 1; i <= m; i++) {
                    const int64_t row = start_ + count_ - i;
                    const DeviceCsr &Hc = res.Hc[i], &Gc = res.Gc[i];
                    refs.push_back({Hc.rowptr, Hc.col, Hc.val, d_u_ + row * w_});
                    refs.push_back({Gc.rowptr, Gc.col, Gc.val, d_p_ + row * w_});
                }
                std::vector<double> bc(w_);
                if (hipMemcpy(d_refs_, refs.data(), sizeof(CsrRef) * refs.size(), hipMemcpyHostToDevice) != hipSuccess ||
                    launch_csr_matvec_sum(d_refs_, (int)refs.size(), w_, d_y_, nullptr) != hipSuccess ||
                    hipMemcpy(bc.data(), d_y_, sizeof(double) * w_, hipMemcpyDeviceToHost) != hipSuccess)
                    return fail("compressed history sum failed");
                const std::vector<double> rec = wavelet_reconstruct(bc.data(), n3);
                for (int64_t r = 0; r < n3; r++) b[r] += rec[r];
            }
            const double t = since(t0);
            phase_ms[0] += t, t_parts += t;
        }
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    // without the local frame update, gravity and the translational fictitious force do not cancel (object.cpp:51-54)
    if (!mc.update_local_frame && use_gravity) {
        double rg[3];
        matTvec3(rotation, gravitational_constant, rg);
        const Dense &B0 = mc.B_matrices[0];
        for (int64_t r = 0; r < n3; r++) b[r] -= B0(r, 0) * rg[0] + B0(r, 1) * rg[1] + B0(r, 2) * rg[2];
    }
    p.assign(n3, 0.0);
    // u_unconstrained = H_0^-1 b (H_inv.solve(b), object.cpp:60), as a product with the resident inverse
    u_unconstrained.assign(n3, 0.0);
    try {
        const auto t0 = Clock::now();
        const auto &res = mc.resident();
        product(res.H_inv, mc.use_compressed_matrices ? &res.H_inv_c : nullptr, b, u_unconstrained);  // object.cpp:56-60
        const double t = since(t0);
        phase_ms[1] += t, t_parts += t;
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    // V_estimate_global (object.cpp:63-81)
    double Rm[9];
    double tr[3];
    if (is_fixed) {
        std::copy(rotation, rotation + 9, Rm);
        std::copy(translation, translation + 3, tr);
    } else {
        double w[3], Rw[9];
        for (int k = 0; k < 3; k++) w[k] = dt * rotational_vel[k];
        rotation_from_vector(w, Rw);
        matmul3(Rw, rotation, Rm);
        for (int k = 0; k < 3; k++) tr[k] = translation[k] + dt * translational_vel[k];
        if (use_gravity)
            for (int k = 0; k < 3; k++) tr[k] += dt * dt * gravitational_constant[k];
    }
    V_estimate_global.resize(n3);
    for (int64_t v = 0; v < nv; v++) {
        double x[3];
        for (int k = 0; k < 3; k++) x[k] = mc.V[3 * v + k] + u_unconstrained[3 * v + k];
        matvec3(Rm, x, &V_estimate_global[3 * v]);
        for (int k = 0; k < 3; k++) V_estimate_global[3 * v + k] += tr[k];
    }
    phase_ms[4] += since(t_start) - t_parts;
    return true;
}

void ElastodynamicsCQBEMObject::update_V_estimate_matrices() {
    const ElastodynamicsCQBEMMatrixCollection &mc = *mc_;
    const int64_t nv = mc.num_vertices, n3 = 3 * nv;
    if (is_fixed) {
        std::copy(rotation, rotation + 9, rotation_estimate);
    } else {
        double w[3], Rw[9];
        for (int k = 0; k < 3; k++) w[k] = dt * rotational_vel[k];
        rotation_from_vector(w, Rw);
        matmul3(Rw, rotation, rotation_estimate);
    }
    rot_trans_acc.assign(n3 * 3, 0.0);
    rot_rot_acc.assign(n3 * 3, 0.0);
    if (!is_fixed && mc.update_local_frame) {
        // dt^2 R and -dt^2 R [V_local,v]_x per vertex (V_local about the centroid at the origin)
        for (int64_t v = 0; v < nv; v++) {
            const double *x = &V_local[3 * v];
            const double X[9] = {0, -x[2], x[1], x[2], 0, -x[0], -x[1], x[0], 0};
            double RX[9];
            matmul3(rotation, X, RX);
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) {
                    rot_trans_acc[(3 * v + a) * 3 + c] = dt * dt * rotation[3 * a + c];
                    rot_rot_acc[(3 * v + a) * 3 + c] = -dt * dt * RX[3 * a + c];
                }
        }
    }
    V_const = V_estimate_global;
    if (!is_fixed && !mc.update_local_frame)
        for (int64_t v = 0; v < nv; v++)
            for (int k = 0; k < 3; k++) {
                V_const[3 * v + k] -= dt * translational_vel[k];
                if (use_gravity) V_const[3 * v + k] -= dt * dt * gravitational_constant[k];
            }
}

void ElastodynamicsCQBEMObject::push_history(const std::vector<double> &u, const std::vector<double> &pv, const double acc[6]) {
    const ElastodynamicsCQBEMMatrixCollection &mc = *mc_;
    const int64_t n3 = 3 * mc.num_vertices;
    const int64_t H = (int64_t)mc.max_time_history;
    if (H == 0) return;  // L = 1: no history term
    if ((int64_t)hist_acc_vectors.size() >= H) {
        hist_acc_vectors.pop_front(), hist_u_vectors.pop_front(), hist_p_vectors.pop_front();
    }
    hist_acc_vectors.emplace_back(acc, acc + 6);
    hist_u_vectors.push_back(u);
    hist_p_vectors.push_back(pv);
    // device rows: drop the oldest when full, slide the window back to row 0 at the buffer's end
    if (count_ >= H) start_++, count_--;
    if (start_ + count_ == cap_) {
        if (count_ > 0 &&
            (hipMemcpy(d_u_, d_u_ + start_ * w_, sizeof(double) * count_ * w_, hipMemcpyDeviceToDevice) != hipSuccess ||
             hipMemcpy(d_p_, d_p_ + start_ * w_, sizeof(double) * count_ * w_, hipMemcpyDeviceToDevice) != hipSuccess ||
             hipMemcpy(d_acc_, d_acc_ + start_ * 6, sizeof(double) * count_ * 6, hipMemcpyDeviceToDevice) != hipSuccess))
            throw std::runtime_error("history slide failed");
        start_ = 0;
    }
    const int64_t row = start_ + count_;
    // compressed matrices keep compressed histories (object.cpp:216-222)
    const std::vector<double> uc = mc.use_compressed_matrices ? wavelet_decompose_vec(u.data(), n3) : u;
    const std::vector<double> pc = mc.use_compressed_matrices ? wavelet_decompose_vec(pv.data(), n3) : pv;
    if (hipMemcpy(d_u_ + row * w_, uc.data(), sizeof(double) * w_, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_p_ + row * w_, pc.data(), sizeof(double) * w_, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_acc_ + row * 6, acc, sizeof(double) * 6, hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("history upload failed");
    count_++;
}

bool ElastodynamicsCQBEMObject::confirm_next_state() {
    const auto t_start = Clock::now();
    double t_parts = 0;
    const ElastodynamicsCQBEMMatrixCollection &mc = *mc_;
    const int64_t nv = mc.num_vertices, n3 = 3 * nv;
    if (!d_b_) return fail("confirm_next_state before estimate_next_state");
    if ((int64_t)p.size() != n3) return fail("p must hold 3 nv tractions");
    std::vector<double> u(n3), up(n3, 0.0);
    try {
        if (has_constraints) {
            // u_p = u_p_mat p on the device (object.cpp:136); V_global from the constraint solve's linearisation
            const auto &res = mc.resident();
            const auto t0 = Clock::now();
            product(res.u_p, mc.use_compressed_matrices ? &res.u_p_c : nullptr, p, up);  // object.cpp:131-136
            const double t = since(t0);
            phase_ms[2] += t, t_parts += t;
            for (int64_t r = 0; r < n3; r++) u[r] = up[r] + u_unconstrained[r];
            V_global.resize(n3);
            for (int64_t v = 0; v < nv; v++) {
                matvec3(rotation_estimate, &up[3 * v], &V_global[3 * v]);
                for (int k = 0; k < 3; k++) V_global[3 * v + k] += V_const[3 * v + k];
            }
            if (!is_fixed && mc.update_local_frame) {
                double ta[3], ra[3];
                mul3n(mc.translational_acc_mat, p.data(), ta);
                mul3n(mc.rotational_acc_mat, p.data(), ra);
                for (int64_t r = 0; r < n3; r++)
                    for (int c = 0; c < 3; c++) V_global[r] += rot_trans_acc[3 * r + c] * ta[c] + rot_rot_acc[3 * r + c] * ra[c];
            }
        } else {
            u = u_unconstrained;
        }
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    for (int64_t r = 0; r < n3; r++) V_local[r] = mc.V[r] + u[r];

    double tacc[3] = {0, 0, 0}, racc[3] = {0, 0, 0};
    auto set_global_from_local = [&]() {
        V_global.resize(n3);
        for (int64_t v = 0; v < nv; v++) {
            matvec3(rotation, &V_local[3 * v], &V_global[3 * v]);
            for (int k = 0; k < 3; k++) V_global[3 * v + k] += translation[k];
        }
    };
    if (mc.update_local_frame) {
        double old_rotation[9];
        std::copy(rotation, rotation + 9, old_rotation);
        // the new frame (object.cpp:155-167)
        mul3n(mc.translational_acc_mat, p.data(), tacc);
        mul3n(mc.rotational_acc_mat, p.data(), racc);
        if (!is_fixed) {
            double t[3];
            matvec3(rotation, tacc, t);
            for (int k = 0; k < 3; k++) translational_vel[k] += dt * t[k] + (use_gravity ? dt * gravitational_constant[k] : 0.0);
            matvec3(rotation, racc, t);
            for (int k = 0; k < 3; k++) rotational_vel[k] += dt * t[k];
            for (int k = 0; k < 3; k++) translation[k] += dt * translational_vel[k];
            double w[3], Rw[9];
            for (int k = 0; k < 3; k++) w[k] = dt * rotational_vel[k];
            rotation_from_vector(w, Rw);
            matmul3(Rw, rotation, rotation);
        }
        if (!has_constraints) set_global_from_local();

        // drift elimination (object.cpp:171-211)
        double cmv[3];
        volume_centroid(V_local, mc.F, cmv);
        std::vector<double> Vv(n3), V0(n3);
        for (int64_t v = 0; v < nv; v++)
            for (int k = 0; k < 3; k++) Vv[3 * v + k] = V_local[3 * v + k] - cmv[k], V0[3 * v + k] = mc.V[3 * v + k] - mc.cm[k];
        std::vector<double> Vn = Vv, V0n = V0;
        normalize_rows(Vn), normalize_rows(V0n);
        double Fm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // init_V_vec^T diag(w) V_vec
        for (int64_t v = 0; v < nv; v++) {
            const double w = mc.rotation_removal_weights[v];
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) Fm[3 * a + c] += V0n[3 * v + a] * w * Vn[3 * v + c];
        }
        double FmT[9], Rr[9];
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) FmT[3 * a + c] = Fm[3 * c + a];
        polar_factor(FmT, Rr);  // V U^T of Fm = U S V^T
        std::vector<double> du(n3);
        for (int64_t r = 0; r < n3; r++) du[r] = Vv[r] - V0[r];
        double tad[3], t1[3], t2[3];
        mul3n(mc.B_trans_inv_H, du.data(), tad);
        mul3n(mc.B_trans_inv_G_prime, p.data(), t1);
        const std::vector<double> t3 = mc.B_trans_inv.solve(b.data(), 1);
        for (int k = 0; k < 3; k++) tad[k] = tad[k] - t1[k] - t3[k];
        for (int64_t v = 0; v < nv; v++) {
            double x[3];
            matTvec3(Rr, &Vv[3 * v], x);  // row (V_local,v - cm) R
            for (int k = 0; k < 3; k++) du[3 * v + k] = x[k] - V0[3 * v + k];
        }
        double aad[3];
        mul3n(mc.B_euler_inv_H, du.data(), aad);
        mul3n(mc.B_euler_inv_G_prime, p.data(), t1);
        const std::vector<double> t4 = mc.B_euler_inv.solve(b.data(), 1);
        matvec3(mc.B_euler_inv_B_trans.data.data(), tad, t2);
        for (int k = 0; k < 3; k++) aad[k] = aad[k] - t1[k] - t4[k] - t2[k];
        for (int64_t r = 0; r < n3; r++) {
            double s = up[r] + u_unconstrained[r];
            for (int k = 0; k < 3; k++) s += mc.H_inv_B_trans(r, k) * tad[k] + mc.H_inv_B_euler(r, k) * aad[k];
            u[r] = s;
        }
        for (int k = 0; k < 3; k++) tacc[k] += tad[k], racc[k] += aad[k];
        if (!is_fixed) {
            double t[3];
            matvec3(old_rotation, tad, t);
            for (int k = 0; k < 3; k++) {
                translational_vel[k] += dt * t[k];
                translation[k] += dt * dt * t[k];
            }
            matvec3(old_rotation, aad, t);
            for (int k = 0; k < 3; k++) rotational_vel[k] += dt * t[k];
            double w[3], Rw[9];
            for (int k = 0; k < 3; k++) w[k] = dt * rotational_vel[k];
            rotation_from_vector(w, Rw);
            matmul3(Rw, old_rotation, rotation);
        }
        for (int64_t r = 0; r < n3; r++) V_local[r] = mc.V[r] + u[r];
    } else {
        if (use_gravity) {
            double t[3];
            matTvec3(rotation, gravitational_constant, t);
            for (int k = 0; k < 3; k++) tacc[k] = -t[k];
        }
        if (!has_constraints) set_global_from_local();
    }
    const double acc[6] = {tacc[0], tacc[1], tacc[2], racc[0], racc[1], racc[2]};
    try {
        const auto t0 = Clock::now();
        push_history(u, p, acc);
        const double t = since(t0);
        phase_ms[3] += t, t_parts += t;
    } catch (const std::exception &e) {
        return fail(e.what());
    }
    phase_ms[4] += since(t_start) - t_parts;
    return true;
}

}  // namespace cqbem
