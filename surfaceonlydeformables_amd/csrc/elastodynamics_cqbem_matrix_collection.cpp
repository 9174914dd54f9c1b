// C++ facade over include/cqbem.h with the reference's collection interface
// (src/elastodynamics_cqbem_matrix_collection.cpp:13-278, src/physics_base_matrix_collection.cpp:32-57,
// src/elastics_bem_matrix_collection.cpp:16-41).
#include "elastodynamics_cqbem_matrix_collection.hpp"

#include "cqbem_rules.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <stdexcept>

namespace cqbem {

// ----------------------------------------------------------------- JSON --
namespace {

struct Parser {
    const std::string &s;
    size_t p = 0;
    explicit Parser(const std::string &t) : s(t) {}
    [[noreturn]] void err(const char *what) { throw std::runtime_error(std::string("json: ") + what + " at offset " + std::to_string(p)); }
    void ws() {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\n' || s[p] == '\t' || s[p] == '\r')) p++;
    }
    bool lit(const char *w) {
        const size_t n = std::char_traits<char>::length(w);
        if (s.compare(p, n, w) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    std::string str() {
        if (s[p] != '"') err("expected string");
        p++;
        std::string out;
        while (p < s.size() && s[p] != '"') {
            if (s[p] == '\\') {
                p++;
                if (p >= s.size()) err("bad escape");
                const char c = s[p];
                out += c == 'n' ? '\n' : c == 't' ? '\t' : c;
            } else {
                out += s[p];
            }
            p++;
        }
        if (p >= s.size()) err("unterminated string");
        p++;
        return out;
    }
    Json value() {
        ws();
        if (p >= s.size()) err("unexpected end");
        Json j;
        const char c = s[p];
        if (c == '{') {
            j.kind = Json::Object;
            p++;
            ws();
            if (s[p] == '}') return p++, j;
            for (;;) {
                ws();
                std::string k = str();
                ws();
                if (s[p++] != ':') err("expected ':'");
                j.obj[k] = value();
                ws();
                if (s[p] == ',') {
                    p++;
                    continue;
                }
                if (s[p] == '}') return p++, j;
                err("expected ',' or '}'");
            }
        }
        if (c == '[') {
            j.kind = Json::Array;
            p++;
            ws();
            if (s[p] == ']') return p++, j;
            for (;;) {
                j.arr.push_back(value());
                ws();
                if (s[p] == ',') {
                    p++;
                    continue;
                }
                if (s[p] == ']') return p++, j;
                err("expected ',' or ']'");
            }
        }
        if (c == '"') {
            j.kind = Json::String;
            j.str = str();
            return j;
        }
        if (lit("true")) return j.kind = Json::Bool, j.b = true, j;
        if (lit("false")) return j.kind = Json::Bool, j.b = false, j;
        if (lit("null")) return j;
        char *end = nullptr;
        j.num = std::strtod(s.c_str() + p, &end);
        if (end == s.c_str() + p) err("bad value");
        p = end - s.c_str();
        j.kind = Json::Number;
        return j;
    }
};

}  // namespace

Json Json::parse(const std::string &text) {
    Parser ps(text);
    Json j = ps.value();
    ps.ws();
    if (ps.p != text.size()) ps.err("trailing characters");
    return j;
}

const Json &Json::operator[](const std::string &k) const {
    static const Json null_value;
    if (kind != Object) throw std::runtime_error("json: not an object when reading '" + k + "'");
    auto it = obj.find(k);
    if (it == obj.end()) throw std::runtime_error("json: missing key '" + k + "'");
    return it->second;
    (void)null_value;
}
double Json::number() const {
    if (kind != Number) throw std::runtime_error("json: expected a number");
    return num;
}
const std::string &Json::string() const {
    if (kind != String) throw std::runtime_error("json: expected a string");
    return str;
}
bool Json::boolean() const {
    if (kind != Bool) throw std::runtime_error("json: expected a boolean");
    return b;
}

// ------------------------------------------------------------------- OBJ --
bool load_obj(const std::string &path, std::vector<double> &V, std::vector<int32_t> &F, std::string &err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open mesh file " + path;
        return false;
    }
    V.clear(), F.clear();
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        ls >> tag;
        if (tag == "v") {
            double x, y, z;
            ls >> x >> y >> z;
            V.insert(V.end(), {x, y, z});
        } else if (tag == "f") {
            std::vector<int> idx;
            std::string tok;
            while (ls >> tok) {
                const int k = std::stoi(tok.substr(0, tok.find('/')));
                idx.push_back(k > 0 ? k - 1 : (int)(V.size() / 3) + k);
            }
            for (size_t k = 1; k + 1 < idx.size(); k++) F.insert(F.end(), {idx[0], idx[k], idx[k + 1]});
        }
    }
    if (V.empty() || F.empty()) {
        err = "empty mesh in " + path;
        return false;
    }
    return true;
}

// ------------------------------------------------------------- collection --
namespace {
std::mutex g_ctx_mu;
cqbem_context *g_ctx = nullptr;  // one GPU context per process, shared by all collections

cqbem_context *context(std::string &err) {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (!g_ctx && cqbem_create(&g_ctx, 0, 0) != CQBEM_OK) {
        g_ctx = nullptr;
        err = "no HIP device available for the CQBEM precompute";
    }
    return g_ctx;
}
}  // namespace

ElastodynamicsCQBEMMatrixCollection::ElastodynamicsCQBEMMatrixCollection(double dt, bool enable_cuda,
                                                                         unsigned int cuda_thread_per_block)
    : dt_(dt), enable_cuda_(enable_cuda), cuda_thread_per_block_(cuda_thread_per_block) {}

ElastodynamicsCQBEMMatrixCollection::~ElastodynamicsCQBEMMatrixCollection() = default;

bool ElastodynamicsCQBEMMatrixCollection::fail(const std::string &msg) {
    error_ = msg;
    return false;
}

bool ElastodynamicsCQBEMMatrixCollection::init(const Json &config) {
    try {
        const Json &m = config["mesh"];
        std::vector<double> v;
        std::vector<int32_t> f;
        if (!load_obj(m["file"].string(), v, f, error_)) return false;
        if (m.contains("scale"))
            for (double &x : v) x *= m["scale"].number();
        if (m.contains("enable_traction_discontinuity") && m["enable_traction_discontinuity"].boolean()) {
            // physics_base_matrix_collection.cpp:68-71,203-365 (vertex reordering off)
            const double deg = m.contains("sharp_edge_threshold_degree") ? m["sharp_edge_threshold_degree"].number() : 20.0;
            std::vector<double> vs;
            std::vector<int32_t> fs, vm;
            std::string err;
            if (split_traction_discontinuity(v.data(), (int)(v.size() / 3), f.data(), (int)(f.size() / 3), deg, vs, fs, vm,
                                             err) != CQBEM_OK)
                return fail("traction-discontinuity split: " + err);
            return init_mesh(config, std::move(vs), std::move(fs), std::move(vm));
        }
        return init_mesh(config, std::move(v), std::move(f), {});
    } catch (const std::exception &e) {
        return fail(e.what());
    }
}

bool ElastodynamicsCQBEMMatrixCollection::init(const Json &config, const std::vector<double> &v,
                                               const std::vector<int32_t> &f) {
    return init_mesh(config, v, f, {});
}

bool ElastodynamicsCQBEMMatrixCollection::init_mesh(const Json &config, std::vector<double> v, std::vector<int32_t> f,
                                                    std::vector<int32_t> vmap) {
    try {
        V = std::move(v), F = std::move(f), vertex_map_inverse = std::move(vmap);
        num_vertices = (int64_t)(V.size() / 3);
        return run(config);
    } catch (const std::exception &e) {
        return fail(e.what());
    }
}

bool ElastodynamicsCQBEMMatrixCollection::run(const Json &config) {
    if (!enable_cuda_) return fail("the CQBEM precompute of this build runs on the GPU only (enable_cuda = false)");
    const Json &mat = config["material"], &sim = config["simulation"];
    rho = mat["rho"].number();
    nu = mat["nu"].number();
    mu = mat["mu"].number();
    // centre on the volume centroid (physics_base_matrix_collection.cpp:41-45)
    {
        double vol = 0, c[3] = {0, 0, 0};
        for (size_t t = 0; t < F.size(); t += 3) {
            const double *a = &V[3 * F[t]], *b = &V[3 * F[t + 1]], *d = &V[3 * F[t + 2]];
            const double tv = (a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                               a[2] * (b[0] * d[1] - b[1] * d[0])) / 6.0;
            vol += tv;
            for (int k = 0; k < 3; k++) c[k] += tv * (a[k] + b[k] + d[k]) / 4.0;
        }
        if (!(vol > 0)) return fail("mesh must be closed and oriented outward (non-positive volume)");
        for (int k = 0; k < 3; k++) cm[k] = c[k] / vol;
        for (int64_t i = 0; i < num_vertices; i++)
            for (int k = 0; k < 3; k++) V[3 * i + k] -= cm[k];
        mass = vol * rho;
    }
    if (dt_ != sim["dt"].number()) return fail("dt for simulator and elastodynamics BEM need to be the same.");
    const std::string ms = sim["multistep_method"].string();
    if (ms == "BDF1")
        multistep_method = BDF1;
    else if (ms == "BDF2")
        multistep_method = BDF2;
    else
        return fail("Unsupported multistep method detected. Valid options are \"BDF1\"/\"BDF2\"");
    use_galerkin = sim.contains("use_galerkin") && sim["use_galerkin"].boolean();
    update_local_frame = !sim.contains("update_local_frame") || sim["update_local_frame"].boolean();
    const int order = (int)sim["gaussian_quadrature_order"].number();
    const int sub = (int)sim["quadrature_subdivision"].number();
    // material constants and history length (matrix_collection.cpp:59-69)
    const double lambda = 2 * nu * mu / (1. - 2. * nu);
    c1 = std::sqrt((lambda + 2. * mu) / rho);
    c2 = std::sqrt(mu / rho);
    if (sim.contains("cq_length")) {
        max_time_history = (size_t)sim["cq_length"].number() - 1;
    } else {
        double maxd = 0.0;
        for (int64_t j = 0; j < num_vertices; j++)
            for (int64_t i = 0; i < num_vertices; i++) {
                const double dx = V[3 * i] - V[3 * j], dy = V[3 * i + 1] - V[3 * j + 1], dz = V[3 * i + 2] - V[3 * j + 2];
                maxd = std::max(maxd, std::sqrt(dx * dx + dy * dy + dz * dz));
            }
        max_time_history = (size_t)std::floor(maxd / (c2 * dt_) + 2);
    }
    R = sim.contains("cq_radius") ? sim["cq_radius"].number() : std::pow(1e-10, 1. / (2. * max_time_history));
    const size_t L = max_time_history + 1;

    std::string err;
    cqbem_context *ctx = context(err);
    if (!ctx) return fail(err);
    const int64_t n3 = 3 * num_vertices;
    std::vector<double> g((size_t)L * n3 * n3), h((size_t)L * n3 * n3), b((size_t)L * n3 * 6);
    cqbem_mesh mesh{V.data(), num_vertices, F.data(), (int64_t)(F.size() / 3),
                    vertex_map_inverse.empty() ? nullptr : vertex_map_inverse.data()};
    cqbem_params prm{c1, c2, rho, nu, dt_, R, (int32_t)L, multistep_method == BDF1 ? CQBEM_BDF1 : CQBEM_BDF2,
                     order, sub, use_galerkin ? 1 : 0, CQBEM_FP64, CQBEM_ALL};
    cqbem_outputs out{g.data(), h.data(), nullptr, nullptr, b.data(), 0};
    const auto t0 = std::chrono::steady_clock::now();
    if (cqbem_assemble(ctx, &mesh, &prm, &out) != CQBEM_OK) return fail(cqbem_last_error(ctx));
    last_assembly_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    G_matrices.assign(L, Dense(n3, n3));
    H_matrices.assign(L, Dense(n3, n3));
    B_matrices.assign(L, Dense(n3, 6));
    for (size_t n = 0; n < L; n++) {
        std::copy(g.begin() + n * n3 * n3, g.begin() + (n + 1) * n3 * n3, G_matrices[n].data.begin());
        for (int64_t e = 0; e < n3 * n3; e++) H_matrices[n].data[e] = -h[n * n3 * n3 + e];  // matrix_collection.cpp:140
        std::copy(b.begin() + n * n3 * 6, b.begin() + (n + 1) * n3 * 6, B_matrices[n].data.begin());
    }
    (void)cuda_thread_per_block_;  // launch shapes are chosen by the library
    const auto t1 = std::chrono::steady_clock::now();
    if (!post_process()) return false;
    last_factor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    return true;
}

// What init() derives from G_0, H_0 and B_0 for the time stepping: the external acceleration matrices
// (physics_base_matrix_collection.cpp:40-50,374-436), G' and the least-squares / LU solves
// (matrix_collection.cpp:174-215), the rotation-removal weights (:307-322) and u_p's diagonal blocks
// (:217-227).  The LU of H_0 and its solves run on the GPU.
bool ElastodynamicsCQBEMMatrixCollection::post_process() {
    const int64_t nv = num_vertices, n3 = 3 * nv, nf = (int64_t)(F.size() / 3);
    auto vtx = [&](int64_t v, int k) { return V[3 * v + k]; };
    std::vector<double> A2(nf), N(3 * nf);  // doubled areas, unit normals
    for (int64_t f = 0; f < nf; f++) {
        const int32_t *t = &F[3 * f];
        double e1[3], e2[3], c[3];
        for (int k = 0; k < 3; k++) e1[k] = vtx(t[1], k) - vtx(t[0], k), e2[k] = vtx(t[2], k) - vtx(t[0], k);
        c[0] = e1[1] * e2[2] - e1[2] * e2[1], c[1] = e1[2] * e2[0] - e1[0] * e2[2], c[2] = e1[0] * e2[1] - e1[1] * e2[0];
        A2[f] = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (int k = 0; k < 3; k++) N[3 * f + k] = c[k] / A2[f];
    }
    // inertia about the volume centroid of the centred mesh, order-4 triangle rule
    double cmc[3] = {0, 0, 0}, vol = 0;
    for (int64_t f = 0; f < nf; f++) {
        const double *a = &V[3 * F[3 * f]], *b = &V[3 * F[3 * f + 1]], *d = &V[3 * F[3 * f + 2]];
        const double tv = (a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                           a[2] * (b[0] * d[1] - b[1] * d[0])) / 6.0;
        vol += tv;
        for (int k = 0; k < 3; k++) cmc[k] += tv * (a[k] + b[k] + d[k]) / 4.0;
    }
    for (int k = 0; k < 3; k++) cmc[k] /= vol;
    std::fill(I, I + 9, 0.0);
    const auto rule = tri_rule(4);
    for (int64_t f = 0; f < nf; f++) {
        const double *n = &N[3 * f];
        for (const TriPoint &q : rule) {
            double r[3];
            for (int k = 0; k < 3; k++)
                r[k] = q.b[0] * vtx(F[3 * f], k) + q.b[1] * vtx(F[3 * f + 1], k) + q.b[2] * vtx(F[3 * f + 2], k) - cmc[k];
            const double w = q.w * 0.5 * A2[f];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
                    const double v = i == j ? (rho / 3.) * (r[i1] * r[i1] * r[i1] * n[i1] + r[i2] * r[i2] * r[i2] * n[i2])
                                            : (-rho / 4.) * (r[j] * r[i] * r[i] * n[i] + r[i] * r[j] * r[j] * n[j]);
                    I[3 * i + j] += w * v;
                }
        }
    }
    // external acceleration matrices (3 x 3 nv)
    translational_acc_mat = Dense(3, n3);
    rotational_acc_mat = Dense(3, n3);
    Dense rot_raw(3, n3);
    for (int64_t f = 0; f < nf; f++) {
        double sum[3];
        for (int k = 0; k < 3; k++) sum[k] = vtx(F[3 * f], k) + vtx(F[3 * f + 1], k) + vtx(F[3 * f + 2], k);
        for (int c = 0; c < 3; c++) {
            const int64_t j = F[3 * f + c];
            for (int k = 0; k < 3; k++) translational_acc_mat(k, 3 * j + k) += A2[f] / 6.;
            double v[3];
            for (int k = 0; k < 3; k++) v[k] = A2[f] / 6. * 0.25 * (sum[k] + vtx(j, k));
            // cross-product matrix [v]_x
            const double X[3][3] = {{0, -v[2], v[1]}, {v[2], 0, -v[0]}, {-v[1], v[0], 0}};
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) rot_raw(a, 3 * j + b) += X[a][b];
        }
    }
    for (double &x : translational_acc_mat.data) x *= 1. / mass;
    {
        // I^-1 by the adjugate
        const double *m = I;
        const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
        if (!(std::fabs(det) > 0)) return fail("singular inertia tensor");
        const double inv[9] = {(m[4] * m[8] - m[5] * m[7]) / det, (m[2] * m[7] - m[1] * m[8]) / det, (m[1] * m[5] - m[2] * m[4]) / det,
                               (m[5] * m[6] - m[3] * m[8]) / det, (m[0] * m[8] - m[2] * m[6]) / det, (m[2] * m[3] - m[0] * m[5]) / det,
                               (m[3] * m[7] - m[4] * m[6]) / det, (m[1] * m[6] - m[0] * m[7]) / det, (m[0] * m[4] - m[1] * m[3]) / det};
        for (int a = 0; a < 3; a++)
            for (int64_t c = 0; c < n3; c++)
                rotational_acc_mat(a, c) = inv[3 * a] * rot_raw(0, c) + inv[3 * a + 1] * rot_raw(1, c) + inv[3 * a + 2] * rot_raw(2, c);
    }
    // B_0 halves, H_0 (un-negated) and G'
    std::vector<double> Bt((size_t)(n3 * 3)), Be((size_t)(n3 * 3));
    for (int64_t r = 0; r < n3; r++)
        for (int k = 0; k < 3; k++) Bt[3 * r + k] = B_matrices[0](r, k), Be[3 * r + k] = B_matrices[0](r, 3 + k);
    std::vector<double> H0(H_matrices[0].data.size());
    for (size_t e = 0; e < H0.size(); e++) H0[e] = -H_matrices[0].data[e];
    u_p = G_matrices[0];  // becomes G', then H_0^-1 G'
    if (update_local_frame)
        for (int64_t r = 0; r < n3; r++)
            for (int64_t c = 0; c < n3; c++) {
                double v = 0;
                for (int k = 0; k < 3; k++) v += Bt[3 * r + k] * translational_acc_mat(k, c) + Be[3 * r + k] * rotational_acc_mat(k, c);
                u_p(r, c) += v;
            }
    std::string err;
    if (!B_trans_inv.compute(Bt, n3, 3, err) || !B_euler_inv.compute(Be, n3, 3, err)) return fail(err);
    auto as3 = [&](std::vector<double> x, int64_t cols) {
        Dense d(3, cols);
        d.data = std::move(x);
        return d;
    };
    B_trans_inv_H = as3(B_trans_inv.solve(H0.data(), n3), n3);
    B_euler_inv_H = as3(B_euler_inv.solve(H0.data(), n3), n3);
    B_trans_inv_G_prime = as3(B_trans_inv.solve(u_p.data.data(), n3), n3);
    B_euler_inv_G_prime = as3(B_euler_inv.solve(u_p.data.data(), n3), n3);
    B_euler_inv_B_trans = as3(B_euler_inv.solve(Bt.data(), 3), 3);
    // rotation-removal weights
    rotation_removal_weights.assign(nv, 0.0);
    for (int64_t f = 0; f < nf; f++) {
        double yn = 0;
        for (int k = 0; k < 3; k++) yn += vtx(F[3 * f], k) * N[3 * f + k];
        const double val = (1. / 6.) * A2[f] * yn;
        for (int c = 0; c < 3; c++) rotation_removal_weights[F[3 * f + c]] += val;
    }
    // LU of H_0 on the GPU and its solves
    if (!H_inv.factor(n3, H0.data(), err)) return fail(err);
    H0.clear(), H0.shrink_to_fit();
    H_inv_B_trans = Dense(n3, 3), H_inv_B_euler = Dense(n3, 3);
    H_inv_B_trans.data = Bt, H_inv_B_euler.data = Be;
    if (!H_inv.solve(u_p.data.data(), n3, err) || !H_inv.solve(H_inv_B_trans.data.data(), 3, err) ||
        !H_inv.solve(H_inv_B_euler.data.data(), 3, err))
        return fail(err);
    u_p_diag = Dense(nv, 9);
    for (int64_t v = 0; v < nv; v++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) u_p_diag(v, 3 * a + b) = u_p(3 * v + a, 3 * v + b);
    return true;
}

}  // namespace cqbem
