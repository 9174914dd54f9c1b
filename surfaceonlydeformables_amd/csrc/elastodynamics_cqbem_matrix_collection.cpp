// C++ facade over include/cqbem.h with the reference's collection interface
// (src/elastodynamics_cqbem_matrix_collection.cpp:13-278, src/physics_base_matrix_collection.cpp:32-57,
// src/elastics_bem_matrix_collection.cpp:16-41).
#include "elastodynamics_cqbem_matrix_collection.hpp"

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
        if (m.contains("enable_traction_discontinuity") && m["enable_traction_discontinuity"].boolean())
            return fail("traction-discontinuity vertex splitting is not part of this build");
        return init(config, v, f);
    } catch (const std::exception &e) {
        return fail(e.what());
    }
}

bool ElastodynamicsCQBEMMatrixCollection::init(const Json &config, const std::vector<double> &v,
                                               const std::vector<int32_t> &f) {
    try {
        V = v, F = f;
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
    cqbem_mesh mesh{V.data(), num_vertices, F.data(), (int64_t)(F.size() / 3), nullptr};
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
    return true;
}

}  // namespace cqbem
