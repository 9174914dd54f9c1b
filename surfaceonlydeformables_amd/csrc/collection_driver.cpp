// Command-line driver of the C++ collection facade: reads a scene JSON (mesh file, material,
// simulation keys as in the reference's data/*.json), runs init(), and writes G, H (negated, as the
// object reads them) and B as raw little-endian doubles to <out>_{G,H,B}.bin, the matrices init()
// derives for the time stepping to <out>_post_<name>.bin, and a JSON line with L, nv and timings.
// --simulate scene.json steps tractions.bin out_prefix: additionally steps one ElastodynamicsCQBEMObject
// (scene key "object": its init keys; "gravity": 3 values, default (0, -9.8, 0)) for `steps` steps; row k of
// tractions.bin (steps x 3 nv doubles) is the step's p, applied as a constraint result when not all zero.
// Per step it appends translation, rotation, translational_vel, rotational_vel, the newest acceleration
// (6), V_local and V_global to <out>_traj.bin; out_prefix '-' writes no files.  step_ms is the mean step
// time once the history is full (steps >= L), after the first step otherwise.
// Other modes (host only): --mesh scene.json (the prepared mesh), --cache-key scene.json, --dump doc.json.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "elastodynamics_cqbem_matrix_collection.hpp"
#include "elastodynamics_cqbem_object.hpp"

int main(int argc, char **argv) {
    if (argc == 3 && (std::string(argv[1]) == "--cache-key" || std::string(argv[1]) == "--dump")) {
        // --cache-key scene.json: per elastodynamicsCQBEM object group of a full scene file, the key of the
        // reference's matrix cache and the JSON it hashes; --dump doc.json: the compact serialisation
        std::ifstream in(argv[2]);
        std::stringstream ss;
        ss << in.rdbuf();
        try {
            const cqbem::Json doc = cqbem::Json::parse(ss.str());
            if (std::string(argv[1]) == "--dump") {
                std::cout << doc.dump() << "\n";
                return 0;
            }
            for (const auto &g : doc["physics_object_groups"].arr) {
                const cqbem::Json &cfg = g["physics_object_group"];
                if (cfg.kind != cqbem::Json::Object || !cfg.contains("type") || cfg["type"].string() != "elastodynamicsCQBEM")
                    continue;
                cqbem::Json stripped;
                const std::string key = cqbem::cache_key(cfg, &stripped);
                std::cout << key << "\t" << stripped.dump() << "\n";
            }
        } catch (const std::exception &e) {
            std::cerr << e.what() << std::endl;
            return 2;
        }
        return 0;
    }
    if (argc == 3 && std::string(argv[1]) == "--mesh") {
        // --mesh scene.json: the prepared mesh (V, F, vmap) of the scene's "mesh" block, as one JSON line
        std::ifstream in(argv[2]);
        std::stringstream ss;
        ss << in.rdbuf();
        try {
            const cqbem::Json doc = cqbem::Json::parse(ss.str());
            std::vector<double> V;
            std::vector<int32_t> F, vm;
            std::string err;
            if (!cqbem::prepare_mesh(doc["mesh"], V, F, vm, err)) {
                std::cerr << err << std::endl;
                return 1;
            }
            auto put = [](const char *name, const auto &xs, bool last) {
                std::printf("\"%s\": [", name);
                for (size_t k = 0; k < xs.size(); k++) std::printf(k ? ", %.17g" : "%.17g", (double)xs[k]);
                std::printf(last ? "]" : "], ");
            };
            std::printf("{");
            put("V", V, false), put("F", F, false), put("vmap", vm, true);
            std::printf("}\n");
        } catch (const std::exception &e) {
            std::cerr << e.what() << std::endl;
            return 2;
        }
        return 0;
    }
    const bool simulate = argc == 6 && std::string(argv[1]) == "--simulate";
    const std::string scene_path = simulate ? argv[2] : (argc == 3 ? argv[1] : "");
    const std::string pre = simulate ? argv[5] : (argc == 3 ? argv[2] : "");
    if (!simulate && argc != 3) {
        std::cerr << "usage: collection_driver scene.json out_prefix   (out_prefix '-': no files, timings only)" << std::endl;
        return 2;
    }
    std::ifstream in(scene_path);
    std::stringstream ss;
    ss << in.rdbuf();
    cqbem::Json cfg;
    try {
        cfg = cqbem::Json::parse(ss.str());
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 2;
    }
    const double dt = cfg["simulation"]["dt"].number();
    const bool enable_cuda = !cfg.contains("enable_cuda") || cfg["enable_cuda"].boolean();
    auto colp = std::make_shared<cqbem::ElastodynamicsCQBEMMatrixCollection>(dt, enable_cuda, 32);
    cqbem::ElastodynamicsCQBEMMatrixCollection &col = *colp;
    if (!col.init(cfg)) {
        std::cerr << "init failed: " << col.last_error() << std::endl;
        return 1;
    }
    if (pre == "-" && !simulate) {
        std::printf("{\"L\": %zu, \"nv\": %lld, \"mass\": %.17g, \"ms\": %.3f, \"factor_ms\": %.3f, \"cached\": %s}\n",
                    col.max_time_history + 1, (long long)col.num_vertices, col.mass, col.last_assembly_ms, col.last_factor_ms,
                    col.loaded_from_cache ? "true" : "false");
        return 0;
    }
    const bool files = pre != "-";  // '-' with --simulate: timings only
    auto dump = [&](const std::string &name, const std::vector<cqbem::Dense> &mats) {
        if (!files) return;
        std::ofstream o(pre + "_" + name + ".bin", std::ios::binary);
        for (auto &m : mats) o.write(reinterpret_cast<const char *>(m.data.data()), m.data.size() * sizeof(double));
    };
    dump("G", col.G_matrices), dump("H", col.H_matrices), dump("B", col.B_matrices);
    auto dump1 = [&](const std::string &name, const std::vector<double> &v) {
        if (!files) return;
        std::ofstream o(pre + "_post_" + name + ".bin", std::ios::binary);
        o.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(double));
    };
    dump1("V", col.V);  // the assembled vertices: after the mesh steps, centred
    dump1("u_p", col.u_p.data), dump1("u_p_diag", col.u_p_diag.data);
    dump1("H_inv_B_trans", col.H_inv_B_trans.data), dump1("H_inv_B_euler", col.H_inv_B_euler.data);
    dump1("B_trans_inv_H", col.B_trans_inv_H.data), dump1("B_euler_inv_H", col.B_euler_inv_H.data);
    dump1("B_trans_inv_G_prime", col.B_trans_inv_G_prime.data), dump1("B_euler_inv_G_prime", col.B_euler_inv_G_prime.data);
    dump1("B_euler_inv_B_trans", col.B_euler_inv_B_trans.data);
    dump1("translational_acc_mat", col.translational_acc_mat.data), dump1("rotational_acc_mat", col.rotational_acc_mat.data);
    dump1("rotation_removal_weights", col.rotation_removal_weights);
    dump1("I", std::vector<double>(col.I, col.I + 9));
    dump1("V", col.V);
    dump1("original_cm", std::vector<double>(col.original_cm, col.original_cm + 3));
    double step_ms = 0.0, first_step_ms = 0.0, upload_ms = 0.0, phase_step_ms[5] = {0, 0, 0, 0, 0};
    if (simulate) {
        const long steps = std::stol(argv[3]);
        const int64_t n3 = 3 * col.num_vertices;
        std::vector<double> tr((size_t)(steps * n3));
        std::ifstream tin(argv[4], std::ios::binary);
        if (!tin.read(reinterpret_cast<char *>(tr.data()), (std::streamsize)(tr.size() * sizeof(double)))) {
            std::cerr << "cannot read " << steps << " x " << n3 << " tractions from " << argv[4] << std::endl;
            return 2;
        }
        double g[3] = {0.0, -9.8, 0.0};
        if (cfg.contains("gravity"))
            for (int k = 0; k < 3; k++) g[k] = cfg["gravity"].arr.at(k).number();
        cqbem::ElastodynamicsCQBEMObject obj(colp, dt, g);
        if (!obj.init(cfg.contains("object") ? cfg["object"] : cqbem::Json())) {
            std::cerr << "object init failed: " << obj.last_error() << std::endl;
            return 1;
        }
        std::ofstream o;
        if (files) o.open(pre + "_traj.bin", std::ios::binary);
        auto put = [&](const double *x, size_t n) {
            if (files) o.write(reinterpret_cast<const char *>(x), n * sizeof(double));
        };
        auto now = [] { return std::chrono::steady_clock::now(); };
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        auto tu = now();
        try {
            (void)col.resident();  // the one-time upload of G, H, B, u_p to the device, outside the step times
        } catch (const std::exception &e) {
            std::cerr << e.what() << std::endl;
            return 1;
        }
        upload_ms = ms(tu, now());
        // step_ms: mean over the steps once the history is full (k >= max_time_history), or after the first
        const long k0 = std::max(1L, std::min((long)col.max_time_history, steps - 1));
        auto t0 = now();
        for (long k = 0; k < steps; k++) {
            if (k == 1) first_step_ms = ms(t0, now());
            if (k == k0) {
                t0 = now();
                std::fill(obj.phase_ms, obj.phase_ms + 5, 0.0);
            }
            if (!obj.estimate_next_state()) {
                std::cerr << "step " << k << ": " << obj.last_error() << std::endl;
                return 1;
            }
            const double *row = tr.data() + k * n3;
            if (std::any_of(row, row + n3, [](double x) { return x != 0.0; })) {
                obj.update_V_estimate_matrices();
                obj.p.assign(row, row + n3);
                obj.has_constraints = true;
            }
            if (!obj.confirm_next_state()) {
                std::cerr << "step " << k << ": " << obj.last_error() << std::endl;
                return 1;
            }
            put(obj.translation, 3), put(obj.rotation, 9), put(obj.translational_vel, 3), put(obj.rotational_vel, 3);
            put(obj.hist_acc_vectors.empty() ? std::vector<double>(6, 0.0).data() : obj.hist_acc_vectors.back().data(), 6);
            put(obj.V_local.data(), obj.V_local.size()), put(obj.V_global.data(), obj.V_global.size());
        }
        if (steps == 1) first_step_ms = ms(t0, now());
        else step_ms = ms(t0, now()) / (double)(steps - k0);
        for (int k = 0; k < 5; k++) phase_step_ms[k] = obj.phase_ms[k] / (double)std::max(1L, steps - k0);
    }
    std::printf("{\"L\": %zu, \"nv\": %lld, \"R\": %.17g, \"c1\": %.17g, \"c2\": %.17g, \"mass\": %.17g, "
                "\"ms\": %.3f, \"factor_ms\": %.3f, \"upload_ms\": %.3f, \"first_step_ms\": %.3f, \"step_ms\": %.3f, "
                "\"step_phases_ms\": {\"history\": %.3f, \"solve\": %.3f, \"u_p_p\": %.3f, \"history_upload\": %.3f, "
                "\"host\": %.3f}, \"factor_phases_ms\": {\"geometry\": %.3f, \"g_prime\": %.3f, \"least_squares\": %.3f, "
                "\"lu\": %.3f, \"lu_solves\": %.3f}, \"cached\": %s, \"compressed_loaded\": %d}\n",
                col.max_time_history + 1, (long long)col.num_vertices, col.R, col.c1, col.c2, col.mass, col.last_assembly_ms,
                col.last_factor_ms, upload_ms, first_step_ms, step_ms, phase_step_ms[0], phase_step_ms[1], phase_step_ms[2],
                phase_step_ms[3], phase_step_ms[4], col.post_process_phase_ms[0], col.post_process_phase_ms[1],
                col.post_process_phase_ms[2], col.post_process_phase_ms[3], col.post_process_phase_ms[4],
                col.loaded_from_cache ? "true" : "false", col.compressed_loaded);
    return 0;
}
