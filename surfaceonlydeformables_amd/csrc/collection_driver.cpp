// Command-line driver of the C++ collection facade: reads a scene JSON (mesh file, material,
// simulation keys as in the reference's data/*.json), runs init(), and writes G, H (negated, as the
// object reads them) and B as raw little-endian doubles to <out>_{G,H,B}.bin, the matrices init()
// derives for the time stepping to <out>_post_<name>.bin, and a JSON line with L, nv and timings.
// Other modes (host only): --mesh scene.json (the prepared mesh), --cache-key scene.json, --dump doc.json.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "elastodynamics_cqbem_matrix_collection.hpp"

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
    if (argc != 3) {
        std::cerr << "usage: collection_driver scene.json out_prefix   (out_prefix '-': no files, timings only)" << std::endl;
        return 2;
    }
    std::ifstream in(argv[1]);
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
    cqbem::ElastodynamicsCQBEMMatrixCollection col(dt, enable_cuda, 32);
    if (!col.init(cfg)) {
        std::cerr << "init failed: " << col.last_error() << std::endl;
        return 1;
    }
    const std::string pre = argv[2];
    if (pre == "-") {
        std::printf("{\"L\": %zu, \"nv\": %lld, \"mass\": %.17g, \"ms\": %.3f, \"factor_ms\": %.3f, \"cached\": %s}\n",
                    col.max_time_history + 1, (long long)col.num_vertices, col.mass, col.last_assembly_ms, col.last_factor_ms,
                    col.loaded_from_cache ? "true" : "false");
        return 0;
    }
    auto dump = [&](const std::string &name, const std::vector<cqbem::Dense> &mats) {
        std::ofstream o(pre + "_" + name + ".bin", std::ios::binary);
        for (auto &m : mats) o.write(reinterpret_cast<const char *>(m.data.data()), m.data.size() * sizeof(double));
    };
    dump("G", col.G_matrices), dump("H", col.H_matrices), dump("B", col.B_matrices);
    auto dump1 = [&](const std::string &name, const std::vector<double> &v) {
        std::ofstream o(pre + "_post_" + name + ".bin", std::ios::binary);
        o.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(double));
    };
    dump1("u_p", col.u_p.data), dump1("u_p_diag", col.u_p_diag.data);
    dump1("H_inv_B_trans", col.H_inv_B_trans.data), dump1("H_inv_B_euler", col.H_inv_B_euler.data);
    dump1("B_trans_inv_H", col.B_trans_inv_H.data), dump1("B_euler_inv_H", col.B_euler_inv_H.data);
    dump1("B_trans_inv_G_prime", col.B_trans_inv_G_prime.data), dump1("B_euler_inv_G_prime", col.B_euler_inv_G_prime.data);
    dump1("B_euler_inv_B_trans", col.B_euler_inv_B_trans.data);
    dump1("translational_acc_mat", col.translational_acc_mat.data), dump1("rotational_acc_mat", col.rotational_acc_mat.data);
    dump1("rotation_removal_weights", col.rotation_removal_weights);
    dump1("I", std::vector<double>(col.I, col.I + 9));
    std::printf("{\"L\": %zu, \"nv\": %lld, \"R\": %.17g, \"c1\": %.17g, \"c2\": %.17g, \"mass\": %.17g, "
                "\"ms\": %.3f, \"factor_ms\": %.3f, \"cached\": %s}\n",
                col.max_time_history + 1, (long long)col.num_vertices, col.R, col.c1, col.c2, col.mass, col.last_assembly_ms,
                col.last_factor_ms, col.loaded_from_cache ? "true" : "false");
    return 0;
}
