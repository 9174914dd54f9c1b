// Command-line driver of the C++ collection facade: reads a scene JSON (mesh file, material,
// simulation keys as in the reference's data/*.json), runs init(), and writes G, H (negated, as the
// object reads them) and B as raw little-endian doubles to <out>_{G,H,B}.bin, plus L and nv.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "elastodynamics_cqbem_matrix_collection.hpp"

int main(int argc, char **argv) {
    if (argc != 3) {
        std::cerr << "usage: collection_driver scene.json out_prefix" << std::endl;
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
    auto dump = [&](const std::string &name, const std::vector<cqbem::Dense> &mats) {
        std::ofstream o(pre + "_" + name + ".bin", std::ios::binary);
        for (auto &m : mats) o.write(reinterpret_cast<const char *>(m.data.data()), m.data.size() * sizeof(double));
    };
    dump("G", col.G_matrices), dump("H", col.H_matrices), dump("B", col.B_matrices);
    std::printf("{\"L\": %zu, \"nv\": %lld, \"R\": %.17g, \"c1\": %.17g, \"c2\": %.17g, \"ms\": %.3f}\n",
                col.max_time_history + 1, (long long)col.num_vertices, col.R, col.c1, col.c2, col.last_assembly_ms);
    return 0;
}
