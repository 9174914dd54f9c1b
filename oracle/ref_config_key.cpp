// TEST INFRASTRUCTURE ONLY: builds against the reference's own vendored JSON library
// (include/json.hpp, nlohmann::json) to print, for a scene file, the key under which the
// reference caches an elastodynamicsCQBEM object group's matrices
// (src/elastodynamics_cqbem_matrix_collection.cpp:41-57): the group's JSON with
// simulation.compression_ratio and simulation.update_local_frame removed, its compact dump, and
// the std::hash of that dump in upper-case hex.  Output per group: "<hex>\t<dump>".
// With --dump FILE it prints the compact dump of a whole JSON document (number formatting checks).
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "json.hpp"

int main(int argc, char **argv) {
    using nlohmann::json;
    if (argc == 3 && std::string(argv[1]) == "--dump") {
        std::ifstream in(argv[2]);
        json doc;
        in >> doc;
        std::cout << doc.dump() << "\n";
        return 0;
    }
    if (argc != 2) {
        std::cerr << "usage: config_key scene.json | config_key --dump doc.json\n";
        return 2;
    }
    std::ifstream in(argv[1]);
    json scene;
    in >> scene;
    for (auto &group : scene["physics_object_groups"]) {
        const json &cfg = group["physics_object_group"];
        if (!cfg.is_object() || cfg["type"] != "elastodynamicsCQBEM") continue;
        json copy = cfg;
        for (const char *k : {"compression_ratio", "update_local_frame"})
            if (copy["simulation"].contains(k)) copy["simulation"].erase(copy["simulation"].find(k));
        std::ostringstream ss;
        ss << std::uppercase << std::hex << std::hash<std::string>()(copy.dump());
        std::cout << ss.str() << "\t" << copy.dump() << "\n";
    }
    return 0;
}
