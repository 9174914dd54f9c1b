// Traction-discontinuity vertex split of a closed triangle mesh, the mesh step of
// PhysicsBaseMatrixCollection::load_mesh (src/physics_base_matrix_collection.cpp:203-365).
//
// Edges whose two face normals differ by more than a threshold angle are sharp.  Every closed
// loop of sharp edges is walked once; each time the walk leaves a vertex along a new sharp edge
// the vertex gets a new copy, and the faces of the sector between the incoming and the outgoing
// sharp edge (counter-clockwise around the vertex) are relabelled to that copy.  The copies are
// then placed right after their original vertex, and vertex_map_inverse maps every output vertex
// to its input vertex.  The walk order (sharp edges taken in ascending (from, to) order, the next
// sharp edge searched counter-clockwise around the current vertex) decides which sector keeps the
// original index, so the output numbering is the reference's.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cqbem.h"

namespace cqbem {

namespace {

struct SplitError {
    std::string msg;
};

// Neighbours of every vertex in counter-clockwise order (seen from outside) around it: for each
// face (s, d, v) the pair (d, v) is a step of s's one-ring, chained d -> v -> ...  (igl::adjacency_list
// with sorted = true; the reference only walks the ring cyclically, so its starting neighbour is
// immaterial).
std::vector<std::vector<int>> ccw_rings(const int32_t *F, int nf, int nv) {
    std::vector<std::vector<std::pair<int, int>>> steps(nv);
    for (int f = 0; f < nf; f++)
        for (int j = 0; j < 3; j++) steps[F[3 * f + j]].push_back({F[3 * f + (j + 1) % 3], F[3 * f + (j + 2) % 3]});
    std::vector<std::vector<int>> ring(nv);
    for (int s = 0; s < nv; s++) {
        const auto &st = steps[s];
        if (st.empty()) continue;
        std::map<int, int> next;  // d -> v
        for (const auto &e : st) {
            if (next.count(e.first)) throw SplitError{"mesh is not manifold at vertex " + std::to_string(s)};
            next[e.first] = e.second;
        }
        int c = st[0].first;
        for (size_t k = 0; k < st.size(); k++) {
            ring[s].push_back(c);
            auto it = next.find(c);
            if (it == next.end()) throw SplitError{"mesh is not closed at vertex " + std::to_string(s)};
            c = it->second;
        }
        if (c != st[0].first) throw SplitError{"one-ring of vertex " + std::to_string(s) + " is not a single loop"};
    }
    return ring;
}

int position(const std::vector<int> &ring, int v) {
    for (size_t k = 0; k < ring.size(); k++)
        if (ring[k] == v) return (int)k;
    throw SplitError{"inconsistent one-ring"};
}

}  // namespace

int split_traction_discontinuity(const double *V, int nv, const int32_t *F_in, int nf, double threshold_deg,
                                 std::vector<double> &V_out, std::vector<int32_t> &F_out, std::vector<int32_t> &vmap_out,
                                 std::string &err) {
    try {
        std::vector<int32_t> F(F_in, F_in + 3 * (size_t)nf);
        // half-edge e = c nf + f runs F[f][c+1] -> F[f][c+2] (the oriented facets of face f)
        auto he_from = [&](int e) { return F_in[3 * (e % nf) + (e / nf + 1) % 3]; };
        auto he_to = [&](int e) { return F_in[3 * (e % nf) + (e / nf + 2) % 3]; };
        // unit face normals
        std::vector<double> N(3 * (size_t)nf);
        for (int f = 0; f < nf; f++) {
            const double *a = V + 3 * F_in[3 * f], *b = V + 3 * F_in[3 * f + 1], *c = V + 3 * F_in[3 * f + 2];
            const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, w[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
            const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int k = 0; k < 3; k++) N[3 * f + k] = len > 0 ? n[k] / len : 0.0;
        }
        // unique edges and their half-edges; sharp when the normals differ by more than the threshold
        // (igl::sharp_edges, as called at physics_base_matrix_collection.cpp:218)
        std::map<std::pair<int, int>, std::vector<int>> uedges;
        for (int e = 0; e < 3 * nf; e++) {
            const int a = he_from(e), b = he_to(e);
            uedges[{std::min(a, b), std::max(a, b)}].push_back(e);
        }
        const double angle = threshold_deg * M_PI / 180.0;
        std::map<std::pair<int, int>, int> sharp;  // directed sharp edges (to, from) of each half-edge -> half-edge
        for (const auto &kv : uedges) {
            const auto &h = kv.second;
            bool is_sharp = false;
            for (size_t i = 0; i < h.size() && !is_sharp; i++)
                for (size_t j = i + 1; j < h.size() && !is_sharp; j++) {
                    const double *n1 = &N[3 * (h[i] % nf)], *n2 = &N[3 * (h[j] % nf)];
                    is_sharp = std::abs(std::acos(n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2])) > angle;
                }
            if (!is_sharp) continue;
            if (h.size() != 2 || he_to(h[0]) != he_from(h[1]) || he_from(h[0]) != he_to(h[1]))
                throw SplitError{"sharp edge is not shared by two consistently oriented faces"};
            for (int e : h) sharp[{he_to(e), he_from(e)}] = e;
        }
        std::map<std::pair<int, int>, int> directed;  // (to, from) -> half-edge, every half-edge
        for (int e = 0; e < 3 * nf; e++) directed[{he_to(e), he_from(e)}] = e;
        const std::vector<std::vector<int>> ring = ccw_rings(F_in, nf, nv);

        std::map<int, std::vector<int>> copies;  // input vertex -> its output indices (itself first)
        int next_index = nv;
        auto add_copy = [&](int v) {
            auto it = copies.find(v);
            if (it == copies.end()) copies[v].push_back(v);
            else it->second.push_back(next_index++);
        };
        auto last_copy = [&](int v) {
            auto it = copies.find(v);
            if (it == copies.end() || it->second.empty()) throw SplitError{"open sharp-edge chain at vertex " + std::to_string(v)};
            return it->second.back();
        };
        auto relabel = [&](int e, int slot_offset, int value) { F[3 * (e % nf) + (e / nf + slot_offset) % 3] = value; };

        while (!sharp.empty()) {
            auto cur = sharp.begin();
            add_copy(cur->first.first);
            const int start = cur->first.first, start2 = cur->first.second;
            while (cur != sharp.end()) {  // one loop of sharp edges
                const std::pair<int, int> key = cur->first;
                const int e = cur->second;
                sharp.erase(cur);
                const int a = key.first, b = key.second;
                const std::vector<int> &rb = ring[b];
                const int n = (int)rb.size();
                const int first_pos = position(rb, a);
                int pos = first_pos;
                do {  // next sharp edge leaving b, counter-clockwise from a (a itself last)
                    pos = (pos + 1) % n;
                    cur = sharp.find({b, rb[pos]});
                    if (cur != sharp.end()) break;
                } while (pos != first_pos);
                const bool back_at_start = b == start;
                if (back_at_start) cur = sharp.end();
                if (cur != sharp.end()) add_copy(b);
                relabel(e, 2, last_copy(a));
                relabel(e, 1, last_copy(b));
                if (cur != sharp.end())  // the sector between the two sharp edges takes b's new copy
                    for (int p = first_pos; p != pos;) {
                        p = (p + 1) % n;
                        relabel(directed.at({b, rb[p]}), 2, last_copy(b));
                    }
                if (back_at_start) {  // close the loop: the sector up to the loop's first edge
                    const int stop = position(rb, start2);
                    for (int p = first_pos; p != stop;) {
                        p = (p + 1) % n;
                        relabel(directed.at({b, rb[p]}), 2, last_copy(b));
                    }
                }
            }
        }

        // copies follow their original; vertex_map_inverse maps output -> input vertex
        for (int v = 0; v < nv; v++)
            if (!copies.count(v)) copies[v].push_back(v);
        std::vector<int> order;  // output position -> pre-order index
        std::vector<int> pre_map(next_index);
        for (int v = 0; v < nv; v++)
            for (int w : copies[v]) order.push_back(w), pre_map[w] = v;
        std::vector<int> newpos(next_index);
        for (int i = 0; i < next_index; i++) newpos[order[i]] = i;
        V_out.assign(3 * (size_t)next_index, 0.0);
        vmap_out.assign(next_index, 0);
        for (int i = 0; i < next_index; i++) {
            const int src = pre_map[order[i]];
            for (int k = 0; k < 3; k++) V_out[3 * i + k] = V[3 * src + k];
            vmap_out[i] = src;
        }
        F_out.resize(3 * (size_t)nf);
        for (size_t k = 0; k < F.size(); k++) F_out[k] = newpos[F[k]];
        return (int)CQBEM_OK;
    } catch (const SplitError &e) {
        err = e.msg;
        return (int)CQBEM_ERR_ARG;
    } catch (const std::out_of_range &) {
        err = "inconsistent half-edge structure";
        return (int)CQBEM_ERR_ARG;
    }
}

}  // namespace cqbem

extern "C" int cqbem_split_traction_discontinuity(const double *V, int64_t nv, const int32_t *F, int64_t nf,
                                                  double threshold_deg, double *V_out, int32_t *F_out,
                                                  int32_t *vertex_map_inverse, int64_t *nv_out) {
    if (!V || !F || !nv_out || nv <= 0 || nf <= 0 || nv > INT32_MAX / 4 || nf > INT32_MAX / 4) return CQBEM_ERR_ARG;
    for (int64_t k = 0; k < 3 * nf; k++)
        if (F[k] < 0 || F[k] >= nv) return CQBEM_ERR_ARG;
    std::vector<double> Vs;
    std::vector<int32_t> Fs, vm;
    std::string err;
    const int rc = cqbem::split_traction_discontinuity(V, (int)nv, F, (int)nf, threshold_deg, Vs, Fs, vm, err);
    if (rc != CQBEM_OK) return rc;
    const int64_t n = (int64_t)vm.size();
    if (V_out || F_out || vertex_map_inverse) {
        if (*nv_out < n) return CQBEM_ERR_ARG;
        if (V_out) std::copy(Vs.begin(), Vs.end(), V_out);
        if (F_out) std::copy(Fs.begin(), Fs.end(), F_out);
        if (vertex_map_inverse) std::copy(vm.begin(), vm.end(), vertex_map_inverse);
    }
    *nv_out = n;
    return CQBEM_OK;
}
