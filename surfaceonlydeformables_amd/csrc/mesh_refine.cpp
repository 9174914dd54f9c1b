// Mesh subdivision and decimation: the geometry steps of PhysicsBaseMatrixCollection::load_mesh between the
// scaling and the normal flip (src/physics_base_matrix_collection.cpp:109-151).  The reference takes them from
// libigl v2.4.0 (cmake/libigl.cmake; the library is not vendored): igl::upsample, igl::loop,
// igl::false_barycentric_subdivision and igl::decimate.  They are restated here from their published algorithms.
//
// Upsample and Loop insert one vertex per edge, numbered after the old vertices in the order in which their edge
// is first met when the faces are scanned in order, a face's edges taken as (v0,v1), (v1,v2), (v2,v0).  Face
// (v0,v1,v2) with edge vertices m01, m12, m20 becomes (v0,m01,m20), (v1,m12,m01), (m01,m12,m20), (m12,v2,m20).
// The new positions are the product S V of a sparse stencil matrix S; each row sums its columns in ascending
// order, one product and one addition at a time (the order of a column-major sparse-times-dense product).
//   Upsample: edge vertex (a + b) / 2.
//   Loop (Warren's weights): edge vertex 3/8 (a + b) + 1/8 (c + d), c and d the vertices opposite the edge,
//   (a + b) / 2 on a boundary edge; an interior vertex of valence n keeps 1 - n beta of itself and takes beta of
//   each neighbour, beta = 3/16 for n = 3 and 3/(8n) otherwise; a boundary vertex keeps 3/4 and takes 1/8 of each
//   of its two boundary neighbours.
// FalseBarycentric appends the barycentre ((a + b) + c) / 3 of face f as vertex nv + f and splits the face into
// (v0,v1,c), (v1,v2,c), (v2,v0,c).
//
// Decimation collapses edges greedily, cheapest first, until at most max_faces faces remain (the first collapse
// is made before the face count is tested).  Cost of an edge: its length; placement: its midpoint.  Edges are
// numbered in ascending (min, max) vertex order; the queue orders (cost, edge number), and an edge whose
// neighbourhood changes is re-queued with its new cost.  Collapsing (s, d), s < d, moves s to the placement,
// relabels d to s, removes the two faces of the edge and, of each removed face's two other edges, keeps the one
// at s.  A collapse is refused (and the edge re-queued at infinite cost) unless the closed one-rings of s and d
// share exactly four vertices (s, d and the two opposite vertices: the link condition), or when s and d are two
// vertices of a lone tetrahedron.  The surviving faces keep their order, and unreferenced vertices are dropped
// with the survivors in their order.  Closed, consistently oriented, edge-manifold meshes only: the BEM needs
// closed surfaces, and the reference's pipeline refuses a mesh without positive volume after this step.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <iterator>
#include <limits>
#include <queue>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/cqbem.h"

// each product rounded before it is added, as the reference's host code computes it
#pragma STDC FP_CONTRACT OFF

namespace cqbem {

namespace {

using Tri = std::array<int32_t, 3>;

struct EdgeKey {
    int32_t a, b;  // a < b
    bool operator<(const EdgeKey &o) const { return a != o.a ? a < o.a : b < o.b; }
    bool operator==(const EdgeKey &o) const { return a == o.a && b == o.b; }
};

EdgeKey key(int32_t x, int32_t y) { return x < y ? EdgeKey{x, y} : EdgeKey{y, x}; }

// The faces' edges, as the edge vertices of upsample and loop number them: ni[3 f + j] is the number (from 0) of
// the undirected edge (F[f][j], F[f][j+1]); partner[3 f + j] is the slot 3 g + j' of the same edge in the face g
// on its other side, or -1 at a boundary.
struct EdgeNumbering {
    std::vector<int32_t> ni;
    std::vector<int64_t> partner;
    int32_t count = 0;
};

bool number_edges(const std::vector<Tri> &F, EdgeNumbering &out, std::string &err) {
    const size_t nf = F.size();
    std::vector<std::pair<EdgeKey, int64_t>> all;  // (edge, 3 f + j)
    all.reserve(3 * nf);
    for (size_t f = 0; f < nf; f++)
        for (int j = 0; j < 3; j++) all.push_back({key(F[f][j], F[f][(j + 1) % 3]), (int64_t)(3 * f + j)});
    std::sort(all.begin(), all.end(), [](const auto &x, const auto &y) {
        return x.first < y.first || (x.first == y.first && x.second < y.second);
    });
    out.ni.assign(3 * nf, -1);
    out.partner.assign(3 * nf, -1);
    for (size_t i = 0; i < all.size();) {
        size_t k = i + 1;
        while (k < all.size() && all[k].first == all[i].first) k++;
        if (k - i > 2) {
            err = "mesh is not edge-manifold: edge (" + std::to_string(all[i].first.a) + ", " +
                  std::to_string(all[i].first.b) + ") has " + std::to_string(k - i) + " faces";
            return false;
        }
        if (k - i == 2) out.partner[all[i].second] = all[i + 1].second, out.partner[all[i + 1].second] = all[i].second;
        i = k;
    }
    // numbered in order of first appearance (face order, then j)
    out.count = 0;
    for (size_t s = 0; s < 3 * nf; s++) {
        if (out.ni[s] >= 0) continue;
        out.ni[s] = out.count;
        if (out.partner[s] >= 0) out.ni[out.partner[s]] = out.count;
        out.count++;
    }
    return true;
}

// One row of S V: the stencil's (column, weight) pairs, duplicates merged, summed in ascending column order.
void stencil_row(std::vector<std::pair<int32_t, double>> &st, const std::vector<double> &V, double *out) {
    std::sort(st.begin(), st.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    size_t w = 0;
    for (size_t i = 0; i < st.size(); i++) {
        if (w > 0 && st[w - 1].first == st[i].first)
            st[w - 1].second += st[i].second;
        else
            st[w++] = st[i];
    }
    st.resize(w);
    for (int c = 0; c < 3; c++) {
        double acc = 0.0;
        for (const auto &e : st) {
            const double t = e.second * V[3 * (size_t)e.first + c];
            acc = acc + t;
        }
        out[c] = acc;
    }
}

bool split_four(std::vector<double> &V, std::vector<Tri> &F, bool loop, std::string &err) {
    EdgeNumbering en;
    if (!number_edges(F, en, err)) return false;
    const size_t nv = V.size() / 3, nf = F.size();
    std::vector<double> W(3 * (nv + (size_t)en.count));
    std::vector<std::pair<int32_t, double>> st;
    if (!loop) {
        std::copy(V.begin(), V.end(), W.begin());
    } else {
        // neighbours and boundary neighbours of every old vertex
        std::vector<std::vector<int32_t>> nb(nv), bnb(nv);
        for (size_t s = 0; s < 3 * nf; s++) {
            if (en.partner[s] >= 0 && en.partner[s] < (int64_t)s) continue;  // each undirected edge once
            const int32_t a = F[s / 3][s % 3], b = F[s / 3][(s % 3 + 1) % 3];
            nb[a].push_back(b), nb[b].push_back(a);
            if (en.partner[s] < 0) bnb[a].push_back(b), bnb[b].push_back(a);
        }
        for (size_t i = 0; i < nv; i++) {
            st.clear();
            if (!bnb[i].empty()) {
                if (bnb[i].size() != 2) {
                    err = "mesh boundary is not manifold at vertex " + std::to_string(i);
                    return false;
                }
                st.push_back({bnb[i][0], 1. / 8.});
                st.push_back({bnb[i][1], 1. / 8.});
                st.push_back({(int32_t)i, 3. / 4.});
            } else if (!nb[i].empty()) {
                const double dn = (double)nb[i].size();
                const double beta = nb[i].size() == 3 ? 3. / 16. : 3. / 8. / dn;
                for (int32_t x : nb[i]) st.push_back({x, beta});
                st.push_back({(int32_t)i, 1. - dn * beta});
            } else {
                st.push_back({(int32_t)i, 1.0});  // a vertex no face uses: kept in place
            }
            stencil_row(st, V, &W[3 * i]);
        }
    }
    std::vector<char> done(en.count, 0);
    for (size_t f = 0; f < nf; f++)
        for (int j = 0; j < 3; j++) {
            const size_t s = 3 * f + j;
            const int32_t m = en.ni[s];
            if (done[m]) continue;
            done[m] = 1;
            const int32_t a = F[f][j], b = F[f][(j + 1) % 3];
            st.clear();
            if (!loop || en.partner[s] < 0) {
                st.push_back({a, 1. / 2.});
                st.push_back({b, 1. / 2.});
            } else {
                st.push_back({a, 3. / 8.});
                st.push_back({b, 3. / 8.});
                st.push_back({F[f][(j + 2) % 3], 1. / 8.});
                const int64_t q = en.partner[s];  // the vertex of the other face opposite the edge
                st.push_back({F[q / 3][(q % 3 + 2) % 3], 1. / 8.});
            }
            stencil_row(st, V, &W[3 * (nv + (size_t)m)]);
        }
    std::vector<Tri> G(4 * nf);
    for (size_t f = 0; f < nf; f++) {
        const int32_t v0 = F[f][0], v1 = F[f][1], v2 = F[f][2];
        const int32_t m01 = (int32_t)nv + en.ni[3 * f], m12 = (int32_t)nv + en.ni[3 * f + 1],
                      m20 = (int32_t)nv + en.ni[3 * f + 2];
        G[4 * f] = {v0, m01, m20};
        G[4 * f + 1] = {v1, m12, m01};
        G[4 * f + 2] = {m01, m12, m20};
        G[4 * f + 3] = {m12, v2, m20};
    }
    V.swap(W);
    F.swap(G);
    return true;
}

void false_barycentric(std::vector<double> &V, std::vector<Tri> &F) {
    const size_t nv = V.size() / 3, nf = F.size();
    V.resize(3 * (nv + nf));
    std::vector<Tri> G(3 * nf);
    for (size_t f = 0; f < nf; f++) {
        for (int c = 0; c < 3; c++) {
            double acc = 0.0;
            for (int j = 0; j < 3; j++) acc = acc + V[3 * (size_t)F[f][j] + c];
            V[3 * (nv + f) + c] = acc / 3.0;
        }
        const int32_t b = (int32_t)(nv + f);
        G[3 * f] = {F[f][0], F[f][1], b};
        G[3 * f + 1] = {F[f][1], F[f][2], b};
        G[3 * f + 2] = {F[f][2], F[f][0], b};
    }
    F.swap(G);
}

// ---- decimation ----

struct Decimator {
    std::vector<double> V;
    std::vector<Tri> F;
    std::vector<EdgeKey> E;                        // (min, max); {-1, -1} once removed
    std::vector<std::array<int32_t, 2>> EF, EI;    // faces of an edge and the corner opposite it in each
    std::vector<std::array<int32_t, 3>> FE;        // edge opposite each corner of a face
    std::vector<std::vector<int32_t>> VF;          // live faces at each vertex
    std::vector<int32_t> stamp;                    // queue entries of an edge are current at its stamp; -1: removed
    std::vector<std::array<double, 3>> place;
    using Entry = std::tuple<double, int32_t, int32_t>;
    std::priority_queue<Entry, std::vector<Entry>, std::greater<Entry>> Q;

    bool build(std::string &err) {
        const size_t nf = F.size();
        std::vector<std::pair<EdgeKey, int64_t>> all;  // (edge opposite corner k of f, 3 f + k)
        all.reserve(3 * nf);
        for (size_t f = 0; f < nf; f++)
            for (int k = 0; k < 3; k++) all.push_back({key(F[f][(k + 1) % 3], F[f][(k + 2) % 3]), (int64_t)(3 * f + k)});
        std::sort(all.begin(), all.end(), [](const auto &x, const auto &y) {
            return x.first < y.first || (x.first == y.first && x.second < y.second);
        });
        FE.assign(nf, {-1, -1, -1});
        for (size_t i = 0; i < all.size(); i += 2) {
            if (i + 1 >= all.size() || !(all[i + 1].first == all[i].first) ||
                (i + 2 < all.size() && all[i + 2].first == all[i].first)) {
                err = "decimation needs a closed edge-manifold mesh: edge (" + std::to_string(all[i].first.a) + ", " +
                      std::to_string(all[i].first.b) + ") does not have exactly two faces";
                return false;
            }
            const int64_t p = all[i].second, q = all[i + 1].second;
            // consistent orientation: the two faces run the edge in opposite directions
            const int32_t pf = (int32_t)(p / 3), pk = (int32_t)(p % 3), qf = (int32_t)(q / 3), qk = (int32_t)(q % 3);
            if (F[pf][(pk + 1) % 3] != F[qf][(qk + 2) % 3]) {
                err = "decimation needs a consistently oriented mesh: faces " + std::to_string(pf) + " and " +
                      std::to_string(qf) + " run their shared edge the same way";
                return false;
            }
            const int32_t e = (int32_t)E.size();
            E.push_back(all[i].first);
            EF.push_back({pf, qf});
            EI.push_back({pk, qk});
            FE[pf][pk] = e, FE[qf][qk] = e;
        }
        VF.assign(V.size() / 3, {});
        for (size_t f = 0; f < nf; f++)
            for (int k = 0; k < 3; k++) VF[F[f][k]].push_back((int32_t)f);
        stamp.assign(E.size(), 0);
        place.assign(E.size(), {0, 0, 0});
        for (size_t e = 0; e < E.size(); e++) Q.emplace(cost_and_place((int32_t)e), (int32_t)e, 0);
        return true;
    }

    double cost_and_place(int32_t e) {
        const double *a = &V[3 * (size_t)E[e].a], *b = &V[3 * (size_t)E[e].b];
        const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
        for (int c = 0; c < 3; c++) place[e][c] = 0.5 * (a[c] + b[c]);
        const double sq = dx * dx + dy * dy;
        return std::sqrt(sq + dz * dz);
    }

    std::vector<int32_t> closed_ring(int32_t v) const {
        std::vector<int32_t> r;
        for (int32_t f : VF[v])
            for (int k = 0; k < 3; k++) r.push_back(F[f][k]);
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        return r;
    }

    bool collapsible(int32_t s, int32_t d) const {
        const std::vector<int32_t> Ns = closed_ring(s), Nd = closed_ring(d);
        std::vector<int32_t> both;
        std::set_intersection(Ns.begin(), Ns.end(), Nd.begin(), Nd.end(), std::back_inserter(both));
        if (both.size() != 4) return false;
        if (Ns.size() == 4 && Nd.size() == 4) {
            std::vector<int32_t> any;
            std::set_union(Ns.begin(), Ns.end(), Nd.begin(), Nd.end(), std::back_inserter(any));
            if (any.size() == 4) return false;  // a lone tetrahedron
        }
        return true;
    }

    void drop_face_at(int32_t v, int32_t f) {
        auto &l = VF[v];
        l.erase(std::find(l.begin(), l.end(), f));
    }

    void kill_edge(int32_t e) {
        E[e] = {-1, -1};
        EF[e] = {-1, -1};
        EI[e] = {-1, -1};
        stamp[e] = -1;
    }

    // Collapses edge e; false (nothing changed) if the link condition refuses it.
    bool collapse(int32_t e) {
        const int32_t s = E[e].a, d = E[e].b;
        if (!collapsible(s, d)) return false;
        std::vector<int32_t> around = VF[s];
        around.insert(around.end(), VF[d].begin(), VF[d].end());
        std::sort(around.begin(), around.end());
        around.erase(std::unique(around.begin(), around.end()), around.end());
        for (int side = 0; side < 2; side++) {
            const int32_t f = EF[e][side];
            int ks = -1, kd = -1;
            for (int k = 0; k < 3; k++) {
                if (F[f][k] == s) ks = k;
                if (F[f][k] == d) kd = k;
            }
            const int32_t ed = FE[f][ks];  // opposite s: the edge (d, x), removed
            const int32_t es = FE[f][kd];  // opposite d: the edge (s, x), kept
            const int side_g = EF[ed][0] == f ? 1 : 0;
            const int32_t g = EF[ed][side_g], kg = EI[ed][side_g];
            FE[g][kg] = es;
            const int side_f = EF[es][0] == f ? 0 : 1;
            EF[es][side_f] = g;
            EI[es][side_f] = kg;
            kill_edge(ed);
            for (int k = 0; k < 3; k++) drop_face_at(F[f][k], f);
            F[f] = {-1, -1, -1};
        }
        for (int32_t f : VF[d]) {
            for (int k = 0; k < 3; k++)
                if (F[f][k] == d) F[f][k] = s;
            for (int k = 0; k < 3; k++) {
                EdgeKey &x = E[FE[f][k]];
                if (x.a == d || x.b == d) x = key(x.a == d ? s : x.a, x.b == d ? s : x.b);
            }
            VF[s].push_back(f);
        }
        VF[d].clear();
        const std::array<double, 3> p = place[e];
        for (int c = 0; c < 3; c++) V[3 * (size_t)s + c] = p[c], V[3 * (size_t)d + c] = p[c];
        kill_edge(e);
        std::vector<int32_t> touched;
        for (int32_t f : around)
            if (F[f][0] >= 0)
                for (int k = 0; k < 3; k++) touched.push_back(FE[f][k]);
        std::sort(touched.begin(), touched.end());
        touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
        for (int32_t x : touched) {
            const double c = cost_and_place(x);
            Q.emplace(c, x, ++stamp[x]);
        }
        return true;
    }

    void run(int64_t max_faces) {
        int64_t m = (int64_t)F.size();
        for (;;) {
            int32_t e = -1;
            while (!Q.empty()) {
                const Entry top = Q.top();
                if (std::get<0>(top) == std::numeric_limits<double>::infinity()) break;
                Q.pop();
                if (std::get<2>(top) == stamp[std::get<1>(top)]) {
                    e = std::get<1>(top);
                    break;
                }
            }
            if (e < 0) return;  // nothing left that can be collapsed
            if (collapse(e)) {
                m -= 2;
                if (m <= max_faces) return;
            } else {
                Q.emplace(std::numeric_limits<double>::infinity(), e, ++stamp[e]);
            }
        }
    }
};

}  // namespace

int subdivide_mesh(std::vector<double> &V, std::vector<int32_t> &F, int scheme, int levels, std::string &err) {
    if (scheme != CQBEM_SUBDIV_UPSAMPLE && scheme != CQBEM_SUBDIV_LOOP && scheme != CQBEM_SUBDIV_FALSE_BARYCENTRIC)
        return err = "unknown subdivision scheme " + std::to_string(scheme), (int)CQBEM_ERR_ARG;
    if (levels < 0) return err = "negative subdivision level", (int)CQBEM_ERR_ARG;
    std::vector<Tri> T(F.size() / 3);
    for (size_t f = 0; f < T.size(); f++) T[f] = {F[3 * f], F[3 * f + 1], F[3 * f + 2]};
    const int grow = scheme == CQBEM_SUBDIV_FALSE_BARYCENTRIC ? 3 : 4;
    {
        int64_t nf = (int64_t)T.size();  // the final face count, before any work
        for (int l = 0; l < levels; l++)
            if ((nf *= grow) > INT32_MAX / 4)
                return err = "subdivided mesh would exceed 2^29 faces", (int)CQBEM_ERR_ARG;
    }
    for (int l = 0; l < levels; l++) {
        // faces and vertices both stay below 2^31 / 4 (edges < 3/2 faces)
        if ((int64_t)T.size() * grow > INT32_MAX / 4 || (int64_t)(V.size() / 3) + 2 * (int64_t)T.size() > INT32_MAX / 4)
            return err = "subdivided mesh would exceed 2^29 faces or vertices", (int)CQBEM_ERR_ARG;
        if (scheme == CQBEM_SUBDIV_FALSE_BARYCENTRIC)
            false_barycentric(V, T);
        else if (!split_four(V, T, scheme == CQBEM_SUBDIV_LOOP, err))
            return (int)CQBEM_ERR_ARG;
    }
    F.resize(3 * T.size());
    for (size_t f = 0; f < T.size(); f++)
        for (int k = 0; k < 3; k++) F[3 * f + k] = T[f][k];
    return CQBEM_OK;
}

int decimate_mesh(std::vector<double> &V, std::vector<int32_t> &F, int64_t max_faces, std::string &err) {
    if (max_faces < 0) return err = "negative face count", (int)CQBEM_ERR_ARG;
    Decimator D;
    D.V = V;
    D.F.resize(F.size() / 3);
    for (size_t f = 0; f < D.F.size(); f++) D.F[f] = {F[3 * f], F[3 * f + 1], F[3 * f + 2]};
    if (!D.build(err)) return (int)CQBEM_ERR_ARG;
    D.run(max_faces);
    // surviving faces in order; referenced vertices in order
    const size_t nv = V.size() / 3;
    std::vector<int32_t> remap(nv, -1);
    for (const Tri &t : D.F)
        if (t[0] >= 0)
            for (int k = 0; k < 3; k++) remap[t[k]] = 0;
    int32_t n = 0;
    std::vector<double> W;
    for (size_t i = 0; i < nv; i++)
        if (remap[i] == 0) {
            remap[i] = n++;
            for (int c = 0; c < 3; c++) W.push_back(D.V[3 * i + c]);
        }
    std::vector<int32_t> G;
    for (const Tri &t : D.F)
        if (t[0] >= 0)
            for (int k = 0; k < 3; k++) G.push_back(remap[t[k]]);
    V.swap(W);
    F.swap(G);
    return CQBEM_OK;
}

}  // namespace cqbem

namespace {

int check_mesh(const double *V, int64_t nv, const int32_t *F, int64_t nf) {
    if (!V || !F || nv <= 0 || nf <= 0 || nv > INT32_MAX / 4 || nf > INT32_MAX / 4) return CQBEM_ERR_ARG;
    for (int64_t k = 0; k < 3 * nf; k++)
        if (F[k] < 0 || F[k] >= nv) return CQBEM_ERR_ARG;
    for (int64_t f = 0; f < nf; f++)  // a face repeating a vertex has no edges to split or collapse
        if (F[3 * f] == F[3 * f + 1] || F[3 * f + 1] == F[3 * f + 2] || F[3 * f + 2] == F[3 * f]) return CQBEM_ERR_ARG;
    return CQBEM_OK;
}

int copy_out(const std::vector<double> &Vr, const std::vector<int32_t> &Fr, double *V_out, int32_t *F_out,
             int64_t *nv_out, int64_t *nf_out) {
    const int64_t nv = (int64_t)(Vr.size() / 3), nf = (int64_t)(Fr.size() / 3);
    if (V_out || F_out) {
        if (*nv_out < nv || *nf_out < nf) return CQBEM_ERR_ARG;
        if (V_out) std::copy(Vr.begin(), Vr.end(), V_out);
        if (F_out) std::copy(Fr.begin(), Fr.end(), F_out);
    }
    *nv_out = nv, *nf_out = nf;
    return CQBEM_OK;
}

}  // namespace

extern "C" int cqbem_subdivide_mesh(const double *V, int64_t nv, const int32_t *F, int64_t nf, int32_t scheme,
                                    int32_t levels, double *V_out, int32_t *F_out, int64_t *nv_out, int64_t *nf_out) {
    if (!nv_out || !nf_out) return CQBEM_ERR_ARG;
    if (const int rc = check_mesh(V, nv, F, nf)) return rc;
    std::vector<double> Vr(V, V + 3 * nv);
    std::vector<int32_t> Fr(F, F + 3 * nf);
    std::string err;
    if (const int rc = cqbem::subdivide_mesh(Vr, Fr, scheme, levels, err)) return rc;
    return copy_out(Vr, Fr, V_out, F_out, nv_out, nf_out);
}

extern "C" int cqbem_decimate_mesh(const double *V, int64_t nv, const int32_t *F, int64_t nf, int64_t max_faces,
                                   double *V_out, int32_t *F_out, int64_t *nv_out, int64_t *nf_out) {
    if (!nv_out || !nf_out) return CQBEM_ERR_ARG;
    if (const int rc = check_mesh(V, nv, F, nf)) return rc;
    std::vector<double> Vr(V, V + 3 * nv);
    std::vector<int32_t> Fr(F, F + 3 * nf);
    std::string err;
    if (const int rc = cqbem::decimate_mesh(Vr, Fr, max_faces, err)) return rc;
    return copy_out(Vr, Fr, V_out, F_out, nv_out, nf_out);
}
