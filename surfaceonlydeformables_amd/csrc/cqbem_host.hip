// Host driver behind include/cqbem.h: job preparation (normals, vertex stars,
// quadrature and transform tables), device memory under an HBM budget,
// row-slab scheduling of the quadrature kernels, the inverse CQ transform and
// the C / B post-processing.  Replaces compute_{H,G,B_trans,B_angular}_matrices
// and compute_B_euler_matrices of the reference's
// ElastodynamicsCQBEMMatrixCollection (src/elastodynamics_cqbem_*.cu,
// src/elastodynamics_cqbem_matrix_collection.cpp:280-305).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <deque>
#include <exception>
#include <stdexcept>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cqbem.h"
#include "cqbem_galerkin.hpp"
#include "cqbem_wavelet.hpp"
#include "cqbem_pool.hpp"
#include "cqbem_kernels.hpp"
#include "cqbem_mixed.hpp"
#include "cqbem_rules.hpp"

using namespace cqbem;

// Device scratch of cqbem_assemble_noalloc: one array per GPU the library is loaded on, held until the process
// ends.  Host code never uses its name as a device address: the address of each device's copy comes from
// hipGetSymbolAddress (cqbem_create, with that device current) and is kept in noalloc_scratch_ptr.
static constexpr size_t kNoAllocScratchBytes = (size_t)CQBEM_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static constexpr int kMaxDevices = 64;
static std::mutex noalloc_setup_mu;
static void *noalloc_scratch_ptr[kMaxDevices];  // per device, set once
static std::mutex noalloc_call_mu[kMaxDevices];  // calls on one device share its array: serialized in-process

namespace {

struct Err {
    int code;
    std::string msg;
};

#define HIPCHK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            throw Err{e_ == hipErrorOutOfMemory ? CQBEM_ERR_NOMEM : CQBEM_ERR_HIP,                       \
                      std::string(#expr) + ": " + hipGetErrorString(e_)};                               \
    } while (0)

// Device allocation owned by a job; released on scope exit.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool pooled = false;  // from the running job's arena (returned with it)
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { dev_free(p, pooled); }
    void alloc(size_t n) {
        dev_free(p, pooled), p = nullptr;
        bytes = n;
        if (n) HIPCHK(dev_alloc(&p, n, &pooled));
    }
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

template <class T> void upload(DBuf &b, const std::vector<T> &v, hipStream_t st) {
    b.alloc(v.size() * sizeof(T));
    if (!v.empty()) HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

bool mixed_float_g();  // CQBEM_MIXED_G=float (below)

void check_inputs(const cqbem_mesh *mesh, const cqbem_params *p) {
    if (!mesh || !p) throw Err{CQBEM_ERR_ARG, "null mesh or params"};
    if (!mesh->V || !mesh->F || mesh->nv < 3 || mesh->nf < 1) throw Err{CQBEM_ERR_ARG, "empty mesh"};
    if (mesh->nv > (1 << 24) || mesh->nf > (1 << 26)) throw Err{CQBEM_ERR_ARG, "mesh too large"};
    for (int64_t k = 0; k < 3 * mesh->nf; k++)
        if (mesh->F[k] < 0 || mesh->F[k] >= mesh->nv) throw Err{CQBEM_ERR_ARG, "face index out of range"};
    if (const int *vm = mesh->vertex_map_inverse) {
        // copies of one original vertex must sit at one position (the split only duplicates vertices)
        std::unordered_map<int, int64_t> first;
        for (int64_t k = 0; k < mesh->nv; k++) {
            if (vm[k] < 0 || vm[k] >= mesh->nv) throw Err{CQBEM_ERR_ARG, "vertex_map_inverse entry out of range"};
            auto it = first.emplace(vm[k], k).first;
            for (int d = 0; d < 3; d++)
                if (mesh->V[3 * k + d] != mesh->V[3 * it->second + d])
                    throw Err{CQBEM_ERR_ARG, "vertices mapped to one original vertex differ in position"};
        }
    }
    if (!(p->c1 > 0) || !(p->c2 > 0) || !(p->rho > 0) || !(p->dt > 0) || !(p->R > 0))
        throw Err{CQBEM_ERR_ARG, "c1, c2, rho, dt and R must be positive"};
    if (p->L < 2 || p->L > 4096) throw Err{CQBEM_ERR_ARG, "L must be in [2, 4096]"};
    if (p->bdf != CQBEM_BDF1 && p->bdf != CQBEM_BDF2) throw Err{CQBEM_ERR_ARG, "bdf must be 1 or 2"};
    if (p->order < 1 || p->order > 5) throw Err{CQBEM_ERR_ARG, "order must be in [1, 5]"};
    if (p->subdivision < 1 || p->subdivision > 16) throw Err{CQBEM_ERR_ARG, "subdivision must be in [1, 16]"};
    if (p->precision != CQBEM_FP64 && p->precision != CQBEM_MIXED) throw Err{CQBEM_ERR_ARG, "unknown precision"};
    if (p->precision == CQBEM_MIXED && p->galerkin && mixed_float_g())
        throw Err{CQBEM_ERR_UNSUPPORTED, "CQBEM_MIXED_G=float is implemented for collocation only"};
    if (p->storage != CQBEM_STORE_F64 && p->storage != CQBEM_STORE_F32) throw Err{CQBEM_ERR_ARG, "unknown storage"};
    if ((p->which & ~CQBEM_ALL) != 0) throw Err{CQBEM_ERR_ARG, "unknown matrix bits in `which`"};
    if (p->quad_policy != CQBEM_QUAD_REFERENCE && p->quad_policy != CQBEM_QUAD_BANDED)
        throw Err{CQBEM_ERR_ARG, "unknown quad_policy"};
    if (p->quad_policy == CQBEM_QUAD_BANDED) {
        // the bands apply to the pair kernel: collocation G/H, in FP64 or CQBEM_MIXED with G in that kernel
        if (p->galerkin || (p->precision == CQBEM_MIXED && mixed_float_g()))
            throw Err{CQBEM_ERR_UNSUPPORTED, "CQBEM_QUAD_BANDED is implemented for collocation G/H in the pair kernel"};
        if (p->far_order < 0 || p->far_order > 5 || p->far_subdivision < 0 || p->far_subdivision > 16 ||
            !(p->band_kappa >= 0) || !(p->band_kappa_b >= 0))
            throw Err{CQBEM_ERR_ARG, "far rule or band threshold out of range"};
    }
    // the device sincos reduces |Im(s) r / c| <= max|s_l| diam / min(c1, c2) exactly below kTrigArgMax
    double lo[3] = {mesh->V[0], mesh->V[1], mesh->V[2]}, hi[3] = {lo[0], lo[1], lo[2]};
    for (int64_t k = 0; k < mesh->nv; k++)
        for (int d = 0; d < 3; d++) {
            const double x = mesh->V[3 * k + d];
            if (!std::isfinite(x)) throw Err{CQBEM_ERR_ARG, "non-finite vertex coordinate"};
            lo[d] = std::min(lo[d], x), hi[d] = std::max(hi[d], x);
        }
    const double diam = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                  (hi[2] - lo[2]) * (hi[2] - lo[2]));
    double smax = 0.0;
    for (const auto &s : cq_frequencies(p->L, p->R, p->dt, p->bdf)) smax = std::max(smax, std::abs(s));
    if (smax * diam / std::min(p->c1, p->c2) >= kTrigArgMax)
        throw Err{CQBEM_ERR_UNSUPPORTED, "max |s| * diameter / c exceeds 1e6: time step too small for the mesh size"};
}

}  // namespace

// --------------------------------------------------------------- context --
// A serving job: owns copies of its mesh and runs on one stream slot in a worker thread.
struct JobRecord {
    uint64_t id = 0;
    std::vector<double> V;
    std::vector<int32_t> F, vmap;
    cqbem_mesh mesh{};
    cqbem_params params{};
    cqbem_outputs out{};
    uint64_t bytes = 0;           // scratch reservation, released when the job finishes
    uint64_t resident_bytes = 0;  // outputs reservation of a resident job, released by cqbem_release
    bool resident = false;
    int slot = -1;
    std::atomic<int> done{0};
    std::atomic<int> cancel{0};  // set by cqbem_cancel; the worker stops at its next slab or phase boundary
    int status = 0;
    std::string error;
    std::thread worker;
};

struct cqbem_context {
    int device = 0;
    uint64_t budget = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    std::string last_error;
    double timings[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::mutex mu;
    // serving
    std::atomic<uint64_t> next_job{1};
    uint64_t reserved = 0;
    std::unordered_map<uint64_t, std::shared_ptr<JobRecord>> jobs;
    std::vector<hipStream_t> slots;
    std::vector<uint64_t> slot_job;
    DevicePool pool;  // job buffers, reused across jobs
    // scenes: the branches of a captured graph are recorded on these streams (fork and join through the events)
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_ev;
};

namespace {

struct Prepared {
    int nv, nf, L, Lf;
    std::vector<double> N, area;
    std::vector<int> vs_ptr, vs_face, sg_ptr, sg_face;
    // pair form: patches of pair_patch_size() vertices, passes of pair_pass_faces() faces
    std::vector<int> pair_patch_pass, pair_patch_verts;
    std::vector<int2> pair_pass_info;   // (first face entry, nfaces | nsub << 16)
    std::vector<int4> pair_pass_faces;  // (face, corner codes), code = slot | sub-step << 10, -1 outside
    int pair_npatches = 0;
    int npts = 0;               // points of the composite triangle rule
    std::vector<double4> quad;  // the rule, padded to an even count with a zero-weight point
    std::vector<double4> pair_pts;  // per padded point: (b0, b1, b2, w), (phi_F own, phi_F partner, 0, 0)
    // CQBEM_QUAD_BANDED: the far rule's pair point table and, per patch, its bounding sphere's centre and the far
    // distance (radius + kappa h)
    std::vector<double4> far_pair_pts;
    int far_npts = 0;  // points of the far rule (unpadded)
    std::vector<double4> patch_sphere;
    std::vector<double> patch_weight;  // per patch: sum over its faces of 1 / (patches listing the face)
    // the B kernels' face bands: the far rule's points (b0, b1, b2, w) and, per face, its bounding sphere's centre
    // and the far distance (radius + kappa h)
    std::vector<double4> far_quad, face_sphere;  // per patch: sum over its faces of 1 / (patches listing the face)
    std::vector<double2> line;
    std::vector<Freq> freqs;
    std::vector<double2> tab;
    Mat mat;
};

// CQBEM_QUAD_BANDED default: far when the row point is kBandKappa largest face diameters beyond the patch's bounding
// sphere.  Tuned on icosphere level 4 at L = 32 (surfaceonlydeformables_amd/tune_bands.py, table in
// surfaceonlydeformables_amd/band_tables/): the smallest kappa whose added error stays within 5% of the reference
// rule's largest per-weight error against a rule of twice the subdivision (3.4% for G, 0.6% for H).
constexpr double kBandKappa = 6.0;
// The B kernels take face bands only when asked (band_kappa_b > 0).  B_trans and B_angular are closed-surface
// integrals that cancel to a small remainder (their reference-rule error is 50x below G's), so the far rule's
// per-face errors, small in absolute terms, stand out against it: tuned the same way, no threshold up to 16 edge
// lengths keeps the added error within 5% of B's reference-rule error on icosphere level 4.

Prepared prepare(const cqbem_mesh *mesh, const cqbem_params *p) {
    Prepared P;
    P.nv = (int)mesh->nv, P.nf = (int)mesh->nf, P.L = p->L, P.Lf = p->L / 2 + 1;
    const double *V = mesh->V;
    const int *F = mesh->F;
    const int *vm = mesh->vertex_map_inverse;
    P.N.resize(3 * (size_t)P.nf), P.area.resize(P.nf);
    for (int f = 0; f < P.nf; f++) {
        const V3 c = crs(sub(vtx(V, F[3 * f + 1]), vtx(V, F[3 * f])), sub(vtx(V, F[3 * f + 2]), vtx(V, F[3 * f])));
        const double n = nrm(c);
        if (!(n > 0)) throw Err{CQBEM_ERR_ARG, "degenerate (zero-area) face " + std::to_string(f)};
        P.N[3 * f] = c.x / n, P.N[3 * f + 1] = c.y / n, P.N[3 * f + 2] = c.z / n, P.area[f] = 0.5 * n;
    }
    // vertex stars
    P.vs_ptr.assign(P.nv + 1, 0);
    for (int k = 0; k < 3 * P.nf; k++) P.vs_ptr[F[k] + 1]++;
    for (int v = 0; v < P.nv; v++) P.vs_ptr[v + 1] += P.vs_ptr[v];
    P.vs_face.resize(3 * (size_t)P.nf);
    {
        std::vector<int> pos(P.vs_ptr.begin(), P.vs_ptr.end() - 1);
        for (int f = 0; f < P.nf; f++)
            for (int c = 0; c < 3; c++) P.vs_face[pos[F[3 * f + c]]++] = (f << 2) | c;
    }
    // singular faces per collocation vertex: faces with a corner equal to (or mapped like) the vertex;
    // the first such corner in face order is the polar/Guiggiani centre (G_collocation.cu:78-88)
    P.sg_ptr.assign(P.nv + 1, 0);
    if (!vm) {
        P.sg_ptr = P.vs_ptr;
        P.sg_face = P.vs_face;
    } else {
        std::unordered_map<int, std::vector<int>> groups;
        for (int v = 0; v < P.nv; v++) groups[vm[v]].push_back(v);
        std::vector<std::vector<int>> lists(P.nv);
        for (int i = 0; i < P.nv; i++) {
            std::vector<int> faces;
            for (int v : groups[vm[i]])
                for (int t = P.vs_ptr[v]; t < P.vs_ptr[v + 1]; t++) faces.push_back(P.vs_face[t] >> 2);
            std::sort(faces.begin(), faces.end());
            faces.erase(std::unique(faces.begin(), faces.end()), faces.end());
            for (int f : faces) {
                int c = 0;
                while (c < 3 && !(F[3 * f + c] == i || vm[F[3 * f + c]] == vm[i])) c++;
                lists[i].push_back((f << 2) | c);
            }
        }
        for (int i = 0; i < P.nv; i++) P.sg_ptr[i + 1] = P.sg_ptr[i] + (int)lists[i].size();
        for (auto &l : lists) P.sg_face.insert(P.sg_face.end(), l.begin(), l.end());
    }
    // Column vertices in the Morton order of their positions (21 bits per axis over the bounding box).
    std::vector<int> order(P.nv);
    {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int v = 0; v < P.nv; v++)
            for (int k = 0; k < 3; k++) lo[k] = std::min(lo[k], V[3 * v + k]), hi[k] = std::max(hi[k], V[3 * v + k]);
        const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
        auto spread = [](uint64_t x) {
            x &= 0x1fffff;
            x = (x | x << 32) & 0x1f00000000ffffULL;
            x = (x | x << 16) & 0x1f0000ff0000ffULL;
            x = (x | x << 8) & 0x100f00f00f00f00fULL;
            x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
            x = (x | x << 2) & 0x1249249249249249ULL;
            return x;
        };
        std::vector<std::pair<uint64_t, int>> key(P.nv);
        for (int v = 0; v < P.nv; v++) {
            uint64_t c[3];
            for (int k = 0; k < 3; k++) c[k] = (uint64_t)std::min(2097151.0, (V[3 * v + k] - lo[k]) / ext * 2097151.0);
            key[v] = {spread(c[0]) | spread(c[1]) << 1 | spread(c[2]) << 2, v};
        }
        std::sort(key.begin(), key.end());
        for (int r = 0; r < P.nv; r++) order[r] = key[r].second;
    }
    // Pair-form patches, grown one at a time.  A patch lists every face with a corner in it; the list is ordered by
    // a vertex-disjoint colouring and cut into passes of pair_pass_faces() faces.  A patch starts from the first
    // unassigned vertex in Morton order and takes, among the unassigned vertices that share a face with it, the
    // one adding the fewest faces to its list, then the one with the most assigned neighbours, then the first in
    // Morton order.  It closes at pair_patch_size() vertices, or when the next vertex would take its face list
    // beyond kPairTargetPasses full passes; a patch whose neighbours run out (a mesh in several pieces) goes on
    // from the next unassigned vertex in Morton order.  Compact patches list fewer faces twice than runs of the
    // Morton order (icosphere level 4: 1.14 face listings per face instead of 1.22).  A patch's slots are its
    // vertices in ascending order.  Within a pass every in-patch corner gets a sub-step: its rank among the pass's
    // updates of that vertex, in face order.
    {
        constexpr int kPairTargetPasses = 4;
        const int C = pair_patch_size(), PF = pair_pass_faces(), cap = kPairTargetPasses * PF;
        std::vector<int> nb_ptr(P.nv + 1, 0), nb;  // distinct neighbours: the other corners of the vertex's faces
        {
            std::vector<int> tmp;
            for (int v = 0; v < P.nv; v++) {
                tmp.clear();
                for (int t = P.vs_ptr[v]; t < P.vs_ptr[v + 1]; t++)
                    for (int c = 0; c < 3; c++)
                        if (F[3 * (P.vs_face[t] >> 2) + c] != v) tmp.push_back(F[3 * (P.vs_face[t] >> 2) + c]);
                std::sort(tmp.begin(), tmp.end());
                tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
                nb.insert(nb.end(), tmp.begin(), tmp.end());
                nb_ptr[v + 1] = (int)nb.size();
            }
        }
        std::vector<int> rank(P.nv), slot(P.nv), patch(P.nv, -1), mark(P.nf, -1);
        std::vector<int> fresh(P.nv, 0), nadj(P.nv, 0), infront(P.nv, -1);  // fresh: faces not yet in the list
        for (int r = 0; r < P.nv; r++) rank[order[r]] = r;
        auto count_fresh = [&](int u, int pc) {
            int n = 0;
            for (int t = P.vs_ptr[u]; t < P.vs_ptr[u + 1]; t++) n += mark[P.vs_face[t] >> 2] != pc;
            return n;
        };
        auto better = [&](int a, int b) {
            if (fresh[a] != fresh[b]) return fresh[a] < fresh[b];
            if (nadj[a] != nadj[b]) return nadj[a] > nadj[b];
            return rank[a] < rank[b];
        };
        std::vector<std::vector<int>> lists;
        std::vector<int> front, members;
        int next = 0;  // Morton cursor: every vertex before it is assigned
        for (;;) {
            while (next < P.nv && patch[order[next]] >= 0) next++;
            if (next == P.nv) break;
            const int pc = (int)lists.size();
            lists.emplace_back();
            front.clear(), members.clear();
            while ((int)members.size() < C) {
                int v;
                size_t at = front.size();
                if (front.empty()) {
                    while (next < P.nv && patch[order[next]] >= 0) next++;
                    if (next == P.nv) break;
                    v = order[next], fresh[v] = count_fresh(v, pc);
                } else {
                    at = 0;
                    for (size_t k = 1; k < front.size(); k++)
                        if (better(front[k], front[at])) at = k;
                    v = front[at];
                }
                if (!members.empty() && (int)lists[pc].size() + fresh[v] > cap) break;
                if (at < front.size()) front[at] = front.back(), front.pop_back();
                for (int t = P.vs_ptr[v]; t < P.vs_ptr[v + 1]; t++) {
                    const int f = P.vs_face[t] >> 2;
                    if (mark[f] == pc) continue;
                    mark[f] = pc, lists[pc].push_back(f);
                    for (int c = 0; c < 3; c++)
                        if (infront[F[3 * f + c]] == pc) fresh[F[3 * f + c]]--;
                }
                patch[v] = pc, members.push_back(v);
                for (int k = nb_ptr[v]; k < nb_ptr[v + 1]; k++) {
                    const int u = nb[k];
                    if (patch[u] >= 0) continue;
                    nadj[u]++;
                    if (infront[u] != pc) infront[u] = pc, fresh[u] = count_fresh(u, pc), front.push_back(u);
                }
            }
            std::sort(members.begin(), members.end());
            for (size_t k = 0; k < members.size(); k++) slot[members[k]] = (int)k;
        }
        P.pair_npatches = (int)lists.size();
        P.pair_patch_verts.assign((size_t)P.pair_npatches * C, -1);
        for (int v = 0; v < P.nv; v++) P.pair_patch_verts[(size_t)patch[v] * C + slot[v]] = v;
        P.pair_patch_pass.assign(1, 0);
        std::vector<int> cnt(C);
        std::vector<std::vector<int>> used;  // per colour, per slot: taken
        for (int pc = 0; pc < P.pair_npatches; pc++) {
            auto &l = lists[pc];
            std::sort(l.begin(), l.end());
            // greedy colouring: faces of one colour share no in-patch vertex; ordering the passes by colour keeps
            // the updates per vertex and pass (the sub-steps) near 2-3 instead of the full valence
            used.clear();
            std::vector<std::pair<int, int>> keyed;
            keyed.reserve(l.size());
            for (int f : l) {
                int colour = 0;
                for (;; colour++) {
                    if (colour == (int)used.size()) used.emplace_back(C, 0);
                    bool ok = true;
                    for (int c = 0; c < 3; c++)
                        if (patch[F[3 * f + c]] == pc && used[colour][slot[F[3 * f + c]]]) ok = false;
                    if (ok) break;
                }
                for (int c = 0; c < 3; c++)
                    if (patch[F[3 * f + c]] == pc) used[colour][slot[F[3 * f + c]]] = 1;
                keyed.emplace_back(colour, f);
            }
            std::sort(keyed.begin(), keyed.end());
            for (size_t k = 0; k < l.size(); k++) l[k] = keyed[k].second;
            for (size_t b = 0; b < l.size(); b += PF) {
                const int nf = (int)std::min<size_t>(PF, l.size() - b);
                std::fill(cnt.begin(), cnt.end(), 0);
                int nsub = 0;
                const int first = (int)P.pair_pass_faces.size();
                for (int k = 0; k < nf; k++) {
                    const int f = l[b + k];
                    int code[3];
                    for (int c = 0; c < 3; c++) {
                        const int v = F[3 * f + c];
                        if (patch[v] != pc) {
                            code[c] = -1;
                            continue;
                        }
                        const int sub = cnt[slot[v]]++;
                        nsub = std::max(nsub, sub + 1);
                        code[c] = slot[v] | sub << 10;
                    }
                    P.pair_pass_faces.push_back(make_int4(f, code[0], code[1], code[2]));
                }
                if (nsub > 0x7fff) throw Err{CQBEM_ERR_UNSUPPORTED, "vertex valence too large for the pair schedule"};
                P.pair_pass_info.push_back(make_int2(first, nf | nsub << 16));
            }
            P.pair_patch_pass.push_back((int)P.pair_pass_info.size());
        }
        if (p->quad_policy == CQBEM_QUAD_BANDED) {
            // per patch: bounding sphere of its faces (centre: mean of their corners) and largest face diameter
            const double kappa = p->band_kappa > 0 ? p->band_kappa : kBandKappa;
            for (int pc = 0; pc < P.pair_npatches; pc++) {
                double c[3] = {0, 0, 0};
                for (int f : lists[pc])
                    for (int k = 0; k < 3; k++)
                        for (int d = 0; d < 3; d++) c[d] += V[3 * F[3 * f + k] + d];
                const double inv = 1.0 / (3.0 * (double)lists[pc].size());
                for (double &x : c) x *= inv;
                double rad = 0, h = 0;
                for (int f : lists[pc])
                    for (int k = 0; k < 3; k++) {
                        const V3 a = vtx(V, F[3 * f + k]), b = vtx(V, F[3 * f + (k + 1) % 3]);
                        rad = std::max(rad, nrm(sub(a, V3{c[0], c[1], c[2]})));
                        h = std::max(h, nrm(sub(a, b)));
                    }
                // the radius is widened by 1e-12 of itself so that rounding never puts a corner of the patch's own
                // faces outside it (such a row is then never far from its own patch, whatever kappa)
                P.patch_sphere.push_back(make_double4(c[0], c[1], c[2], rad * (1 + 1e-12) + kappa * h));
            }
            // per face: the same construction (centre: mean of its corners; radius: the farthest corner; h: its
            // longest edge), for the B kernels, which sum each row over the faces
            const double kappa_b = p->band_kappa_b;
            for (int f = 0; f < P.nf && kappa_b > 0; f++) {
                double c[3] = {0, 0, 0};
                for (int k = 0; k < 3; k++)
                    for (int d = 0; d < 3; d++) c[d] += V[3 * F[3 * f + k] + d];
                for (double &x : c) x *= 1.0 / 3.0;
                double rad = 0, h = 0;
                for (int k = 0; k < 3; k++) {
                    const V3 a = vtx(V, F[3 * f + k]), b = vtx(V, F[3 * f + (k + 1) % 3]);
                    rad = std::max(rad, nrm(sub(a, V3{c[0], c[1], c[2]})));
                    h = std::max(h, nrm(sub(a, b)));
                }
                P.face_sphere.push_back(make_double4(c[0], c[1], c[2], rad * (1 + 1e-12) + kappa_b * h));
            }
            std::vector<int> mult(P.nf, 0);
            for (const auto &l : lists)
                for (int f : l) mult[f]++;
            for (const auto &l : lists) {
                double w = 0;
                for (int f : l) w += 1.0 / mult[f];
                P.patch_weight.push_back(w);
            }
        }
    }
    for (const TriPoint &t : composite_tri_rule(p->order, p->subdivision)) P.quad.push_back(make_double4(t.b[0], t.b[1], t.b[2], t.w));
    P.npts = (int)P.quad.size();
    if (P.quad.size() & 1) P.quad.push_back(make_double4(1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 0.0));
    // pair-kernel point table: even points go to lane A (corner 0), odd points to lane B (corner 2)
    for (size_t q = 0; q < P.quad.size(); q++) {
        const double4 &Q = P.quad[q], &Qp = P.quad[q ^ 1];
        P.pair_pts.push_back(Q);
        P.pair_pts.push_back((q & 1) ? make_double4(Q.z, Qp.z, 0.0, 0.0) : make_double4(Q.x, Qp.x, 0.0, 0.0));
    }
    if (p->quad_policy == CQBEM_QUAD_BANDED) {
        std::vector<double4> fq;
        for (const TriPoint &t : composite_tri_rule(p->far_order > 0 ? p->far_order : p->order,
                                                    p->far_subdivision > 0 ? p->far_subdivision : 1))
            fq.push_back(make_double4(t.b[0], t.b[1], t.b[2], t.w));
        P.far_npts = (int)fq.size();
        P.far_quad = fq;
        if (fq.size() & 1) fq.push_back(make_double4(1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 0.0));
        for (size_t q = 0; q < fq.size(); q++) {
            const double4 &Q = fq[q], &Qp = fq[q ^ 1];
            P.far_pair_pts.push_back(Q);
            P.far_pair_pts.push_back((q & 1) ? make_double4(Q.z, Qp.z, 0.0, 0.0) : make_double4(Q.x, Qp.x, 0.0, 0.0));
        }
    }
    for (auto &xw : composite_line01(p->order, p->subdivision)) P.line.push_back(make_double2(xw.first, xw.second));
    const auto s = cq_frequencies(p->L, p->R, p->dt, p->bdf);
    for (auto sl : s) {
        Freq f;
        f.s = mkc(sl.real(), sl.imag());
        const std::complex<double> a = sl / p->c1, b = sl / p->c2, is = 1.0 / sl, is2 = is * is;
        f.sc1 = mkc(a.real(), a.imag()), f.sc2 = mkc(b.real(), b.imag());
        f.is = mkc(is.real(), is.imag()), f.is2 = mkc(is2.real(), is2.imag());
        P.freqs.push_back(f);
    }
    for (auto t : transform_table(p->L, p->R)) P.tab.push_back(make_double2(t.real(), t.imag()));
    Mat &m = P.mat;
    m.c1 = p->c1, m.c2 = p->c2, m.rho = p->rho;
    m.ic1 = 1.0 / p->c1, m.ic2 = 1.0 / p->c2, m.ic1s = m.ic1 * m.ic1, m.ic2s = m.ic2 * m.ic2;
    m.kap = p->c2 * p->c2 / (p->c1 * p->c1), m.six_c2s = 6.0 * p->c2 * p->c2;
    m.g_scale = 1.0 / (4.0 * kPi * p->rho);
    return P;
}

struct DevMesh {
    DBuf V, F, N, area, vs_ptr, vs_face, sg_ptr, sg_face, vmap, quad, line, freqs, tab;
    DBuf pair_patch_pass, pair_pass_info, pair_pass_faces, pair_patch_verts, pair_pts;
    DBuf far_pts, patch_sphere, far_quad, face_sphere;  // CQBEM_QUAD_BANDED
    Geom g;
};

// Collocation G/H regular-face kernel variant: "pair" (default: compact vertex patches, each face evaluated once
// per patch by a lane pair) or "star" (one lane per column vertex over its vertex star, each face
// evaluated once per corner); CQBEM_COLLOC_KERNEL=star selects the latter.
bool colloc_star() {
    const char *e = std::getenv("CQBEM_COLLOC_KERNEL");
    return e && !std::strcmp(e, "star");
}

// CQBEM_MIXED's far-field G: by default in the FP64 pair kernel together with H (G shares H's exponentials there,
// which costs less than a float kernel of its own on MI355X); CQBEM_MIXED_G=float selects the float star kernel
// (the reference's own typedef for G).
bool mixed_float_g() {
    const char *e = std::getenv("CQBEM_MIXED_G");
    return e && !std::strcmp(e, "float");
}

// The band descriptor of a prepared job (empty unless CQBEM_QUAD_BANDED).
Bands bands_of(const DevMesh &d, const Prepared &P) {
    Bands b;
    if (P.far_pair_pts.empty()) return b;
    b.far_pts = d.far_pts.as<double4>(), b.far_rounds = (int)P.far_pair_pts.size() / 4;
    b.sphere = d.patch_sphere.as<double4>();
    return b;
}

// The B kernels' face bands of a prepared job (empty unless CQBEM_QUAD_BANDED with band_kappa_b > 0).
FaceBands fbands_of(const DevMesh &d, const Prepared &P) {
    FaceBands b;
    if (P.face_sphere.empty()) return b;
    b.far_quad = d.far_quad.as<double4>(), b.far_npts = P.far_npts, b.sphere = d.face_sphere.as<double4>();
    return b;
}

// Host replica of the pair kernel's per-tile rule choice (CQBEM_QUAD_BANDED): whether tile (patch pc, row point x)
// takes the far rule.
bool tile_is_far(const Prepared &P, int pc, const double *x) {
    if (P.far_pair_pts.empty()) return false;
    return beyond_sphere(x[0], x[1], x[2], P.patch_sphere[pc]);
}

// Host replica of the B kernels' per-face rule choice: whether face f takes the far rule for row point x.
bool face_is_far(const Prepared &P, int f, const double *x) {
    if (P.face_sphere.empty()) return false;
    return beyond_sphere(x[0], x[1], x[2], P.face_sphere[f]);
}

// Regular-face evaluations of row vertex i over all frequencies (each face and point once, whichever patches list
// the face): (nf - singular faces) x points x Lf for the reference rule; per tile's rule when banded (the singular
// faces of a row lie in near tiles: the patch listing them contains the row vertex).
double regular_evals(const Prepared &P, const double *V, int i) {
    const int nsing = P.sg_ptr[i + 1] - P.sg_ptr[i];
    if (P.far_pair_pts.empty()) return (double)(P.nf - nsing) * P.npts * P.Lf;
    double e = -(double)nsing * P.npts;
    for (int pc = 0; pc < P.pair_npatches; pc++) e += P.patch_weight[pc] * (tile_is_far(P, pc, V + 3 * i) ? P.far_npts : P.npts);
    return e * P.Lf;
}

hipError_t launch_gh(const DevMesh &d, const Prepared &P, const Freq *freqs, int l0, int nl, int row0, int nrows, cd *UG,
                     cd *UH, long long ld, long long lstride, hipStream_t st) {
    if (colloc_star() && P.far_pair_pts.empty())
        return launch_colloc_gh(d.g, d.quad.as<double4>(), P.npts, freqs, l0, nl, P.mat, row0, nrows, UG, UH, ld, lstride, st);
    return launch_colloc_gh_pair(d.g, d.pair_pts.as<double4>(), (int)P.quad.size(), freqs, l0, nl, P.mat, row0, nrows,
                                 nullptr, d.pair_patch_pass.as<int>(), d.pair_pass_info.as<int2>(), d.pair_pass_faces.as<int4>(),
                                 d.pair_patch_verts.as<int>(), P.pair_npatches, UG, UH, ld, lstride, st, bands_of(d, P));
}

void upload_mesh(DevMesh &d, const cqbem_mesh *mesh, const Prepared &P, hipStream_t st) {
    d.V.alloc(sizeof(double) * 3 * P.nv);
    HIPCHK(hipMemcpyAsync(d.V.p, mesh->V, sizeof(double) * 3 * P.nv, hipMemcpyHostToDevice, st));
    d.F.alloc(sizeof(int) * 3 * P.nf);
    HIPCHK(hipMemcpyAsync(d.F.p, mesh->F, sizeof(int) * 3 * P.nf, hipMemcpyHostToDevice, st));
    if (mesh->vertex_map_inverse) {
        d.vmap.alloc(sizeof(int) * P.nv);
        HIPCHK(hipMemcpyAsync(d.vmap.p, mesh->vertex_map_inverse, sizeof(int) * P.nv, hipMemcpyHostToDevice, st));
    }
    upload(d.N, P.N, st), upload(d.area, P.area, st), upload(d.vs_ptr, P.vs_ptr, st), upload(d.vs_face, P.vs_face, st);
    upload(d.sg_ptr, P.sg_ptr, st), upload(d.sg_face, P.sg_face, st), upload(d.quad, P.quad, st), upload(d.line, P.line, st);
    upload(d.freqs, P.freqs, st), upload(d.tab, P.tab, st);
    upload(d.pair_patch_pass, P.pair_patch_pass, st), upload(d.pair_pass_info, P.pair_pass_info, st);
    upload(d.pair_pass_faces, P.pair_pass_faces, st), upload(d.pair_patch_verts, P.pair_patch_verts, st);
    upload(d.pair_pts, P.pair_pts, st);
    if (!P.far_pair_pts.empty()) {
        upload(d.far_pts, P.far_pair_pts, st), upload(d.patch_sphere, P.patch_sphere, st);
        if (!P.face_sphere.empty()) upload(d.far_quad, P.far_quad, st), upload(d.face_sphere, P.face_sphere, st);
    }
    Geom &g = d.g;
    g.V = d.V.as<double>(), g.F = d.F.as<int>(), g.N = d.N.as<double>(), g.area = d.area.as<double>();
    g.vs_ptr = d.vs_ptr.as<int>(), g.vs_face = d.vs_face.as<int>(), g.sg_ptr = d.sg_ptr.as<int>(), g.sg_face = d.sg_face.as<int>();
    g.vmap = mesh->vertex_map_inverse ? d.vmap.as<int>() : nullptr;
    g.nv = P.nv, g.nf = P.nf;
}

size_t matrix_bytes(int nv, int L) { return (size_t)L * 9 * (size_t)nv * nv * sizeof(double); }

// Rows of collocation vertices per slab so that the G/H spectra of one slab fit `bytes`.
int slab_rows(int nv, int Lf, int nmat, size_t bytes) {
    const size_t per_row = (size_t)nmat * Lf * 3 * 3 * (size_t)nv * sizeof(cd);
    size_t rows = per_row ? bytes / per_row : (size_t)nv;
    rows = std::min<size_t>(std::max<size_t>(rows, 1), (size_t)nv);
    return (int)std::min<size_t>(rows, 65535);
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t st;
    double *acc;
    bool on;
    Timer(hipStream_t s, double *dst) : st(s), acc(dst), on(phase_timers_on()) {
        if (!on) return;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        HIPCHK(hipEventRecord(a, st));
    }
    ~Timer() {
        if (!on) return;
        if (hipEventRecord(b, st) == hipSuccess && hipEventSynchronize(b) == hipSuccess) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, a, b) == hipSuccess) *acc += ms;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

// Bound of the Galerkin per-test-point row scratch within a job's scratch budget: larger chunks only save
// launches, the traffic per face is the same.
size_t galerkin_budget(size_t slab_budget) { return std::min<size_t>(slab_budget, (size_t)8 << 30); }

// Device footprint of a job: scratch spectra (one slab), B spectra/outputs and, when
// outputs are not caller-provided device buffers, the staged outputs.
uint64_t footprint(const cqbem_mesh *mesh, const cqbem_params *p, int outputs_on_device, size_t slab_budget) {
    const int nv = (int)mesh->nv, L = p->L, Lf = L / 2 + 1;
    const int nmat = ((p->which & CQBEM_G) ? 1 : 0) + ((p->which & CQBEM_H) ? 1 : 0);
    uint64_t bytes = 0;
    if (nmat) {
        if (p->galerkin) {  // full-matrix spectra
            bytes += (uint64_t)nmat * Lf * 9 * (uint64_t)nv * nv * sizeof(cd);
        } else {
            const int rows = slab_rows(nv, Lf, nmat, slab_budget);
            bytes += (uint64_t)nmat * Lf * 3 * (uint64_t)rows * 3 * nv * sizeof(cd);
        }
        if (!outputs_on_device) bytes += (uint64_t)nmat * matrix_bytes(nv, L);
    }
    if (p->galerkin) {
        // touching-pair scratch: one 270-complex block per ordered pair of touching faces and frequency
        uint64_t npairs = 0;
        for (const auto &c : touching_candidates(mesh->F, (int)mesh->nf, nv, mesh->vertex_map_inverse)) npairs += c.size();
        bytes += npairs * ((uint64_t)Lf * 270 * sizeof(cd) + 256);
        const int npts = (int)composite_tri_rule(p->order, p->subdivision).size();
        const bool B = (p->which & (CQBEM_B_TRANS | CQBEM_B_ANGULAR)) != 0;
        bytes += galerkin_scratch_bytes(nv, (int)mesh->nf, npts, Lf, (p->which & CQBEM_G) != 0, (p->which & CQBEM_H) != 0, B,
                                        galerkin_budget(slab_budget));
    }
    bytes += (uint64_t)Lf * 3 * nv * 12 * sizeof(cd) + (uint64_t)L * 3 * nv * 18 * sizeof(double);
    bytes += (uint64_t)mesh->nf * 64 + (uint64_t)nv * 64 + (1 << 20);
    return bytes;
}

// Runs one job on stream `st`.  `slab_budget` bounds the Laplace-domain scratch.
void run_job(cqbem_context *ctx, hipStream_t st, const cqbem_mesh *mesh, const cqbem_params *p, const cqbem_outputs *out,
             size_t slab_budget, double *tm, bool sync_outputs, JobArena *fixed_arena = nullptr,
             const std::atomic<int> *cancel = nullptr) {
    // a cancelled serving job ends at the next check (the arena then waits for the work already launched)
    auto check_cancel = [cancel]() {
        if (cancel && cancel->load()) throw Err{CQBEM_ERR_CANCELLED, "job cancelled"};
    };
    // every device buffer of the job comes from the context's pool and goes back to it when the stream is done
    // (CQBEM_POOL=0: plain hipMalloc / hipFree per buffer), or from the caller's fixed region
    static const bool use_pool = [] {
        const char *e = std::getenv("CQBEM_POOL");
        return !(e && !std::strcmp(e, "0"));
    }();
    JobArena arena(&ctx->pool, st);
    ArenaScope arena_scope(fixed_arena ? fixed_arena : use_pool ? &arena : nullptr);
    const auto t_host0 = std::chrono::steady_clock::now();
    Prepared P = prepare(mesh, p);
    const int nv = P.nv, L = P.L, Lf = P.Lf;
    const long long n3 = 3LL * nv;
    DevMesh d;
    upload_mesh(d, mesh, P, st);
    std::vector<double> Chost;
    DBuf dC, dCidx, dCval;
    const bool wantG = (p->which & CQBEM_G) && out->G, wantH = (p->which & CQBEM_H) && out->H;
    const bool wantBt = (p->which & CQBEM_B_TRANS) && out->B_trans;
    const bool wantBa = (p->which & CQBEM_B_ANGULAR) && (out->B_angular || out->B);
    const bool wantB = wantBt || wantBa || out->B;
    tm[5] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count();

    // outputs: caller device buffers, or staged device buffers copied back at the end; with CQBEM_STORE_F32 the
    // caller's G and H buffers hold float (the transform rounds each weight once)
    DBuf sG, sH, sBt, sBa, sB6;
    double *G = nullptr, *H = nullptr;
    const bool f32 = p->storage == CQBEM_STORE_F32;
    if (f32 && !out->on_device) throw Err{CQBEM_ERR_ARG, "CQBEM_STORE_F32 needs device outputs"};
    if (wantG) {
        if (out->on_device) G = out->G;
        else sG.alloc(matrix_bytes(nv, L)), G = sG.as<double>();
    }
    if (wantH) {
        if (out->on_device) H = out->H;
        else sH.alloc(matrix_bytes(nv, L)), H = sH.as<double>();
    }
    // weights n of rows from `row` on: out[n][row..] = transform of the spectra
    auto transform_into = [&](double *dst, const cd *spec, long long E, long long spec_ls, long long row_off) {
        if (f32)
            return launch_transform(spec, E, spec_ls, Lf, d.tab.as<double2>(), L, reinterpret_cast<float *>(dst) + row_off,
                                    n3 * n3, ctx->num_cus, st);
        return launch_transform(spec, E, spec_ls, Lf, d.tab.as<double2>(), L, dst + row_off, n3 * n3, ctx->num_cus, st);
    };
    DBuf specB, specBa;
    double *Wbt = nullptr, *Wba = nullptr, *W6 = nullptr;
    if (wantB) {
        specB.alloc((size_t)Lf * n3 * 3 * sizeof(cd));
        specBa.alloc((size_t)Lf * n3 * 9 * sizeof(cd));
        if (out->on_device && out->B_trans) Wbt = out->B_trans;
        else sBt.alloc((size_t)L * n3 * 3 * sizeof(double)), Wbt = sBt.as<double>();
        if (out->on_device && out->B_angular) Wba = out->B_angular;
        else sBa.alloc((size_t)L * n3 * 9 * sizeof(double)), Wba = sBa.as<double>();
        if (out->B) {
            if (out->on_device) W6 = out->B;
            else sB6.alloc((size_t)L * n3 * 6 * sizeof(double)), W6 = sB6.as<double>();
        }
    }
    const Geom &g = d.g;
    const double4 *quad = d.quad.as<double4>();
    const double2 *line = d.line.as<double2>();
    const Freq *freqs = d.freqs.as<Freq>();
    const double2 *tab = d.tab.as<double2>();
    const int npts = P.npts, nline = (int)P.line.size();
    // mixed precision: far-field B (and, with CQBEM_MIXED_G=float, G) in float, the rest in double
    const bool mixed = p->precision == CQBEM_MIXED, g32 = mixed && mixed_float_g();
    DBuf dfreq32;
    if (mixed) upload(dfreq32, freqs_f32(P.freqs), st);
    const FreqF *freqs32 = dfreq32.as<FreqF>();
    const MatF mat32 = mat_f32(P.mat);

    if (!p->galerkin) {
        const int nmat = (wantG ? 1 : 0) + (wantH ? 1 : 0);
        if (nmat) {
            const int rows = slab_rows(nv, Lf, nmat, slab_budget);
            const long long lstride = 3LL * rows * n3;  // one frequency of one slab
            DBuf specG, specH;
            if (wantG) specG.alloc((size_t)Lf * lstride * sizeof(cd));
            if (wantH) specH.alloc((size_t)Lf * lstride * sizeof(cd));
            for (int r0 = 0; r0 < nv; r0 += rows) {
                check_cancel();
                const int nr = std::min(rows, nv - r0);
                {
                    Timer t(st, &tm[1]);
                    if (!g32) {
                        HIPCHK(launch_gh(d, P, freqs, 0, Lf, r0, nr, specG.as<cd>(), specH.as<cd>(), n3, lstride, st));
                    } else {
                        if (wantH) HIPCHK(launch_gh(d, P, freqs, 0, Lf, r0, nr, nullptr, specH.as<cd>(), n3, lstride, st));
                        if (wantG)
                            HIPCHK(launch_colloc_g32(g, quad, npts, freqs32, 0, Lf, mat32, r0, nr, specG.as<cd>(), n3, lstride,
                                                     st));
                    }
                }
                tm[6] += 1;
                for (int i = r0; i < r0 + nr; i++) tm[7] += regular_evals(P, mesh->V, i);
                {
                    Timer t(st, &tm[3]);
                    if (wantG)
                        HIPCHK(launch_colloc_singular(0, g, line, nline, freqs, 0, Lf, P.mat, r0, nr, specG.as<cd>(), n3,
                                                      lstride, nullptr, nullptr, st));
                    if (wantH)
                        HIPCHK(launch_colloc_singular(2, g, line, nline, freqs, 0, Lf, P.mat, r0, nr, specH.as<cd>(), n3,
                                                      lstride, nullptr, nullptr, st));
                }
                {
                    Timer t(st, &tm[4]);
                    const long long E = 3LL * nr * n3;
                    if (wantG) HIPCHK(transform_into(G, specG.as<cd>(), E, lstride, 3LL * r0 * n3));
                    if (wantH) HIPCHK(transform_into(H, specH.as<cd>(), E, lstride, 3LL * r0 * n3));
                }
            }
        }
        if (wantB) {
            check_cancel();
            {
                Timer t(st, &tm[2]);
                if (mixed)
                    HIPCHK(launch_colloc_b32(g, quad, npts, freqs32, 0, Lf, mat32, 0, nv, specB.as<cd>(), specBa.as<cd>(), st,
                                             fbands_of(d, P)));
                else
                    HIPCHK(launch_colloc_b(g, quad, npts, freqs, 0, Lf, P.mat, 0, nv, specB.as<cd>(), specBa.as<cd>(), st,
                                           fbands_of(d, P)));
            }
            {
                Timer t(st, &tm[3]);
                HIPCHK(launch_colloc_singular(1, g, line, nline, freqs, 0, Lf, P.mat, 0, nv, nullptr, n3, 0, specB.as<cd>(),
                                              specBa.as<cd>(), st));
            }
        }
        if (wantH) {
            Chost = mesh->vertex_map_inverse
                        ? c_matrix_collocation_split(mesh->V, nv, mesh->F, P.nf, P.N.data(), mesh->vertex_map_inverse, p->nu)
                        : c_matrix_collocation(mesh->V, nv, mesh->F, P.nf, P.N.data(), p->nu);
            upload(dC, Chost, st);
            if (f32)
                HIPCHK(launch_add_c_colloc(reinterpret_cast<float *>(H), dC.as<double>(), nv, st));
            else
                HIPCHK(launch_add_c_colloc(H, dC.as<double>(), nv, st));
        }
    } else {
        GalerkinJob gj{g, quad, npts, line, nline, freqs, Lf, P.mat, ctx->num_cus, d.pair_pts.as<double4>(), (int)P.quad.size(),
                       d.pair_patch_pass.as<int>(), d.pair_pass_info.as<int2>(), d.pair_pass_faces.as<int4>(),
                       d.pair_patch_verts.as<int>(), P.pair_npatches, galerkin_budget(slab_budget)};
        if (mixed) gj.freqs32 = freqs32, gj.mat32 = mat32;  // far-field B at the test points in float
        check_cancel();
        const int nmat = (wantG ? 1 : 0) + (wantH ? 1 : 0);
        if (nmat || wantB) {
            // Galerkin spectra are assembled for the full matrix (face pairs scatter to any row)
            DBuf specG, specH;
            const long long lstride = n3 * n3;
            if (wantG) specG.alloc((size_t)Lf * lstride * sizeof(cd));
            if (wantH) specH.alloc((size_t)Lf * lstride * sizeof(cd));
            run_galerkin(gj, mesh->F, mesh->vertex_map_inverse, specG.as<cd>(), specH.as<cd>(),
                         wantB ? specB.as<cd>() : nullptr, wantB ? specBa.as<cd>() : nullptr, tm, st);
            Timer t(st, &tm[4]);
            if (wantG) HIPCHK(transform_into(G, specG.as<cd>(), n3 * n3, n3 * n3, 0));
            if (wantH) HIPCHK(transform_into(H, specH.as<cd>(), n3 * n3, n3 * n3, 0));
        }
        if (wantH) {
            // 1/2 P1 mass matrix per component (elastics_bem_matrix_collection.cpp:447-467), summed per entry on the host
            std::unordered_map<long long, double> acc;
            for (int f = 0; f < P.nf; f++) {
                const double A2 = 2.0 * P.area[f];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        for (int k = 0; k < 3; k++)
                            acc[(3LL * mesh->F[3 * f + a] + k) * n3 + 3LL * mesh->F[3 * f + b] + k] += (a == b) ? A2 / 24.0 : A2 / 48.0;
            }
            std::vector<long long> idx;
            std::vector<double> val;
            for (auto &kv : acc) idx.push_back(kv.first), val.push_back(kv.second);
            upload(dCidx, idx, st), upload(dCval, val, st);
            if (f32)
                HIPCHK(launch_add_triplets(reinterpret_cast<float *>(H), dCidx.as<long long>(), dCval.as<double>(),
                                           (long long)idx.size(), st));
            else
                HIPCHK(launch_add_triplets(H, dCidx.as<long long>(), dCval.as<double>(), (long long)idx.size(), st));
        }
    }
    if (wantB) {
        Timer t(st, &tm[4]);
        HIPCHK(launch_transform(specB.as<cd>(), n3 * 3, n3 * 3, Lf, tab, L, Wbt, n3 * 3, ctx->num_cus, st));
        HIPCHK(launch_transform(specBa.as<cd>(), n3 * 9, n3 * 9, Lf, tab, L, Wba, n3 * 9, ctx->num_cus, st));
        HIPCHK(launch_b_post(Wbt, Wba, W6, g.V, nv, L, st));
    }
    if (!out->on_device && sync_outputs) {
        const auto t0 = std::chrono::steady_clock::now();
        if (wantG) HIPCHK(hipMemcpyAsync(out->G, G, matrix_bytes(nv, L), hipMemcpyDeviceToHost, st));
        if (wantH) HIPCHK(hipMemcpyAsync(out->H, H, matrix_bytes(nv, L), hipMemcpyDeviceToHost, st));
        if (wantB && out->B_trans) HIPCHK(hipMemcpyAsync(out->B_trans, Wbt, (size_t)L * n3 * 3 * 8, hipMemcpyDeviceToHost, st));
        if (wantB && out->B_angular) HIPCHK(hipMemcpyAsync(out->B_angular, Wba, (size_t)L * n3 * 9 * 8, hipMemcpyDeviceToHost, st));
        if (W6) HIPCHK(hipMemcpyAsync(out->B, W6, (size_t)L * n3 * 6 * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        tm[5] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    HIPCHK(hipStreamSynchronize(st));
}

// Stream capture of a launch sequence into a HIP graph that is launched once and waited for.  With capture off
// (CQBEM_SCENE_GRAPH=0) the same launches run directly on their streams and launch() only waits.
bool scene_graphs_on() {
    const char *e = std::getenv("CQBEM_SCENE_GRAPH");
    return !(e && !std::strcmp(e, "0"));
}
struct GraphRun {
    hipStream_t st;
    bool capture;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    explicit GraphRun(hipStream_t s) : st(s), capture(scene_graphs_on()) {
        if (capture) HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    }
    void launch() {
        if (capture) {
            HIPCHK(hipStreamEndCapture(st, &graph));
            HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            HIPCHK(hipGraphLaunch(exec, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        capture = false;
    }
    ~GraphRun() {
        if (capture && !graph) {  // capture abandoned by an exception: end it so the stream is usable again
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(st, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

// Concurrent branches of an uncaptured launch sequence: fork() makes the side streams wait for what `st` has
// queued so far, on(i) is the stream of branch i (round robin), join() makes `st` wait for every branch.  Inactive
// (every branch on `st`) while `st` is being captured: the captured scene graphs are single-stream.
struct Branches {
    hipStream_t st;
    bool active;
    const std::vector<hipStream_t> &side;
    const std::vector<hipEvent_t> &ev;  // side.size() + 1 events
    Branches(cqbem_context *ctx, hipStream_t s, bool captured)
        : st(s), active(!captured && !ctx->side.empty()), side(ctx->side), ev(ctx->side_ev) {}
    void fork() {
        if (!active) return;
        HIPCHK(hipEventRecord(ev.back(), st));
        for (hipStream_t s : side) HIPCHK(hipStreamWaitEvent(s, ev.back(), 0));
    }
    hipStream_t on(size_t i) const { return active ? side[i % side.size()] : st; }
    void join() {
        if (!active) return;
        for (size_t i = 0; i < side.size(); i++) {
            HIPCHK(hipEventRecord(ev[i], side[i]));
            HIPCHK(hipStreamWaitEvent(st, ev[i], 0));
        }
    }
};

// Runs fn(0..n-1) on up to 16 host threads; the first exception is rethrown once all have finished.
template <class Fn> void parallel_for(size_t n, Fn fn) {
    const size_t nt = std::min<size_t>(n, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex em;
    auto work = [&]() {
        for (size_t i; (i = next++) < n;) {
            try {
                fn(i);
            } catch (...) {
                std::lock_guard<std::mutex> l(em);
                if (!err) err = std::current_exception();
            }
        }
    };
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; t++) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
    if (err) std::rethrow_exception(err);
}

// A scene: every object's G/H regular-face tiles go through shared launches of the pair kernel (k_scene_gh_tiles:
// all objects' tiles in one grid, largest pieces first), in waves whose spectra fit the scratch budget; each wave's
// launches -- the shared pair kernel, the float G kernel of objects that use it (CQBEM_MIXED_G=float), the singular
// faces and the transforms of
// every piece -- are captured in one HIP graph and launched once.  B, C and the B post-processing of all objects
// follow in a last graph.  Galerkin objects run one after another through run_job.
void run_scene(cqbem_context *ctx, hipStream_t st, int n, const cqbem_mesh *meshes, const cqbem_params *params,
               const cqbem_outputs *outs, double *tm) {
    if (ctx->side.empty() && !scene_graphs_on()) {  // the branches' streams (default priority), made once per context
        constexpr int kBranches = 8;
        ctx->side.resize(kBranches), ctx->side_ev.resize(kBranches + 1);
        for (auto &s : ctx->side) HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        for (auto &e : ctx->side_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    JobArena arena(&ctx->pool, st);
    ArenaScope arena_scope(&arena);
    struct Obj {
        int k;
        Prepared P;
        DevMesh d;
        DBuf specB, specBa, dC, dfreq32, sBt, sBa;
        double *Wbt = nullptr, *Wba = nullptr;  // B_trans / B_angular weights: the caller's, or staged for B
        bool wantG, wantH, wantB, mixed, g32, f32;
        MatF mat32;
        std::vector<double> C;  // jump term, host copy
    };
    std::vector<std::unique_ptr<Obj>> objs;
    std::vector<int> galerkin;
    size_t fixed = 0;
    const auto t_host0 = std::chrono::steady_clock::now();
    for (int k = 0; k < n; k++) {
        check_inputs(&meshes[k], &params[k]);
        if (!outs[k].on_device) throw Err{CQBEM_ERR_ARG, "assemble_scene: outputs must be device buffers"};
        if (params[k].galerkin) {
            galerkin.push_back(k);
            continue;
        }
        objs.push_back(std::make_unique<Obj>());
        objs.back()->k = k;
    }
    // host tables of every object (mesh tables, rules, frequencies, jump term) on host threads
    parallel_for(objs.size(), [&](size_t j) {
        Obj *o = objs[j].get();
        const int k = o->k;
        const cqbem_params *p = &params[k];
        const cqbem_outputs *out = &outs[k];
        o->P = prepare(&meshes[k], p);
        o->wantG = (p->which & CQBEM_G) && out->G, o->wantH = (p->which & CQBEM_H) && out->H;
        o->wantB = ((p->which & CQBEM_B_TRANS) && out->B_trans) || ((p->which & CQBEM_B_ANGULAR) && out->B_angular) || out->B;
        o->mixed = p->precision == CQBEM_MIXED, o->f32 = p->storage == CQBEM_STORE_F32;
        o->g32 = o->mixed && mixed_float_g();
        o->mat32 = mat_f32(o->P.mat);
        if (o->wantH)
            o->C = meshes[k].vertex_map_inverse
                       ? c_matrix_collocation_split(meshes[k].V, o->P.nv, meshes[k].F, o->P.nf, o->P.N.data(),
                                                    meshes[k].vertex_map_inverse, p->nu)
                       : c_matrix_collocation(meshes[k].V, o->P.nv, meshes[k].F, o->P.nf, o->P.N.data(), p->nu);
    });
    // their device copies, from this thread (the buffers come from the scene's arena)
    for (auto &op : objs) {
        Obj *o = op.get();
        const int k = o->k;
        const cqbem_params *p = &params[k];
        const cqbem_outputs *out = &outs[k];
        upload_mesh(o->d, &meshes[k], o->P, st);
        if (o->mixed) upload(o->dfreq32, freqs_f32(o->P.freqs), st);
        const long long n3 = 3LL * o->P.nv;
        if (o->wantB) {
            o->specB.alloc((size_t)o->P.Lf * n3 * 3 * sizeof(cd));
            o->specBa.alloc((size_t)o->P.Lf * n3 * 9 * sizeof(cd));
            if (out->B_trans) o->Wbt = out->B_trans;
            else o->sBt.alloc((size_t)o->P.L * n3 * 3 * sizeof(double)), o->Wbt = o->sBt.as<double>();
            if (out->B_angular) o->Wba = out->B_angular;
            else o->sBa.alloc((size_t)o->P.L * n3 * 9 * sizeof(double)), o->Wba = o->sBa.as<double>();
        }
        if (o->wantH) upload(o->dC, o->C, st);
        fixed += footprint(&meshes[k], p, 1, 0);
    }
    tm[5] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count();
    // the spectra budget: what the HBM budget and the free memory leave, at most 48 GiB
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    const uint64_t avail = std::min<uint64_t>(ctx->budget, (uint64_t)fr + ctx->pool.cached_bytes());
    if (fixed >= avail) throw Err{CQBEM_ERR_NOMEM, "assemble_scene: the objects do not fit the HBM budget"};
    const uint64_t W = std::min<uint64_t>(avail - fixed, (uint64_t)48 << 30);
    // pieces: row slabs of each object, sized to the budget
    struct Piece {
        Obj *o;
        int row0, nrows;
        uint64_t bytes;
        long long tiles;
    };
    std::vector<Piece> pieces;
    for (auto &op : objs) {
        Obj *o = op.get();
        const int nmat = (o->wantG ? 1 : 0) + (o->wantH ? 1 : 0);
        if (!nmat) continue;
        const uint64_t per_row = (uint64_t)nmat * o->P.Lf * 9 * (uint64_t)o->P.nv * sizeof(cd);
        const int rows = (int)std::min<uint64_t>(std::min<uint64_t>(std::max<uint64_t>(W / per_row, 1), o->P.nv), 65535);
        for (int r0 = 0; r0 < o->P.nv; r0 += rows) {
            const int nr = std::min(rows, o->P.nv - r0);
            pieces.push_back(Piece{o, r0, nr, per_row * nr, (long long)o->P.pair_npatches * o->P.Lf * nr});
        }
    }
    // waves: first fit of the pieces, largest first, into the budget
    std::sort(pieces.begin(), pieces.end(), [](const Piece &a, const Piece &b) { return a.bytes > b.bytes; });
    std::vector<std::vector<Piece>> waves;
    std::vector<uint64_t> wave_bytes;
    for (const Piece &pc : pieces) {
        size_t w = 0;
        while (w < waves.size() && wave_bytes[w] + pc.bytes > W) ++w;
        if (w == waves.size()) waves.emplace_back(), wave_bytes.push_back(0);
        waves[w].push_back(pc), wave_bytes[w] += pc.bytes;
    }
    DBuf spec, dpieces, counter;
    if (!waves.empty()) {
        spec.alloc(*std::max_element(wave_bytes.begin(), wave_bytes.end()));
        counter.alloc(sizeof(unsigned));
    }
    for (auto &wave : waves) {
        // per piece: its spectra in the wave's region, its tiles in one of the two shared launches
        std::vector<ScenePiece> gh, honly, gonly;
        struct Slab {
            Piece pc;
            cd *sG, *sH;
            long long lstride;
        };
        std::vector<Slab> slabs;
        size_t off = 0;
        for (const Piece &pc : wave) {
            Obj *o = pc.o;
            const long long n3 = 3LL * o->P.nv, lstride = 3LL * pc.nrows * n3;
            cd *sG = nullptr, *sH = nullptr;
            if (o->wantG) sG = reinterpret_cast<cd *>(spec.as<char>() + off), off += (size_t)o->P.Lf * lstride * sizeof(cd);
            if (o->wantH) sH = reinterpret_cast<cd *>(spec.as<char>() + off), off += (size_t)o->P.Lf * lstride * sizeof(cd);
            slabs.push_back(Slab{pc, sG, sH, lstride});
            // objects with the float G kernel: H through the shared launch, G through the float kernel
            const bool g_in_pair = o->wantG && !o->g32;
            if (!g_in_pair && !o->wantH) continue;
            ScenePiece sp{o->d.g, o->d.pair_pts.as<double4>(), (int)o->P.quad.size() / 2, o->d.freqs.as<Freq>(), o->P.mat,
                          pc.row0, pc.nrows, o->P.Lf, o->P.pair_npatches, o->d.pair_patch_pass.as<int>(),
                          o->d.pair_pass_info.as<int2>(), o->d.pair_pass_faces.as<int4>(), o->d.pair_patch_verts.as<int>(),
                          g_in_pair ? sG : nullptr, o->wantH ? sH : nullptr, n3, lstride, 0u, bands_of(o->d, o->P)};
            (g_in_pair ? (o->wantH ? gh : gonly) : honly).push_back(sp);
        }
        auto number = [](std::vector<ScenePiece> &v) {
            long long base = 0;
            for (auto &sp : v) {
                sp.base = (unsigned)base;
                base += (long long)sp.npatches * sp.nl * sp.nrows;
            }
            if (base >= (1LL << 31)) throw Err{CQBEM_ERR_UNSUPPORTED, "assemble_scene: too many tiles in one wave"};
            return base;
        };
        const long long ngh = number(gh), nho = number(honly), ngo = number(gonly);
        std::vector<ScenePiece> all(gh);
        all.insert(all.end(), honly.begin(), honly.end());
        all.insert(all.end(), gonly.begin(), gonly.end());
        upload(dpieces, all, st);
        HIPCHK(hipStreamSynchronize(st));  // the piece table is on the device before the capture starts
        const ScenePiece *dgh = dpieces.as<ScenePiece>(), *dho = dgh + gh.size(), *dgo = dho + honly.size();
        const auto t0 = std::chrono::steady_clock::now();
        {
            // the shared launches run one after another on the stream (they share the tile counter)
            GraphRun gr(st);
            if (ngh) HIPCHK(launch_scene_gh_pair(dgh, (int)gh.size(), ngh, counter.as<unsigned>(), true, true, ctx->num_cus, st));
            if (nho) HIPCHK(launch_scene_gh_pair(dho, (int)honly.size(), nho, counter.as<unsigned>(), false, true, ctx->num_cus, st));
            if (ngo) HIPCHK(launch_scene_gh_pair(dgo, (int)gonly.size(), ngo, counter.as<unsigned>(), true, false, ctx->num_cus, st));
            // each slab's float G, singular faces and transforms: one branch per slab (uncaptured), after the shared
            // launches
            Branches br(ctx, st, gr.capture);
            br.fork();
            size_t bi = 0;
            for (const Slab &sb : slabs) {
                const hipStream_t st = br.on(bi++);
                Obj *o = sb.pc.o;
                const Geom &g = o->d.g;
                const long long n3 = 3LL * o->P.nv;
                const Freq *freqs = o->d.freqs.as<Freq>();
                const double2 *line = o->d.line.as<double2>();
                const int nline = (int)o->P.line.size(), Lf = o->P.Lf, L = o->P.L;
                if (o->wantG && o->g32)
                    HIPCHK(launch_colloc_g32(g, o->d.quad.as<double4>(), o->P.npts, o->dfreq32.as<FreqF>(), 0, Lf, o->mat32,
                                             sb.pc.row0, sb.pc.nrows, sb.sG, n3, sb.lstride, st));
                if (o->wantG)
                    HIPCHK(launch_colloc_singular(0, g, line, nline, freqs, 0, Lf, o->P.mat, sb.pc.row0, sb.pc.nrows, sb.sG, n3,
                                                  sb.lstride, nullptr, nullptr, st));
                if (o->wantH)
                    HIPCHK(launch_colloc_singular(2, g, line, nline, freqs, 0, Lf, o->P.mat, sb.pc.row0, sb.pc.nrows, sb.sH, n3,
                                                  sb.lstride, nullptr, nullptr, st));
                const long long E = 3LL * sb.pc.nrows * n3, row_off = 3LL * sb.pc.row0 * n3;
                const cqbem_outputs &out = outs[o->k];
                const double2 *tab = o->d.tab.as<double2>();
                for (int m = 0; m < 2; m++) {
                    const cd *src = m == 0 ? sb.sG : sb.sH;
                    double *dst = m == 0 ? out.G : out.H;
                    if (!src || !(m == 0 ? o->wantG : o->wantH)) continue;
                    if (o->f32)
                        HIPCHK(launch_transform(src, E, sb.lstride, Lf, tab, L, reinterpret_cast<float *>(dst) + row_off, n3 * n3,
                                                ctx->num_cus, st));
                    else
                        HIPCHK(launch_transform(src, E, sb.lstride, Lf, tab, L, dst + row_off, n3 * n3, ctx->num_cus, st));
                }
            }
            br.join();
            gr.launch();
        }
        tm[1] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        tm[6] += 1;
        for (const Piece &pc : wave)
            for (int i = pc.row0; i < pc.row0 + pc.nrows; i++)
                tm[7] += regular_evals(pc.o->P, meshes[pc.o->k].V, i);
    }
    // B, the jump term and the B post-processing of every object, in one graph
    {
        const auto t0 = std::chrono::steady_clock::now();
        GraphRun gr(st);
        Branches br(ctx, st, gr.capture);  // one branch per object
        br.fork();
        size_t bi = 0;
        for (auto &op : objs) {
            const hipStream_t st = br.on(bi++);
            Obj *o = op.get();
            const cqbem_outputs &out = outs[o->k];
            const Geom &g = o->d.g;
            const int nv = o->P.nv, Lf = o->P.Lf, L = o->P.L;
            const long long n3 = 3LL * nv;
            const Freq *freqs = o->d.freqs.as<Freq>();
            if (o->wantH) {
                if (o->f32)
                    HIPCHK(launch_add_c_colloc(reinterpret_cast<float *>(out.H), o->dC.as<double>(), nv, st));
                else
                    HIPCHK(launch_add_c_colloc(out.H, o->dC.as<double>(), nv, st));
            }
            if (!o->wantB) continue;
            if (o->mixed)
                HIPCHK(launch_colloc_b32(g, o->d.quad.as<double4>(), o->P.npts, o->dfreq32.as<FreqF>(), 0, Lf, o->mat32, 0, nv,
                                         o->specB.as<cd>(), o->specBa.as<cd>(), st, fbands_of(o->d, o->P)));
            else
                HIPCHK(launch_colloc_b(g, o->d.quad.as<double4>(), o->P.npts, freqs, 0, Lf, o->P.mat, 0, nv, o->specB.as<cd>(),
                                       o->specBa.as<cd>(), st, fbands_of(o->d, o->P)));
            HIPCHK(launch_colloc_singular(1, g, o->d.line.as<double2>(), (int)o->P.line.size(), freqs, 0, Lf, o->P.mat, 0, nv,
                                          nullptr, n3, 0, o->specB.as<cd>(), o->specBa.as<cd>(), st));
            const double2 *tab = o->d.tab.as<double2>();
            HIPCHK(launch_transform(o->specB.as<cd>(), n3 * 3, n3 * 3, Lf, tab, L, o->Wbt, n3 * 3, ctx->num_cus, st));
            HIPCHK(launch_transform(o->specBa.as<cd>(), n3 * 9, n3 * 9, Lf, tab, L, o->Wba, n3 * 9, ctx->num_cus, st));
            HIPCHK(launch_b_post(o->Wbt, o->Wba, out.B, g.V, nv, L, st));
        }
        br.join();
        gr.launch();
        tm[2] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    // the Galerkin objects' scratch: what the spectra region (held until the scene ends) leaves of the budget
    const uint64_t held = spec.p ? spec.bytes : 0;
    for (int k : galerkin) {
        const size_t slab = (size_t)std::max<uint64_t>(W > held ? W - held : 0, (uint64_t)256 << 20);
        run_job(ctx, st, &meshes[k], &params[k], &outs[k], slab, tm, true);
    }
}

int fail(cqbem_context *ctx, const Err &e) {
    if (ctx) ctx->last_error = e.msg;
    return e.code;
}

}  // namespace

// ================================================================= C-ABI ===
extern "C" {

int cqbem_create(cqbem_context **out, int device, uint64_t hbm_budget_bytes) {
    if (!out) return CQBEM_ERR_ARG;
    *out = nullptr;
    auto ctx = std::make_unique<cqbem_context>();
    try {
        int n = 0;
        HIPCHK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw Err{CQBEM_ERR_ARG, "no such HIP device"};
        HIPCHK(hipSetDevice(device));
        ctx->device = device;
        if (device < kMaxDevices) {
            // the no-allocation entry point's scratch: this device's copy of the array, looked up once
            std::lock_guard<std::mutex> lk(noalloc_setup_mu);
            if (!noalloc_scratch_ptr[device]) {
                void *ptr = nullptr;
                HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
                noalloc_scratch_ptr[device] = ptr;
            }
        }
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        ctx->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(0.9 * (double)fr);
        HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    } catch (const Err &e) {
        return e.code;
    }
    *out = ctx.release();
    return CQBEM_OK;
}

int cqbem_destroy(cqbem_context *ctx) {
    if (!ctx) return CQBEM_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    for (auto &kv : ctx->jobs)
        if (kv.second->worker.joinable()) kv.second->worker.join();
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream), (void)hipStreamDestroy(ctx->stream);
    for (hipStream_t s : ctx->slots) (void)hipStreamSynchronize(s), (void)hipStreamDestroy(s);
    for (hipStream_t s : ctx->side) (void)hipStreamSynchronize(s), (void)hipStreamDestroy(s);
    for (hipEvent_t e : ctx->side_ev) (void)hipEventDestroy(e);
    delete ctx;
    return CQBEM_OK;
}

const char *cqbem_last_error(const cqbem_context *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int cqbem_frequencies(int32_t L, double R, double dt, int32_t bdf, double *s_out) {
    if (L < 2 || !(R > 0) || !(dt > 0) || (bdf != 1 && bdf != 2) || !s_out) return CQBEM_ERR_ARG;
    const auto s = cq_frequencies(L, R, dt, bdf);
    for (size_t l = 0; l < s.size(); l++) s_out[2 * l] = s[l].real(), s_out[2 * l + 1] = s[l].imag();
    return (int)s.size();
}

int cqbem_history(cqbem_context *ctx, int64_t n3, int32_t L, const double *G, const double *H, const double *B,
                  const double *p_hist, const double *u_hist, const double *acc_hist, int32_t nhist, double *b) {
    if (!ctx) return CQBEM_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        if (n3 <= 0 || L < 1 || nhist < 0 || !b) throw Err{CQBEM_ERR_ARG, "history: bad sizes or null output"};
        const int m = std::min<int>(nhist, L - 1);
        if (m > 0 && (((G && !p_hist) || (H && !u_hist) || (B && !acc_hist))))
            throw Err{CQBEM_ERR_ARG, "history: a matrix without its history"};
        HIPCHK(hipSetDevice(ctx->device));
        if (m == 0 || (!G && !H && !B)) {
            HIPCHK(hipMemsetAsync(b, 0, sizeof(double) * n3, ctx->stream));
        } else {
            HIPCHK(launch_history(G, H, B, p_hist, u_hist, acc_hist, n3, m, nhist, b, ctx->stream));
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
    } catch (const Err &e) {
        ctx->last_error = e.msg;
        return e.code;
    }
    return CQBEM_OK;
}

int cqbem_history_f32(cqbem_context *ctx, int64_t n3, int32_t L, const float *G, const float *H, const double *B,
                      const double *p_hist, const double *u_hist, const double *acc_hist, int32_t nhist, double *b) {
    if (!ctx) return CQBEM_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        if (n3 <= 0 || (n3 & 1) || L < 1 || nhist < 0 || !b) throw Err{CQBEM_ERR_ARG, "history_f32: bad sizes or null output"};
        const int m = std::min<int>(nhist, L - 1);
        if (m > 0 && (((G && !p_hist) || (H && !u_hist) || (B && !acc_hist))))
            throw Err{CQBEM_ERR_ARG, "history_f32: a matrix without its history"};
        HIPCHK(hipSetDevice(ctx->device));
        if (m == 0 || (!G && !H && !B)) {
            HIPCHK(hipMemsetAsync(b, 0, sizeof(double) * n3, ctx->stream));
        } else {
            const hipError_t e = launch_history_f32(G, H, B, p_hist, u_hist, acc_hist, n3, m, nhist, b, ctx->stream);
            if (e == hipErrorInvalidValue) throw Err{CQBEM_ERR_ARG, "history_f32: misaligned matrices or histories"};
            HIPCHK(e);
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
    } catch (const Err &e) {
        ctx->last_error = e.msg;
        return e.code;
    }
    return CQBEM_OK;
}

int cqbem_matvec(cqbem_context *ctx, int64_t n, const double *A, const double *x, double *y) {
    if (!ctx) return CQBEM_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        if (n <= 0 || !A || !x || !y) throw Err{CQBEM_ERR_ARG, "matvec: bad size or null pointer"};
        HIPCHK(hipSetDevice(ctx->device));
        HIPCHK(launch_matvec(A, x, n, y, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    } catch (const Err &e) {
        ctx->last_error = e.msg;
        return e.code;
    }
    return CQBEM_OK;
}

int cqbem_wavelet_compress(cqbem_context *ctx, int64_t n, const double *A, int64_t num_elems, int64_t *rowptr, int32_t *col,
                           double *val, int64_t *nnz) {
    if (!ctx) return CQBEM_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        if (n < 2 || n > (int64_t)1 << 24 || !A || !rowptr || !nnz || !col || !val)
            throw Err{CQBEM_ERR_ARG, "wavelet_compress: bad size or null pointer"};
        HIPCHK(hipSetDevice(ctx->device));
        *nnz = wavelet_compress_into(A, n, num_elems, 1.0, rowptr, col, val, ctx->stream);
    } catch (const Err &e) {
        ctx->last_error = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        ctx->last_error = e.what();
        return CQBEM_ERR_HIP;
    }
    return CQBEM_OK;
}

int cqbem_csr_matvec(cqbem_context *ctx, int64_t rows, const int64_t *rowptr, const int32_t *col, const double *val,
                     const double *x, double *y) {
    if (!ctx) return CQBEM_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        if (rows <= 0 || !rowptr || !x || !y) throw Err{CQBEM_ERR_ARG, "csr_matvec: bad size or null pointer"};
        HIPCHK(hipSetDevice(ctx->device));
        HIPCHK(launch_csr_matvec(rowptr, col, val, x, rows, y, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    } catch (const Err &e) {
        ctx->last_error = e.msg;
        return e.code;
    }
    return CQBEM_OK;
}

int cqbem_band_plan(const cqbem_mesh *mesh, const cqbem_params *params, int32_t *patch_of_vertex, uint8_t *far,
                    int32_t *npatches) {
    try {
        if (!mesh || !params || !npatches) throw Err{CQBEM_ERR_ARG, "null argument"};
        check_inputs(mesh, params);
        if (params->quad_policy != CQBEM_QUAD_BANDED) throw Err{CQBEM_ERR_ARG, "quad_policy is not CQBEM_QUAD_BANDED"};
        const Prepared P = prepare(mesh, params);
        *npatches = P.pair_npatches;
        const int PV = pair_patch_size();
        if (patch_of_vertex)
            for (int pc = 0; pc < P.pair_npatches; pc++)
                for (int sl = 0; sl < PV; sl++) {
                    const int v = P.pair_patch_verts[(size_t)pc * PV + sl];
                    if (v >= 0) patch_of_vertex[v] = pc;
                }
        if (far)
            for (int i = 0; i < P.nv; i++)
                for (int pc = 0; pc < P.pair_npatches; pc++)
                    far[(size_t)i * P.pair_npatches + pc] = tile_is_far(P, pc, mesh->V + 3 * i) ? 1 : 0;
    } catch (const Err &e) {
        return e.code;
    } catch (const std::exception &) {
        return CQBEM_ERR_ARG;
    }
    return CQBEM_OK;
}

int cqbem_band_plan_faces(const cqbem_mesh *mesh, const cqbem_params *params, uint8_t *far) {
    try {
        if (!mesh || !params || !far) throw Err{CQBEM_ERR_ARG, "null argument"};
        check_inputs(mesh, params);
        if (params->quad_policy != CQBEM_QUAD_BANDED) throw Err{CQBEM_ERR_ARG, "quad_policy is not CQBEM_QUAD_BANDED"};
        const Prepared P = prepare(mesh, params);
        for (int i = 0; i < P.nv; i++)
            for (int f = 0; f < P.nf; f++) far[(size_t)i * P.nf + f] = face_is_far(P, f, mesh->V + 3 * i) ? 1 : 0;
    } catch (const Err &e) {
        return e.code;
    } catch (const std::exception &) {
        return CQBEM_ERR_ARG;
    }
    return CQBEM_OK;
}

int cqbem_job_footprint(const cqbem_mesh *mesh, const cqbem_params *params, int outputs_on_device, uint64_t *bytes) {
    try {
        check_inputs(mesh, params);
        if (!bytes) throw Err{CQBEM_ERR_ARG, "null bytes"};
        *bytes = footprint(mesh, params, outputs_on_device, (size_t)8 << 30);
    } catch (const Err &e) {
        return e.code;
    }
    return CQBEM_OK;
}

int cqbem_assemble(cqbem_context *ctx, const cqbem_mesh *mesh, const cqbem_params *params, cqbem_outputs *out) {
    if (!ctx) return CQBEM_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    const auto t0 = std::chrono::steady_clock::now();
    double tm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    try {
        check_inputs(mesh, params);
        if (!out) throw Err{CQBEM_ERR_ARG, "null outputs"};
        HIPCHK(hipSetDevice(ctx->device));
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        const uint64_t avail = std::min<uint64_t>(ctx->budget, (uint64_t)fr + ctx->pool.cached_bytes());
        // everything but the Laplace-domain scratch, then give the scratch what is left (capped)
        const uint64_t fixed = footprint(mesh, params, out->on_device, 0);
        if (fixed > avail) throw Err{CQBEM_ERR_NOMEM, "job does not fit the HBM budget"};
        const size_t slab = (size_t)std::min<uint64_t>(avail - fixed, (uint64_t)48 << 30);
        run_job(ctx, ctx->stream, mesh, params, out, slab, tm, true);
    } catch (const Err &e) {
        return fail(ctx, e);
    } catch (const std::exception &e) {
        return fail(ctx, Err{CQBEM_ERR_HIP, e.what()});
    }
    tm[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::copy(tm, tm + 8, ctx->timings);
    return CQBEM_OK;
}

int cqbem_assemble_scene(cqbem_context *ctx, int32_t n, const cqbem_mesh *meshes, const cqbem_params *params,
                         const cqbem_outputs *outs) {
    if (!ctx || n < 0 || (n > 0 && (!meshes || !params || !outs))) return CQBEM_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    const auto t0 = std::chrono::steady_clock::now();
    double tm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    try {
        HIPCHK(hipSetDevice(ctx->device));
        struct TimersOff {  // no event waits inside the captured graphs
            bool prev = phase_timers_on();
            TimersOff() { phase_timers_on() = false; }
            ~TimersOff() { phase_timers_on() = prev; }
        } timers_off;
        run_scene(ctx, ctx->stream, n, meshes, params, outs, tm);
    } catch (const Err &e) {
        return fail(ctx, e);
    } catch (const std::exception &e) {
        return fail(ctx, Err{CQBEM_ERR_HIP, e.what()});
    }
    tm[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::copy(tm, tm + 8, ctx->timings);
    return CQBEM_OK;
}

int cqbem_assemble_noalloc(cqbem_context *ctx, const cqbem_mesh *mesh, const cqbem_params *params,
                           const cqbem_outputs *out) {
    if (!ctx) return CQBEM_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    const auto t0 = std::chrono::steady_clock::now();
    double tm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    try {
        check_inputs(mesh, params);
        if (!out || !out->on_device) throw Err{CQBEM_ERR_ARG, "assemble_noalloc: outputs must be device buffers"};
        void *base = ctx->device < kMaxDevices ? noalloc_scratch_ptr[ctx->device] : nullptr;
        if (!base) throw Err{CQBEM_ERR_UNSUPPORTED, "assemble_noalloc: no scratch on this device"};
        // every device buffer of the job but the caller's outputs, with one row slab of spectra
        const uint64_t fixed = footprint(mesh, params, 1, 0);
        if (fixed > kNoAllocScratchBytes)
            throw Err{CQBEM_ERR_NOMEM, "assemble_noalloc: the job needs " + std::to_string(fixed >> 20) +
                                           " MiB of scratch, more than the " + std::to_string(CQBEM_NOALLOC_SCRATCH_MIB) +
                                           " MiB maximum"};
        const size_t slab = (size_t)(kNoAllocScratchBytes - fixed);
        HIPCHK(hipSetDevice(ctx->device));
        std::lock_guard<std::mutex> dev_lk(noalloc_call_mu[ctx->device]);
        struct TimersOff {
            bool prev = phase_timers_on();
            TimersOff() { phase_timers_on() = false; }
            ~TimersOff() { phase_timers_on() = prev; }
        } timers_off;
        JobArena region(base, kNoAllocScratchBytes, ctx->stream);
        run_job(ctx, ctx->stream, mesh, params, out, slab, tm, true, &region);
    } catch (const Err &e) {
        return fail(ctx, e);
    } catch (const std::exception &e) {
        return fail(ctx, Err{CQBEM_ERR_HIP, e.what()});
    }
    tm[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::copy(tm, tm + 8, ctx->timings);
    return CQBEM_OK;
}

int cqbem_last_timings(const cqbem_context *ctx, double *out, int32_t n) {
    if (!ctx || !out || n < 0) return CQBEM_ERR_ARG;
    const int k = std::min(n, 8);
    std::copy(ctx->timings, ctx->timings + k, out);
    return k;
}

int cqbem_laplace(cqbem_context *ctx, const cqbem_mesh *mesh, const cqbem_params *params, int32_t l, double *U, double *Pm,
                  double *Bt, double *Ba) {
    if (!ctx) return CQBEM_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        check_inputs(mesh, params);
        if (l < 0 || l > params->L / 2) throw Err{CQBEM_ERR_ARG, "frequency index out of range"};
        HIPCHK(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        Prepared P = prepare(mesh, params);
        DevMesh d;
        upload_mesh(d, mesh, P, st);
        const int nv = P.nv;
        const long long n3 = 3LL * nv;
        const bool wG = U && (params->which & CQBEM_G), wH = Pm && (params->which & CQBEM_H);
        const bool wB = (Bt || Ba) && (params->which & (CQBEM_B_TRANS | CQBEM_B_ANGULAR));
        DBuf dU, dP, dBt, dBa;
        if (wG) dU.alloc((size_t)n3 * n3 * sizeof(cd));
        if (wH) dP.alloc((size_t)n3 * n3 * sizeof(cd));
        dBt.alloc((size_t)n3 * 3 * sizeof(cd)), dBa.alloc((size_t)n3 * 9 * sizeof(cd));
        const double4 *quad = d.quad.as<double4>();
        const double2 *line = d.line.as<double2>();
        const Freq *freqs = d.freqs.as<Freq>();
        const int npts = P.npts, nline = (int)P.line.size();
        const bool mixed = params->precision == CQBEM_MIXED;
        DBuf dfreq32;
        if (mixed) upload(dfreq32, freqs_f32(P.freqs), st);
        const MatF mat32 = mat_f32(P.mat);
        if (!params->galerkin && mixed && mixed_float_g()) {  // far-field G and B in float, H in double (as cqbem_assemble)
            if (wH) HIPCHK(launch_gh(d, P, freqs, l, 1, 0, nv, nullptr, dP.as<cd>(), n3, n3 * n3, st));
            if (wG) HIPCHK(launch_colloc_g32(d.g, quad, npts, dfreq32.as<FreqF>(), l, 1, mat32, 0, nv, dU.as<cd>(), n3, n3 * n3, st));
            if (wG) HIPCHK(launch_colloc_singular(0, d.g, line, nline, freqs, l, 1, P.mat, 0, nv, dU.as<cd>(), n3, n3 * n3, nullptr, nullptr, st));
            if (wH) HIPCHK(launch_colloc_singular(2, d.g, line, nline, freqs, l, 1, P.mat, 0, nv, dP.as<cd>(), n3, n3 * n3, nullptr, nullptr, st));
            if (wB) {
                HIPCHK(launch_colloc_b32(d.g, quad, npts, dfreq32.as<FreqF>(), l, 1, mat32, 0, nv, dBt.as<cd>(), dBa.as<cd>(), st));
                HIPCHK(launch_colloc_singular(1, d.g, line, nline, freqs, l, 1, P.mat, 0, nv, nullptr, n3, 0, dBt.as<cd>(),
                                              dBa.as<cd>(), st));
            }
        } else if (!params->galerkin) {
            if (wG || wH) {
                HIPCHK(launch_gh(d, P, freqs, l, 1, 0, nv, dU.as<cd>(), dP.as<cd>(), n3, n3 * n3, st));
                if (wG) HIPCHK(launch_colloc_singular(0, d.g, line, nline, freqs, l, 1, P.mat, 0, nv, dU.as<cd>(), n3, n3 * n3, nullptr, nullptr, st));
                if (wH) HIPCHK(launch_colloc_singular(2, d.g, line, nline, freqs, l, 1, P.mat, 0, nv, dP.as<cd>(), n3, n3 * n3, nullptr, nullptr, st));
            }
            if (wB) {
                if (mixed)
                    HIPCHK(launch_colloc_b32(d.g, quad, npts, dfreq32.as<FreqF>(), l, 1, mat32, 0, nv, dBt.as<cd>(), dBa.as<cd>(), st,
                                             fbands_of(d, P)));
                else
                    HIPCHK(launch_colloc_b(d.g, quad, npts, freqs, l, 1, P.mat, 0, nv, dBt.as<cd>(), dBa.as<cd>(), st,
                                           fbands_of(d, P)));
                HIPCHK(launch_colloc_singular(1, d.g, line, nline, freqs, l, 1, P.mat, 0, nv, nullptr, n3, 0, dBt.as<cd>(),
                                              dBa.as<cd>(), st));
            }
        } else {
            GalerkinJob gj{d.g, quad, npts, line, nline, freqs + l, 1, P.mat, ctx->num_cus, d.pair_pts.as<double4>(),
                           (int)P.quad.size(), d.pair_patch_pass.as<int>(), d.pair_pass_info.as<int2>(),
                           d.pair_pass_faces.as<int4>(), d.pair_patch_verts.as<int>(), P.pair_npatches, (size_t)1 << 30};
            if (mixed) gj.freqs32 = dfreq32.as<FreqF>() + l, gj.mat32 = mat32;
            double tm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            run_galerkin(gj, mesh->F, mesh->vertex_map_inverse, wG ? dU.as<cd>() : nullptr, wH ? dP.as<cd>() : nullptr,
                         wB ? dBt.as<cd>() : nullptr, wB ? dBa.as<cd>() : nullptr, tm, st);
        }
        if (wG) HIPCHK(hipMemcpyAsync(U, dU.p, dU.bytes, hipMemcpyDeviceToHost, st));
        if (wH) HIPCHK(hipMemcpyAsync(Pm, dP.p, dP.bytes, hipMemcpyDeviceToHost, st));
        if (wB && Bt) HIPCHK(hipMemcpyAsync(Bt, dBt.p, dBt.bytes, hipMemcpyDeviceToHost, st));
        if (wB && Ba) HIPCHK(hipMemcpyAsync(Ba, dBa.p, dBa.bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    } catch (const Err &e) {
        return fail(ctx, e);
    } catch (const std::exception &e) {
        return fail(ctx, Err{CQBEM_ERR_HIP, e.what()});
    }
    return CQBEM_OK;
}

// ---- serving ----
// Stream slots per context (concurrent jobs): CQBEM_SLOTS, 1..16, default 4.
static int slot_count() {
    const char *e = std::getenv("CQBEM_SLOTS");
    const int v = e ? std::atoi(e) : 4;
    return std::min(16, std::max(1, v));
}

// Device bytes of the outputs a job writes (G and H at their storage size).
static uint64_t output_bytes(const cqbem_mesh *mesh, const cqbem_params *p, const cqbem_outputs *out) {
    const uint64_t n3 = 3 * (uint64_t)mesh->nv, L = (uint64_t)p->L;
    const uint64_t esz = p->storage == CQBEM_STORE_F32 ? 4 : 8;
    uint64_t b = 0;
    if ((p->which & CQBEM_G) && out->G) b += L * n3 * n3 * esz;
    if ((p->which & CQBEM_H) && out->H) b += L * n3 * n3 * esz;
    if (out->B_trans) b += L * n3 * 3 * 8;
    if (out->B_angular) b += L * n3 * 9 * 8;
    if (out->B) b += L * n3 * 6 * 8;
    return b;
}

static int submit_impl(cqbem_context *ctx, const cqbem_mesh *mesh, const cqbem_params *params, const cqbem_outputs *out,
                       uint64_t *job_id, bool resident) {
    if (!ctx || !job_id) return CQBEM_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        check_inputs(mesh, params);
        if (!out || !out->on_device) throw Err{CQBEM_ERR_ARG, "serving needs device-resident outputs"};
        HIPCHK(hipSetDevice(ctx->device));
        if (ctx->slots.empty()) {
            const int ns = slot_count();
            ctx->slots.resize(ns), ctx->slot_job.assign(ns, 0);
            for (auto &s : ctx->slots) HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        }
        const int nslots = (int)ctx->slots.size();
        // reap finished jobs' slots
        for (int k = 0; k < nslots; k++) {
            const uint64_t jid = ctx->slot_job[k];
            if (jid && (!ctx->jobs.count(jid) || ctx->jobs[jid]->done.load())) ctx->slot_job[k] = 0;
        }
        int slot = -1;
        for (int k = 0; k < nslots && slot < 0; k++)
            if (!ctx->slot_job[k]) slot = k;
        if (slot < 0) {
            ctx->last_error = "no free stream slot";
            return CQBEM_ERR_BUSY;
        }
        // admission by footprint: the job's scratch (and, when resident, its outputs) next to what is reserved
        const uint64_t res = resident ? output_bytes(mesh, params, out) : 0;
        const uint64_t fixed = footprint(mesh, params, 1, 0);  // scratch with one-row slabs
        if (fixed + res > ctx->budget)
            throw Err{CQBEM_ERR_NOMEM, "job footprint (" + std::to_string((fixed + res) >> 20) +
                                           " MiB) exceeds the HBM budget"};
        // the spectra slab: half of what the budget leaves (all of it up to 256 MiB), at most 2 GiB (a slab of that
        // size already holds ~30k tiles of the pair kernel, and smaller slabs leave room for the other slots' jobs);
        // one-row slabs when even that does not fit
        const uint64_t used = ctx->reserved + fixed + res;
        const uint64_t left = ctx->budget > used ? ctx->budget - used : 0;
        size_t slab = (size_t)std::min<uint64_t>(std::max<uint64_t>(left / 2, std::min<uint64_t>(left, (uint64_t)256 << 20)),
                                                 (uint64_t)2 << 30);
        uint64_t scratch = footprint(mesh, params, 1, slab);
        if (ctx->reserved + scratch + res > ctx->budget) slab = 0, scratch = fixed;
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        if (ctx->reserved + scratch + res > ctx->budget || (uint64_t)fr + ctx->pool.cached_bytes() < scratch) {
            ctx->last_error = "HBM budget or device memory held by running and resident jobs";
            return CQBEM_ERR_BUSY;
        }
        const uint64_t bytes = scratch;
        auto job = std::make_shared<JobRecord>();
        job->id = ctx->next_job++;
        job->V.assign(mesh->V, mesh->V + 3 * mesh->nv);
        job->F.assign(mesh->F, mesh->F + 3 * mesh->nf);
        if (mesh->vertex_map_inverse) job->vmap.assign(mesh->vertex_map_inverse, mesh->vertex_map_inverse + mesh->nv);
        job->mesh = cqbem_mesh{job->V.data(), mesh->nv, job->F.data(), mesh->nf,
                               job->vmap.empty() ? nullptr : job->vmap.data()};
        job->params = *params;
        job->out = *out;
        job->bytes = bytes;
        job->resident_bytes = res;
        job->resident = resident;
        job->slot = slot;
        ctx->reserved += bytes + res;
        ctx->slot_job[slot] = job->id;
        ctx->jobs[job->id] = job;
        hipStream_t st = ctx->slots[slot];
        JobRecord *jr = job.get();
        const int dev = ctx->device;
        job->worker = std::thread([ctx, jr, st, dev, slab]() {
            double tm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            try {
                HIPCHK(hipSetDevice(dev));
                run_job(ctx, st, &jr->mesh, &jr->params, &jr->out, slab, tm, false, nullptr, &jr->cancel);
            } catch (const Err &e) {
                jr->status = e.code, jr->error = e.msg;
            } catch (const std::exception &e) {
                jr->status = CQBEM_ERR_HIP, jr->error = e.what();
            }
            {
                std::lock_guard<std::mutex> lk2(ctx->mu);
                ctx->reserved -= jr->bytes;
            }
            jr->done.store(1);
        });
        *job_id = job->id;
    } catch (const Err &e) {
        return fail(ctx, e);
    }
    return CQBEM_OK;
}

int cqbem_submit(cqbem_context *ctx, const cqbem_mesh *mesh, const cqbem_params *params, const cqbem_outputs *out,
                 uint64_t *job_id) {
    return submit_impl(ctx, mesh, params, out, job_id, false);
}

int cqbem_submit_resident(cqbem_context *ctx, const cqbem_mesh *mesh, const cqbem_params *params,
                          const cqbem_outputs *out, uint64_t *job_id) {
    return submit_impl(ctx, mesh, params, out, job_id, true);
}

int cqbem_release(cqbem_context *ctx, uint64_t job_id) {
    if (!ctx) return CQBEM_ERR_ARG;
    std::thread worker;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->jobs.find(job_id);
        if (it == ctx->jobs.end() || !it->second->resident) return CQBEM_ERR_NOTFOUND;
        if (!it->second->done.load()) {
            ctx->last_error = "release: the job is still running";
            return CQBEM_ERR_BUSY;
        }
        worker = std::move(it->second->worker);
        ctx->reserved -= it->second->resident_bytes;
        for (int k = 0; k < (int)ctx->slot_job.size(); k++)
            if (ctx->slot_job[k] == job_id) ctx->slot_job[k] = 0;
        ctx->jobs.erase(it);
    }
    if (worker.joinable()) worker.join();  // finished: returns at once
    return CQBEM_OK;
}

int cqbem_cancel(cqbem_context *ctx, uint64_t job_id) {
    if (!ctx) return CQBEM_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->jobs.find(job_id);
    if (it == ctx->jobs.end()) return CQBEM_ERR_NOTFOUND;
    it->second->cancel.store(1);
    return CQBEM_OK;
}

int cqbem_poll(cqbem_context *ctx, uint64_t job_id, int32_t *done) {
    if (!ctx || !done) return CQBEM_ERR_ARG;
    std::shared_ptr<JobRecord> job;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->jobs.find(job_id);
        if (it == ctx->jobs.end()) return CQBEM_ERR_NOTFOUND;
        job = it->second;
    }
    *done = job->done.load();
    if (*done && job->status) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->last_error = job->error;
        return job->status;
    }
    return CQBEM_OK;
}

int cqbem_wait(cqbem_context *ctx, uint64_t job_id) {
    if (!ctx) return CQBEM_ERR_ARG;
    std::shared_ptr<JobRecord> job;
    std::thread worker;
    {
        // the first waiter takes the worker thread out of the record and joins it; later waiters on the same
        // job wait for its completion flag instead (a std::thread is joined once)
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->jobs.find(job_id);
        if (it == ctx->jobs.end()) return CQBEM_ERR_NOTFOUND;
        job = it->second;
        worker = std::move(job->worker);
    }
    if (worker.joinable())
        worker.join();
    else
        while (!job->done.load()) std::this_thread::sleep_for(std::chrono::microseconds(100));
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (int k = 0; k < (int)ctx->slot_job.size(); k++)
        if (ctx->slot_job[k] == job_id) ctx->slot_job[k] = 0;
    if (!job->resident) ctx->jobs.erase(job_id);  // a resident job's record lives until cqbem_release
    if (job->status) {
        ctx->last_error = job->error;
        return job->status;
    }
    return CQBEM_OK;
}

int cqbem_reserved_bytes(const cqbem_context *ctx, uint64_t *bytes) {
    if (!ctx || !bytes) return CQBEM_ERR_ARG;
    *bytes = ctx->reserved;
    return CQBEM_OK;
}

}  // extern "C"
