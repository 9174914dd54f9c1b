"""CPU restatement of the reference's mesh subdivision and decimation steps.  Test infrastructure: only tests
import it, as the checker of libcqbem.so's cqbem_subdivide_mesh / cqbem_decimate_mesh.

PhysicsBaseMatrixCollection::load_mesh (src/physics_base_matrix_collection.cpp:109-151) calls libigl v2.4.0
(cmake/libigl.cmake; libigl is not vendored in the reference): igl::upsample and igl::loop (lines 121-131, new
positions S * V), igl::false_barycentric_subdivision (lines 132-138) and igl::decimate with max_m =
(Eigen::Index)(F.rows() * decimation_scale) (lines 143-151).  Their published algorithms, restated with plain
Python loops for small meshes:

- upsample / loop: one new vertex per undirected edge, numbered after the old ones in the order in which the edge
  is first met scanning faces in order and each face's edges (v0,v1), (v1,v2), (v2,v0); face (v0,v1,v2) becomes
  (v0,m01,m20), (v1,m12,m01), (m01,m12,m20), (m12,v2,m20).  A new position is a row of the sparse product S V:
  its terms summed in ascending column order, starting from 0.0, each product rounded before the addition.
  Upsample rows: 1/2, 1/2.  Loop rows (Warren's weights): interior edge 3/8, 3/8 on its ends and 1/8, 1/8 on the
  two opposite vertices; boundary edge 1/2, 1/2; interior vertex of valence n: beta on each neighbour and
  1 - n beta on itself, beta = 3/16 (n = 3) or 3/8/n; boundary vertex 1/8 on each boundary neighbour, 3/4 on
  itself.
- false_barycentric_subdivision: barycentre ((a + b) + c) / 3 of face f appended as vertex nv + f; the face
  becomes (v0,v1,c), (v1,v2,c), (v2,v0,c).
- decimate: shortest edge first (cost = edge length, placement = midpoint 0.5 (a + b)); ties by edge number, edges
  numbered in ascending (min, max) order; a collapse of (s, d), s < d, keeps s at the placement, relabels d to s,
  drops the edge's two faces and, of each dropped face's other edges, keeps the one at s under its number; it is
  refused (edge re-queued at infinite cost) unless the closed one-rings of s and d share exactly four vertices and
  s, d are not two vertices of a lone tetrahedron; every edge of the faces around s and d is re-costed after a
  collapse.  The first collapse happens before the face count is tested; the loop ends when the count is at most
  max_faces or nothing collapsible is left.  Surviving faces and referenced vertices keep their order.

Parity: libigl is absent, so these restatements are checked (tests/test_mesh_refine.py) against what mathematics
fixes -- Loop's closed forms on the regular tetrahedron, octahedron and icosahedron (one per stencil branch), exact
edge midpoints, area and volume conservation and the 1/4 and 1/3 child areas of upsample and false_barycentric,
the shortest-first order, the link condition and the preserved topology of decimation.  The numbering of libigl's
own outputs and its tie-breaking among equal-length edges are parity unpinned."""
from __future__ import annotations

import heapq
import math

import numpy as np

UPSAMPLE, LOOP, FALSE_BARYCENTRIC = 0, 1, 2


def _ukey(a, b):
    return (a, b) if a < b else (b, a)


def _row(terms, V):
    """One row of S V: duplicate columns merged, columns ascending, 0.0 + w0 v0 + w1 v1 + ..."""
    merged = {}
    for c, w in terms:
        merged[c] = merged[c] + w if c in merged else w
    out = []
    for k in range(3):
        acc = 0.0
        for c in sorted(merged):
            acc = acc + merged[c] * float(V[c][k])
        out.append(acc)
    return out


def _split_four(V, F, loop):
    nv = len(V)
    order = {}      # undirected edge -> number, in order of first appearance
    sides = {}      # undirected edge -> [(face, j)]
    for f, tri in enumerate(F):
        for j in range(3):
            e = _ukey(int(tri[j]), int(tri[(j + 1) % 3]))
            order.setdefault(e, len(order))
            sides.setdefault(e, []).append((f, j))
    if any(len(s) > 2 for s in sides.values()):
        raise ValueError("mesh is not edge-manifold")
    W = [list(map(float, v)) for v in V]
    if loop:
        nbrs = {i: set() for i in range(nv)}
        bnbrs = {i: [] for i in range(nv)}
        for (a, b), s in sides.items():
            nbrs[a].add(b)
            nbrs[b].add(a)
            if len(s) == 1:
                bnbrs[a].append(b)
                bnbrs[b].append(a)
        for i in range(nv):
            if bnbrs[i]:
                if len(bnbrs[i]) != 2:
                    raise ValueError("boundary is not manifold")
                terms = [(bnbrs[i][0], 1 / 8), (bnbrs[i][1], 1 / 8), (i, 3 / 4)]
            elif nbrs[i]:
                n = len(nbrs[i])
                beta = 3 / 16 if n == 3 else 3 / 8 / n
                terms = [(x, beta) for x in nbrs[i]] + [(i, 1 - n * beta)]
            else:
                terms = [(i, 1.0)]
            W[i] = _row(terms, V)
    new = [None] * len(order)
    for (a, b), m in order.items():
        s = sides[(a, b)]
        if loop and len(s) == 2:
            opp = [int(F[f][(j + 2) % 3]) for f, j in s]
            terms = [(a, 3 / 8), (b, 3 / 8), (opp[0], 1 / 8), (opp[1], 1 / 8)]
        else:
            terms = [(a, 1 / 2), (b, 1 / 2)]
        new[m] = _row(terms, V)
    out_F = []
    for tri in F:
        v0, v1, v2 = (int(x) for x in tri)
        m01 = nv + order[_ukey(v0, v1)]
        m12 = nv + order[_ukey(v1, v2)]
        m20 = nv + order[_ukey(v2, v0)]
        out_F += [(v0, m01, m20), (v1, m12, m01), (m01, m12, m20), (m12, v2, m20)]
    return np.array(W + new, dtype=np.float64).reshape(-1, 3), np.array(out_F, dtype=np.int32)


def _false_barycentric(V, F):
    nv = len(V)
    bc, out_F = [], []
    for f, tri in enumerate(F):
        c = []
        for k in range(3):
            acc = 0.0
            for j in range(3):
                acc = acc + float(V[tri[j]][k])
            c.append(acc / 3.0)
        bc.append(c)
        v0, v1, v2 = (int(x) for x in tri)
        out_F += [(v0, v1, nv + f), (v1, v2, nv + f), (v2, v0, nv + f)]
    return np.vstack([np.asarray(V, dtype=np.float64), np.array(bc)]), np.array(out_F, dtype=np.int32)


def subdivide(V, F, scheme=UPSAMPLE, levels=1):
    V, F = np.asarray(V, dtype=np.float64), np.asarray(F, dtype=np.int32)
    for _ in range(levels):
        if scheme == FALSE_BARYCENTRIC:
            V, F = _false_barycentric(V, F)
        else:
            V, F = _split_four(V, F, scheme == LOOP)
    return V, F


def decimate(V, F, max_faces):
    V = [list(map(float, v)) for v in np.asarray(V, dtype=np.float64)]
    faces = [list(map(int, t)) for t in np.asarray(F)]
    vf = {i: set() for i in range(len(V))}
    for f, t in enumerate(faces):
        for x in t:
            vf[x].add(f)
    keys = sorted({_ukey(t[j], t[(j + 1) % 3]) for t in faces for j in range(3)})
    num = {k: e for e, k in enumerate(keys)}          # live edge -> number
    key_of = dict(enumerate(keys))                     # number -> live edge
    stamp = [0] * len(keys)
    place = [None] * len(keys)

    def cost(e):
        a, b = key_of[e]
        pa, pb = V[a], V[b]
        place[e] = [0.5 * (pa[k] + pb[k]) for k in range(3)]
        dx, dy, dz = pa[0] - pb[0], pa[1] - pb[1], pa[2] - pb[2]
        return math.sqrt(dx * dx + dy * dy + dz * dz)

    def ring(v):
        return {x for f in vf[v] for x in faces[f]}

    Q = [(cost(e), e, 0) for e in range(len(keys))]
    heapq.heapify(Q)
    m = len(faces)
    while True:
        e = -1
        while Q and Q[0][0] != math.inf:
            c, x, st = heapq.heappop(Q)
            if stamp[x] == st:
                e = x
                break
        if e < 0:
            break
        s, d = key_of[e]
        Ns, Nd = ring(s), ring(d)
        if len(Ns & Nd) != 4 or (len(Ns) == 4 and len(Nd) == 4 and len(Ns | Nd) == 4):
            stamp[e] += 1
            heapq.heappush(Q, (math.inf, e, stamp[e]))
            continue
        around = vf[s] | vf[d]
        for f in sorted(vf[s] & vf[d]):               # the two faces of the edge
            x = next(y for y in faces[f] if y != s and y != d)
            dead = num.pop(_ukey(d, x))
            del key_of[dead]
            stamp[dead] = -1
            for y in faces[f]:
                vf[y].discard(f)
            faces[f] = None
        for f in list(vf[d]):
            faces[f] = [s if y == d else y for y in faces[f]]
            vf[s].add(f)
        vf[d] = set()
        for k in [k for k in num if d in k]:          # edges (d, y) become (s, y) under their number
            n = num.pop(k)
            y = k[0] if k[1] == d else k[1]
            num[_ukey(s, y)] = n
            key_of[n] = _ukey(s, y)
        p = place[e]
        V[s], V[d] = list(p), list(p)
        del num[key_of[e]]
        del key_of[e]
        stamp[e] = -1
        touched = sorted({num[_ukey(t[j], t[(j + 1) % 3])] for f in around if faces[f] is not None
                          for t in [faces[f]] for j in range(3)})
        for x in touched:
            stamp[x] += 1
            heapq.heappush(Q, (cost(x), x, stamp[x]))
        m -= 2
        if m <= max_faces:
            break
    live = [t for t in faces if t is not None]
    used = sorted({x for t in live for x in t})
    remap = {v: i for i, v in enumerate(used)}
    return (np.array([V[v] for v in used], dtype=np.float64).reshape(-1, 3),
            np.array([[remap[x] for x in t] for t in live], dtype=np.int32).reshape(-1, 3))
