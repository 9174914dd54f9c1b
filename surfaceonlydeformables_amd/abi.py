"""ctypes binding of include/cqbem.h (libcqbem.so, built in-tree for gfx950).

The product path: every call goes to the HIP library.  If the library is not
built, loading fails with an error instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libcqbem.so")

CQBEM_OK = 0
ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_UNSUPPORTED, ERR_BUSY, ERR_NOTFOUND, ERR_CANCELLED = -1, -2, -3, -4, -5, -6, -7
BDF1, BDF2 = 1, 2
FP64, MIXED = 0, 1
STORE_F64, STORE_F32 = 0, 1
QUAD_REFERENCE, QUAD_BANDED = 0, 1
SUBDIV_UPSAMPLE, SUBDIV_LOOP, SUBDIV_FALSE_BARYCENTRIC = 0, 1, 2
G, H, B_TRANS, B_ANGULAR, ALL = 1, 2, 4, 8, 15

EXPORTED = [
    "cqbem_create", "cqbem_destroy", "cqbem_last_error", "cqbem_assemble", "cqbem_job_footprint",
    "cqbem_laplace", "cqbem_frequencies", "cqbem_last_timings", "cqbem_submit", "cqbem_poll",
    "cqbem_wait", "cqbem_reserved_bytes", "cqbem_history", "cqbem_matvec", "cqbem_split_traction_discontinuity",
    "cqbem_wavelet_compress", "cqbem_csr_matvec", "cqbem_assemble_noalloc", "cqbem_submit_resident", "cqbem_release",
    "cqbem_history_f32", "cqbem_assemble_scene", "cqbem_band_plan", "cqbem_cancel",
    "cqbem_subdivide_mesh", "cqbem_decimate_mesh", "cqbem_band_plan_faces",
]


class cqbem_mesh(C.Structure):
    _fields_ = [("V", C.POINTER(C.c_double)), ("nv", C.c_int64), ("F", C.POINTER(C.c_int32)), ("nf", C.c_int64),
                ("vertex_map_inverse", C.POINTER(C.c_int32))]


class cqbem_params(C.Structure):
    _fields_ = [("c1", C.c_double), ("c2", C.c_double), ("rho", C.c_double), ("nu", C.c_double),
                ("dt", C.c_double), ("R", C.c_double), ("L", C.c_int32), ("bdf", C.c_int32),
                ("order", C.c_int32), ("subdivision", C.c_int32), ("galerkin", C.c_int32),
                ("precision", C.c_int32), ("which", C.c_int32), ("storage", C.c_int32),
                ("quad_policy", C.c_int32), ("far_order", C.c_int32), ("far_subdivision", C.c_int32),
                ("reserved", C.c_int32), ("band_kappa", C.c_double), ("band_kappa_b", C.c_double)]


class cqbem_outputs(C.Structure):
    _fields_ = [("G", C.c_void_p), ("H", C.c_void_p), ("B_trans", C.c_void_p), ("B_angular", C.c_void_p),
                ("B", C.c_void_p), ("on_device", C.c_int32)]


_lib = None


class CQBEMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cqbem error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m surfaceonlydeformables_amd.build` "
                               "(or __graft_entry__.build()) first")
        # One HIP runtime per process: torch ships its own libamdhip64 (same soname as ROCm's).  Loading
        # torch first makes libcqbem bind to that copy, so torch tensors and the library share a runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.cqbem_create.argtypes = [P(C.c_void_p), C.c_int, C.c_uint64]
        L.cqbem_destroy.argtypes = [C.c_void_p]
        L.cqbem_last_error.argtypes = [C.c_void_p]
        L.cqbem_last_error.restype = C.c_char_p
        L.cqbem_assemble.argtypes = [C.c_void_p, P(cqbem_mesh), P(cqbem_params), P(cqbem_outputs)]
        L.cqbem_assemble_noalloc.argtypes = [C.c_void_p, P(cqbem_mesh), P(cqbem_params), P(cqbem_outputs)]
        L.cqbem_assemble_scene.argtypes = [C.c_void_p, C.c_int32, P(cqbem_mesh), P(cqbem_params), P(cqbem_outputs)]
        L.cqbem_job_footprint.argtypes = [P(cqbem_mesh), P(cqbem_params), C.c_int, P(C.c_uint64)]
        L.cqbem_band_plan.argtypes = [P(cqbem_mesh), P(cqbem_params), C.c_void_p, C.c_void_p, P(C.c_int32)]
        L.cqbem_laplace.argtypes = [C.c_void_p, P(cqbem_mesh), P(cqbem_params), C.c_int32, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p]
        L.cqbem_frequencies.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_int32, P(C.c_double)]
        L.cqbem_last_timings.argtypes = [C.c_void_p, P(C.c_double), C.c_int32]
        L.cqbem_submit.argtypes = [C.c_void_p, P(cqbem_mesh), P(cqbem_params), P(cqbem_outputs), P(C.c_uint64)]
        L.cqbem_submit_resident.argtypes = [C.c_void_p, P(cqbem_mesh), P(cqbem_params), P(cqbem_outputs), P(C.c_uint64)]
        L.cqbem_release.argtypes = [C.c_void_p, C.c_uint64]
        L.cqbem_cancel.argtypes = [C.c_void_p, C.c_uint64]
        L.cqbem_history_f32.argtypes = [C.c_void_p, C.c_int64, C.c_int32] + [C.c_void_p] * 6 + [C.c_int32, C.c_void_p]
        L.cqbem_poll.argtypes = [C.c_void_p, C.c_uint64, P(C.c_int32)]
        L.cqbem_wait.argtypes = [C.c_void_p, C.c_uint64]
        L.cqbem_reserved_bytes.argtypes = [C.c_void_p, P(C.c_uint64)]
        L.cqbem_history.argtypes = [C.c_void_p, C.c_int64, C.c_int32] + [C.c_void_p] * 6 + [C.c_int32, C.c_void_p]
        L.cqbem_matvec.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cqbem_wavelet_compress.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                             C.c_void_p, P(C.c_int64)]
        L.cqbem_csr_matvec.argtypes = [C.c_void_p, C.c_int64] + [C.c_void_p] * 5
        L.cqbem_split_traction_discontinuity.argtypes = [P(C.c_double), C.c_int64, P(C.c_int32), C.c_int64, C.c_double,
                                                         P(C.c_double), P(C.c_int32), P(C.c_int32), P(C.c_int64)]
        L.cqbem_band_plan.argtypes = [P(cqbem_mesh), P(cqbem_params), C.c_void_p, C.c_void_p, P(C.c_int32)]
        L.cqbem_band_plan_faces.argtypes = [P(cqbem_mesh), P(cqbem_params), C.c_void_p]
        L.cqbem_subdivide_mesh.argtypes = [P(C.c_double), C.c_int64, P(C.c_int32), C.c_int64, C.c_int32, C.c_int32,
                                           P(C.c_double), P(C.c_int32), P(C.c_int64), P(C.c_int64)]
        L.cqbem_decimate_mesh.argtypes = [P(C.c_double), C.c_int64, P(C.c_int32), C.c_int64, C.c_int64,
                                          P(C.c_double), P(C.c_int32), P(C.c_int64), P(C.c_int64)]
        _lib = L
    return _lib


def make_mesh(V, F, vmap=None):
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int32)
    m = cqbem_mesh(V.ctypes.data_as(C.POINTER(C.c_double)), len(V), F.ctypes.data_as(C.POINTER(C.c_int32)), len(F),
                   None)
    keep = [V, F]
    if vmap is not None:
        vm = np.ascontiguousarray(vmap, dtype=np.int32)
        m.vertex_map_inverse = vm.ctypes.data_as(C.POINTER(C.c_int32))
        keep.append(vm)
    m._keep = keep
    return m


def make_params(c1, c2, rho, nu, dt, R, L, bdf=BDF2, order=5, subdivision=2, galerkin=False, precision=FP64,
                which=ALL, storage=STORE_F64, quad=QUAD_REFERENCE, far_order=0, far_subdivision=0, band_kappa=0.0,
                band_kappa_b=0.0):
    return cqbem_params(c1, c2, rho, nu, dt, R, L, bdf, order, subdivision, int(galerkin), precision, which, storage,
                        quad, far_order, far_subdivision, 0, band_kappa, band_kappa_b)


def split_traction_discontinuity(V, F, threshold_deg=20.0):
    """Vertex split along sharp edges (the reference's enable_traction_discontinuity mesh step):
    returns (V_split, F_split, vertex_map_inverse), ready for make_mesh(V_split, F_split, vmap)."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int32)
    pd, pi = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    n = C.c_int64(0)
    args = (V.ctypes.data_as(pd), len(V), F.ctypes.data_as(pi), len(F), threshold_deg)
    rc = lib().cqbem_split_traction_discontinuity(*args, None, None, None, C.byref(n))
    if rc != CQBEM_OK:
        raise CQBEMError(rc, "vertex split failed (the mesh must be a closed, consistently oriented manifold)")
    Vs, Fs, vm = np.zeros((n.value, 3)), np.zeros((len(F), 3), np.int32), np.zeros(n.value, np.int32)
    rc = lib().cqbem_split_traction_discontinuity(*args, Vs.ctypes.data_as(pd), Fs.ctypes.data_as(pi),
                                                  vm.ctypes.data_as(pi), C.byref(n))
    if rc != CQBEM_OK:
        raise CQBEMError(rc, "vertex split failed")
    return Vs, Fs, vm


def frequencies(L, R, dt, bdf):
    out = np.zeros(2 * (L // 2 + 1))
    n = lib().cqbem_frequencies(L, R, dt, bdf, out.ctypes.data_as(C.POINTER(C.c_double)))
    if n < 0:
        raise CQBEMError(n, "invalid frequency parameters")
    return out[0::2] + 1j * out[1::2]


class Context:
    """One GPU context (device, stream, HBM budget)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self._h = C.c_void_p()
        rc = lib().cqbem_create(C.byref(self._h), device, hbm_budget_bytes)
        if rc != CQBEM_OK:
            raise CQBEMError(rc, "cqbem_create failed (no HIP device?)")

    def close(self):
        if self._h:
            lib().cqbem_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != CQBEM_OK:
            raise CQBEMError(rc, lib().cqbem_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    TIMING_KEYS = ["total", "regular_gh", "regular_b", "singular", "transform", "host", "gh_launches",
                   "gh_evaluations"]

    def last_timings(self):
        t = np.zeros(8)
        n = lib().cqbem_last_timings(self._h, t.ctypes.data_as(C.POINTER(C.c_double)), 8)
        if n < 0:
            self._check(n)
        return dict(zip(self.TIMING_KEYS, t.tolist()))

    def assemble(self, mesh: cqbem_mesh, params: cqbem_params, want=("G", "H", "B_trans", "B_angular", "B")):
        """Host-output assembly: returns dict of numpy arrays (L, 3nv, 3nv) / (L, 3nv, 3) / (L, 3nv, 3, 3) / (L, 3nv, 6)."""
        nv, L = mesh.nv, params.L
        n3 = 3 * nv
        shapes = {"G": (L, n3, n3), "H": (L, n3, n3), "B_trans": (L, n3, 3), "B_angular": (L, n3, 3, 3),
                  "B": (L, n3, 6)}
        bits = {"G": G, "H": H, "B_trans": B_TRANS, "B_angular": B_ANGULAR, "B": B_TRANS | B_ANGULAR}
        res = {}
        out = cqbem_outputs()
        for k in want:
            if params.which & bits[k] == bits[k]:
                res[k] = np.zeros(shapes[k])
                setattr(out, k, res[k].ctypes.data)
        out.on_device = 0
        self._check(lib().cqbem_assemble(self._h, C.byref(mesh), C.byref(params), C.byref(out)))
        return res

    def assemble_device(self, mesh: cqbem_mesh, params: cqbem_params, ptrs: dict):
        """Device-output assembly into caller-owned device buffers {name: device pointer int}."""
        out = cqbem_outputs()
        for k, p in ptrs.items():
            setattr(out, k, p)
        out.on_device = 1
        self._check(lib().cqbem_assemble(self._h, C.byref(mesh), C.byref(params), C.byref(out)))

    def assemble_scene(self, meshes, params, ptrs):
        """All objects at once (cqbem_assemble_scene): meshes and params are lists, ptrs a list of {name: device
        pointer} dicts."""
        n = len(meshes)
        M_ = (cqbem_mesh * n)(*meshes)
        P_ = (cqbem_params * n)(*params)
        O_ = (cqbem_outputs * n)()
        for k, d in enumerate(ptrs):
            for name, p in d.items():
                setattr(O_[k], name, p)
            O_[k].on_device = 1
        self._check(lib().cqbem_assemble_scene(self._h, n, M_, P_, O_))

    def assemble_device_noalloc(self, mesh: cqbem_mesh, params: cqbem_params, ptrs: dict):
        """assemble_device through cqbem_assemble_noalloc: no device allocation, scratch from the library's fixed
        per-device array (CQBEM_NOALLOC_SCRATCH_MIB)."""
        out = cqbem_outputs()
        for k, p in ptrs.items():
            setattr(out, k, p)
        out.on_device = 1
        self._check(lib().cqbem_assemble_noalloc(self._h, C.byref(mesh), C.byref(params), C.byref(out)))

    def laplace(self, mesh: cqbem_mesh, params: cqbem_params, l: int):
        n3 = 3 * mesh.nv
        U = np.zeros((n3, n3), np.complex128) if params.which & G else None
        P = np.zeros((n3, n3), np.complex128) if params.which & H else None
        Bt = np.zeros((n3, 3), np.complex128) if params.which & (B_TRANS | B_ANGULAR) else None
        Ba = np.zeros((n3, 3, 3), np.complex128) if params.which & (B_TRANS | B_ANGULAR) else None
        ptr = lambda a: None if a is None else a.ctypes.data
        self._check(lib().cqbem_laplace(self._h, C.byref(mesh), C.byref(params), l, ptr(U), ptr(P), ptr(Bt), ptr(Ba)))
        return {"U": U, "P": P, "Bt": Bt, "Ba": Ba}

    # serving
    def submit(self, mesh, params, ptrs: dict, resident: bool = False) -> int:
        """Enqueues a job (-1 when it is not admitted now: retry after a running job finishes).  resident: its
        outputs stay reserved after it finishes, until release()."""
        out = cqbem_outputs()
        for k, p in ptrs.items():
            setattr(out, k, p)
        out.on_device = 1
        jid = C.c_uint64()
        fn = lib().cqbem_submit_resident if resident else lib().cqbem_submit
        rc = fn(self._h, C.byref(mesh), C.byref(params), C.byref(out), C.byref(jid))
        if rc == ERR_BUSY:
            return -1
        self._check(rc)
        return jid.value

    def cancel(self, job_id: int):
        """Asks a submitted job to stop at its next slab or phase boundary (cqbem_cancel)."""
        self._check(lib().cqbem_cancel(self._h, job_id))

    def release(self, job_id: int):
        self._check(lib().cqbem_release(self._h, job_id))

    def poll(self, job_id: int) -> bool:
        done = C.c_int32()
        self._check(lib().cqbem_poll(self._h, job_id, C.byref(done)))
        return bool(done.value)

    def wait(self, job_id: int):
        self._check(lib().cqbem_wait(self._h, job_id))

    def reserved_bytes(self) -> int:
        b = C.c_uint64()
        self._check(lib().cqbem_reserved_bytes(self._h, C.byref(b)))
        return b.value

    def history(self, n3: int, L: int, G, H, B, p_hist, u_hist, acc_hist, nhist: int, b):
        """History term of one CQ step on device buffers (device pointers as ints; G, H, B may be 0)."""
        self._check(lib().cqbem_history(self._h, n3, L, G or None, H or None, B or None, p_hist or None,
                                        u_hist or None, acc_hist or None, nhist, b))

    def history_f32(self, n3: int, L: int, G, H, B, p_hist, u_hist, acc_hist, nhist: int, b):
        """History term on G and H stored in float (CQBEM_STORE_F32)."""
        self._check(lib().cqbem_history_f32(self._h, n3, L, G or None, H or None, B or None, p_hist or None,
                                            u_hist or None, acc_hist or None, nhist, b))

    def matvec(self, n: int, A, x, y):
        """y = A x on device buffers (row-major n x n A; device pointers as ints)."""
        self._check(lib().cqbem_matvec(self._h, n, A, x, y))

    def wavelet_compress(self, n: int, A, num_elems: int, rowptr, col, val) -> int:
        """Haar compression of the device matrix A (n x n) into device CSR arrays; returns nnz."""
        nnz = C.c_int64()
        self._check(lib().cqbem_wavelet_compress(self._h, n, A, num_elems, rowptr, col or None, val or None,
                                                 C.byref(nnz)))
        return nnz.value

    def csr_matvec(self, rows: int, rowptr, col, val, x, y):
        """y = M x for a device CSR matrix."""
        self._check(lib().cqbem_csr_matvec(self._h, rows, rowptr, col or None, val or None, x, y))


def wavelet_padded(n: int) -> int:
    """The power of two >= n that a compressed n x n matrix is padded to."""
    np_ = 1
    while np_ < n:
        np_ <<= 1
    return np_


def band_plan(mesh, params):
    """CQBEM_QUAD_BANDED tile plan: (patch_of_vertex (nv,), far (nv, npatches) bool), the same at every frequency."""
    n = C.c_int32()
    rc = lib().cqbem_band_plan(C.byref(mesh), C.byref(params), None, None, C.byref(n))
    if rc != CQBEM_OK:
        raise CQBEMError(rc, "band plan refused")
    pov = np.zeros(mesh.nv, np.int32)
    far = np.zeros((mesh.nv, n.value), np.uint8)
    rc = lib().cqbem_band_plan(C.byref(mesh), C.byref(params), pov.ctypes.data, far.ctypes.data, C.byref(n))
    if rc != CQBEM_OK:
        raise CQBEMError(rc, "band plan refused")
    return pov, far.astype(bool)


def band_plan_faces(mesh, params):
    """CQBEM_QUAD_BANDED face plan of the far-field B: far (nv, nf) bool, the same at every frequency."""
    far = np.zeros((mesh.nv, mesh.nf), np.uint8)
    rc = lib().cqbem_band_plan_faces(C.byref(mesh), C.byref(params), far.ctypes.data)
    if rc != CQBEM_OK:
        raise CQBEMError(rc, "band plan refused")
    return far.astype(bool)


def job_footprint(mesh, params, outputs_on_device=True) -> int:
    b = C.c_uint64()
    rc = lib().cqbem_job_footprint(C.byref(mesh), C.byref(params), int(outputs_on_device), C.byref(b))
    if rc != CQBEM_OK:
        raise CQBEMError(rc, "invalid job")
    return b.value


def _resize_mesh(fn, what, V, F, *extra):
    """Calls a two-pass mesh entry point (sizes, then the mesh) and returns (V_out, F_out)."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.int32)
    pd, pi = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    nv, nf = C.c_int64(0), C.c_int64(0)
    args = (V.ctypes.data_as(pd), len(V), F.ctypes.data_as(pi), len(F), *extra)
    rc = fn(*args, None, None, C.byref(nv), C.byref(nf))
    if rc != CQBEM_OK:
        raise CQBEMError(rc, what + " failed")
    Vo, Fo = np.zeros((nv.value, 3)), np.zeros((nf.value, 3), np.int32)
    rc = fn(*args, Vo.ctypes.data_as(pd), Fo.ctypes.data_as(pi), C.byref(nv), C.byref(nf))
    if rc != CQBEM_OK:
        raise CQBEMError(rc, what + " failed")
    return Vo, Fo


def subdivide_mesh(V, F, scheme=SUBDIV_UPSAMPLE, levels=1):
    """`levels` rounds of subdivision (the reference's mesh.subdivision_level / subdivision_scheme step)."""
    return _resize_mesh(lib().cqbem_subdivide_mesh, "subdivision", V, F, scheme, levels)


def decimate_mesh(V, F, max_faces):
    """Shortest-edge decimation to at most max_faces faces (the reference's mesh.decimation_scale step, which
    passes max_faces = int(len(F) * decimation_scale))."""
    return _resize_mesh(lib().cqbem_decimate_mesh, "decimation", V, F, max_faces)
