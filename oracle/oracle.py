"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

MODE_FP64, MODE_REFMIXED, MODE_LD = 0, 1, 2
MASK_G, MASK_H, MASK_BT, MASK_BA = 1, 2, 4, 8


class orc_params(C.Structure):
    _fields_ = [("c1", C.c_double), ("c2", C.c_double), ("rho", C.c_double), ("nu", C.c_double),
                ("dt", C.c_double), ("R", C.c_double), ("L", C.c_int), ("bdf", C.c_int),
                ("order", C.c_int), ("subdiv", C.c_int), ("galerkin", C.c_int), ("mode", C.c_int),
                ("nthreads", C.c_int)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        d = C.POINTER(C.c_double)
        i = C.POINTER(C.c_int)
        _lib.orc_polar_inv_r.restype = C.c_double
        _lib.orc_far_poly.restype = C.c_double
        _lib.orc_transform.argtypes = [d, C.c_longlong, C.c_int, C.c_double, d]
        _lib.orc_laplace.argtypes = [d, C.c_int, i, C.c_int, i, C.POINTER(orc_params), C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_int, d, d, d, d]
        _lib.orc_assemble.argtypes = [d, C.c_int, i, C.c_int, i, C.POINTER(orc_params), C.c_int,
                                      d, d, d, d, d, d, d, d, d]
        _lib.orc_kernel.argtypes = [C.c_int, C.c_int, d, d, d, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_double, d]
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def params(c1, c2, rho, nu, dt, R, L, bdf=2, order=5, subdiv=2, galerkin=False, mode=MODE_FP64, nthreads=0):
    return orc_params(c1, c2, rho, nu, dt, R, L, bdf, order, subdiv, int(galerkin), mode, nthreads)


def frequencies(L, R, dt, bdf):
    out = np.zeros(2 * (L // 2 + 1))
    lib().orc_frequencies(C.c_int(L), C.c_double(R), C.c_double(dt), C.c_int(bdf), _dp(out))
    return out[0::2] + 1j * out[1::2]


def tri_rule(order):
    b = np.zeros((7, 3)); w = np.zeros(7)
    n = lib().orc_rule_tri(C.c_int(order), _dp(b), _dp(w))
    return b[:n], w[:n]


def tri_rule_composite(order, sub):
    n_max = 7 * sub * sub
    b = np.zeros((n_max, 3)); w = np.zeros(n_max)
    n = lib().orc_rule_tri_composite(C.c_int(order), C.c_int(sub), _dp(b), _dp(w))
    return b[:n], w[:n]


def gauss_rule(order):
    x = np.zeros(5); w = np.zeros(5)
    n = lib().orc_rule_gauss(C.c_int(order), _dp(x), _dp(w))
    return x[:n], w[:n]


def kernel(which, x, y, n, c1, c2, rho, s, mode=MODE_FP64):
    """which: 0 u_hat (3x3), 1 p_hat (3x3), 2 b_hat (3x3), 3 b_hat_ijk (3x3x3)."""
    cnt = 27 if which == 3 else 9
    out = np.zeros(2 * cnt)
    x, y, n = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, n))
    lib().orc_kernel(which, mode, _dp(x), _dp(y), _dp(n), c1, c2, rho, s.real, s.imag, _dp(out))
    v = out[0::2] + 1j * out[1::2]
    return v.reshape(3, 3, 3) if which == 3 else v.reshape(3, 3)


def polar_inv_r(y1, y2, y3, order, sub):
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (y1, y2, y3)]
    return lib().orc_polar_inv_r(_dp(a[0]), _dp(a[1]), _dp(a[2]), C.c_int(order), C.c_int(sub))


def far_poly(y1, y2, y3, a, b, order, sub):
    v = [np.ascontiguousarray(t, dtype=np.float64) for t in (y1, y2, y3, a, b)]
    return lib().orc_far_poly(*[_dp(t) for t in v], C.c_int(order), C.c_int(sub))


def C_collocation(V, F, nu, vmap=None):
    V = np.ascontiguousarray(V, np.float64); F = np.ascontiguousarray(F, np.int32)
    out = np.zeros((3 * len(V), 3))
    if vmap is None:
        lib().orc_C_collocation(_dp(V), C.c_int(len(V)), _ip(F), C.c_int(len(F)), C.c_double(nu), _dp(out))
    else:
        vm = np.ascontiguousarray(vmap, np.int32)
        lib().orc_C_collocation_split(_dp(V), C.c_int(len(V)), _ip(F), C.c_int(len(F)), _ip(vm), C.c_double(nu),
                                      _dp(out))
    return out


def C_galerkin(V, F):
    V = np.ascontiguousarray(V, np.float64); F = np.ascontiguousarray(F, np.int32)
    out = np.zeros((3 * len(V), 3 * len(V)))
    lib().orc_C_galerkin(_dp(V), C.c_int(len(V)), _ip(F), C.c_int(len(F)), _dp(out))
    return out


def transform(spec, L, R):
    """spec: (L//2+1, M) complex -> (L, M) real time-domain weights."""
    spec = np.ascontiguousarray(spec, dtype=np.complex128)
    M = spec.shape[1]
    out = np.zeros((L, M))
    lib().orc_transform(spec.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), M, L, R, _dp(out))
    return out


def laplace(V, F, p, s, mask=15, rows=None, vmap=None):
    """Laplace-domain matrices at complex frequency s. Returns dict of complex arrays."""
    V = np.ascontiguousarray(V, np.float64); F = np.ascontiguousarray(F, np.int32)
    nv = len(V); n3 = 3 * nv
    U = np.zeros((n3, n3), np.complex128) if mask & MASK_G else None
    P = np.zeros((n3, n3), np.complex128) if mask & MASK_H else None
    Bt = np.zeros((n3, 3), np.complex128) if mask & MASK_BT else None
    Ba = np.zeros((n3, 3, 3), np.complex128) if mask & MASK_BA else None
    r0, r1 = (0, nv) if rows is None else rows
    vm = None if vmap is None else np.ascontiguousarray(vmap, np.int32)
    rc = lib().orc_laplace(_dp(V), nv, _ip(F), len(F), _ip(vm), C.byref(p), s.real, s.imag, mask, r0, r1,
                           *[None if a is None else a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
                             for a in (U, P, Bt, Ba)])
    if rc:
        raise RuntimeError(f"orc_laplace failed: {rc}")
    return {"U": U, "P": P, "Bt": Bt, "Ba": Ba}


def assemble(V, F, p, mask=15, b6=True, spectra=False, vmap=None):
    """Time-domain CQ weights: G, H (H_0 includes C), Bt, Ba (with the B_trans x correction), B6."""
    V = np.ascontiguousarray(V, np.float64); F = np.ascontiguousarray(F, np.int32)
    nv = len(V); n3 = 3 * nv; L = p.L; Lf = L // 2 + 1
    G = np.zeros((L, n3, n3)) if mask & MASK_G else None
    H = np.zeros((L, n3, n3)) if mask & MASK_H else None
    Bt = np.zeros((L, n3, 3)) if mask & MASK_BT else None
    Ba = np.zeros((L, n3, 3, 3)) if mask & MASK_BA else None
    B6 = np.zeros((L, n3, 6)) if b6 else None
    sp = {}
    if spectra:
        sp = {"U": np.zeros((Lf, n3, n3), np.complex128) if G is not None else None,
              "P": np.zeros((Lf, n3, n3), np.complex128) if H is not None else None,
              "Bt": np.zeros((Lf, n3, 3), np.complex128) if Bt is not None else None,
              "Ba": np.zeros((Lf, n3, 3, 3), np.complex128) if Ba is not None else None}
    vm = None if vmap is None else np.ascontiguousarray(vmap, np.int32)
    cp = lambda a: None if a is None else a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
    rc = lib().orc_assemble(_dp(V), nv, _ip(F), len(F), _ip(vm), C.byref(p), mask,
                            _dp(G), _dp(H), _dp(Bt), _dp(Ba), _dp(B6),
                            cp(sp.get("U")), cp(sp.get("P")), cp(sp.get("Bt")), cp(sp.get("Ba")))
    if rc:
        raise RuntimeError(f"orc_assemble failed: {rc}")
    out = {"G": G, "H": H, "Bt": Bt, "Ba": Ba, "B6": B6}
    if spectra:
        out["spec"] = sp
    return out
