"""Builds the gfx950 shared library libcqbem.so in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "libcqbem.so")
SOURCES = ["cqbem_kernels.hip", "cqbem_galerkin.hip", "cqbem_mixed.hip", "cqbem_host.hip", "cqbem_wavelet.hip", "cqbem_dense.cpp",
           "elastodynamics_cqbem_matrix_collection.cpp", "elastodynamics_cqbem_object.cpp", "mesh_split.cpp",
           "mesh_refine.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed-FP32 instructions stay disabled in all device code
DEVICE_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def build(verbose: bool = False) -> str:
    objs = []
    build_dir = os.path.join(PKG, "_build")
    os.makedirs(build_dir, exist_ok=True)
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(build_dir, src + ".o")
        objs.append(obj)
        cmd = [HIPCC, *COMMON, "-c", path, "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip", *DEVICE_FLAGS]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {src} ---\n{text}\n")
        elif verbose and text.strip():
            sys.stderr.write(text)
    if failed:
        raise RuntimeError("hipcc failed")
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", OUT + ".tmp", *objs, "-ldl"]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    # command-line driver of the C++ collection facade (links the library in-tree, rpath $ORIGIN)
    drv = os.path.join(PKG, "collection_driver")
    subprocess.run([HIPCC, *COMMON, os.path.join(CSRC, "collection_driver.cpp"), "-o", drv + ".tmp", "-L" + PKG,
                    "-lcqbem", "-Wl,-rpath,$ORIGIN"], check=True)
    os.replace(drv + ".tmp", drv)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
