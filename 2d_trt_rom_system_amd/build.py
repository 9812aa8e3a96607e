"""Build the HIP library libtrt_hip.so in-tree for gfx950 (hipcc, no torch)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtrt_hip.so")
SOURCES = ["capi.hip", "state.hip", "sweep.hip", "sweep_pk.hip", "physics.hip", "loqd.hip", "qd_direct.hip", "qd_nd.hip", "serve.hip", "tune.hip", "restart.hip", "rom.hip", "noalloc.hip", "ncfile.hip", "outputs.hip", "romgrid.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics"]


def build(verbose: bool = False, jobs: int = 8) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(objdir, s.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((s, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        for s, t in failed:
            sys.stderr.write(f"--- {s} ---\n{t}\n")
        raise RuntimeError("hipcc failed for " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
                           "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    build_fortran(objdir, verbose)
    return LIB


FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
FSOURCES = ["trt_capi.f90", "trt_dropin.f90", "trt_mlqd_main.f90"]
FEXE = os.path.join(HERE, "trt_mlqd")


def build_fortran(objdir: str, verbose: bool = False) -> str:
    """The ISO_C_BINDING drop-in modules and the Fortran driver, linked against libtrt_hip.so."""
    fdir = os.path.join(HERE, "fortran")
    objs = []
    for s in FSOURCES:
        obj = os.path.join(objdir, s.replace(".f90", ".o"))
        cmd = [FC, "-O2", "-fPIC", "-module-dir", objdir, "-c", os.path.join(fdir, s), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = FEXE + ".tmp"
    subprocess.check_call([FC, "-o", tmp, *objs, "-L" + HERE, "-ltrt_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FEXE)
    return FEXE


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
