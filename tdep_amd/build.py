"""Build libtc3.so (HIP for gfx950) in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtc3.so")
SOURCES = ["tc3_host.cpp", "tc3_api.cpp", "tc3_isotopes.cpp", "tc3_hdf5.cpp", "tc3_cumulative.cpp", "tc3_harmonic.hip", "tc3_scatter.hip",
           "tc3_kappa.hip", "tc3_mfma.hip", "tc3_fused.hip", "tc3_adjoint.hip", "tc3_fourphonon.hip",
           "tc3_jvp.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops keeps the compiler from emitting packed FP32 VALU instructions.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False, jobs=8):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        lang = ["-x", "hip"]
        cmd = [HIPCC] + FLAGS + lang + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-ldl"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
HOST = os.path.join(HERE, "tc_host")


def build_fortran_host(verbose=False):
    """The Fortran host program (tdep_amd/fortran/tc_host.f90) linked against libtc3.so;
    skipped with a message when no Fortran compiler is present."""
    if not os.path.exists(FC):
        sys.stderr.write(f"{FC} not found: the Fortran host is not built\n")
        return None
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    src = [os.path.join(HERE, "fortran", f) for f in ("tc3_api.f90", "tc_host.f90")]
    tmp = HOST + ".tmp"
    cmd = [FC, "-O2", "-module-dir", objdir, "-o", tmp] + src + ["-L" + HERE, "-ltc3", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode:
        raise RuntimeError("Fortran host build failed:\n" + r.stdout)
    os.replace(tmp, HOST)
    return HOST


def _wait(item):
    name, p = item
    out, _ = p.communicate()
    out = out.decode(errors="replace")
    if p.returncode:
        raise RuntimeError(f"hipcc failed on {name}:\n{out}")
    warn = [l for l in out.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if warn:
        sys.stderr.write(f"[{name}]\n" + "\n".join(warn) + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran_host(verbose="-v" in sys.argv))
