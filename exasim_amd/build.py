"""Build libexahip.so (C++ host code + HIP kernels for gfx950) in-tree with hipcc,
and the stand-alone driver exasim_amd/gput2cEXASIM that links it.

    python -m exasim_amd.build            # -> exasim_amd/libexahip.so, exasim_amd/gput2cEXASIM
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libexahip.so")
EXE = os.path.join(HERE, "gput2cEXASIM")
SOURCES = ["kernels.hip", "gemm.hip", "datain.cpp", "conn.cpp", "bilu0.cpp", "solver.cpp", "gmres.cpp", "batch.cpp", "api.cpp", "noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"


def flags():
    return ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            # MFMA accumulators in VGPRs: without it the loop-carried tiles are copied
            # between VGPRs and AGPRs around every K-step of the GEMM
            "-mllvm", "-amdgpu-mfma-vgpr-form",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
            "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def build(verbose=False, jobs=None):
    """Regenerate the models under models/, compile every source to an object file (in
    parallel) and link the shared library."""
    from . import modelgen
    modelgen.generate_all()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.rsplit(".", 1)[0] + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [HIPCC] + flags() + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        text = "\n".join(l for l in text.splitlines()
                         if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
        if p.returncode != 0:
            failed.append(src)
            print(text, file=sys.stderr)
        elif verbose and text.strip():
            print(text)
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-ldl"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # stand-alone driver with the command line of gput2cEXASIM (host code only)
    tmp = EXE + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
           os.path.join(CSRC, "main_t2c.cpp"), "-o", tmp, "-L" + HERE, "-lexahip",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, EXE)
    build_plugins(verbose)
    return LIB


PLUGINS = os.path.join(HERE, "plugins")


def build_plugins(verbose=False):
    """Build each sample model library under plugins/<name>/ (model functions in text2code's GPU
    form, compiled against include/CudaHip.hpp) into plugins/lib<name>.so."""
    out = []
    for name in sorted(os.listdir(PLUGINS)):
        src = os.path.join(PLUGINS, name)
        if not os.path.isdir(src):
            continue
        cpps = sorted(os.path.join(src, f) for f in os.listdir(src) if f.endswith(".cpp"))
        so = os.path.join(PLUGINS, "lib%s.so" % name)
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=" + ARCH,
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I" + os.path.join(ROOT, "include"), "-I" + src] + cpps + ["-o", so + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            print(r.stdout.decode(errors="replace"), file=sys.stderr)
            raise RuntimeError("hipcc failed for the model library " + name)
        os.replace(so + ".tmp", so)
        out.append(so)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
