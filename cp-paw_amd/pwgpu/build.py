"""Build the pwgpu HIP library in-tree (libpwgpu.so next to this file).

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions are
disabled (-packed-fp32-ops) for every HIP translation unit.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
LIB = os.path.join(HERE, "libpwgpu.so")

SOURCES = ["pw_plan.cpp", "pw_fft.hip", "pw_kernels.hip", "pw_ortho.hip", "pw_waves.cpp"]

FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose=True, jobs=None):
    """Compile every source to an object in parallel, then link libpwgpu.so."""
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [hipcc(), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + \
              [f for f in FLAGS if f != "-shared"]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            sys.stderr.write(text[-4000:])
    if failed:
        msg = "\n".join(f"--- {s} ---\n{t[-6000:]}" for s, t in failed)
        raise RuntimeError("hipcc failed:\n" + msg)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace")[-6000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
