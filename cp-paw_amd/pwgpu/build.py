"""Build the pwgpu HIP library in-tree (libpwgpu.so next to this file) and the
Fortran host layer (cp-paw_amd/fortran: ISO_C_BINDING module, reference-style
wrappers and the pawgpu_cpstep driver) against it.

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions are
disabled (-packed-fp32-ops) for every HIP translation unit.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
LIB = os.path.join(HERE, "libpwgpu.so")

SOURCES = ["pw_plan.cpp", "pw_fft.hip", "pw_kernels.hip", "pw_ortho.hip", "pw_waves.cpp", "pw_serve.cpp",
           "pw_potential.hip", "pw_augment.hip"]

FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


FDIR = os.path.join(os.path.dirname(HERE), "fortran")
FSOURCES = ["pwgpu_c.f90", "pawgpu_restart.f90", "pawgpu.f90", "pawgpu_cpstep.f90"]
FDRIVER = os.path.join(FDIR, "pawgpu_cpstep")


def flang():
    for cand in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang", "/opt/rocm/llvm/bin/flang"):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran(verbose=True):
    """Compile the Fortran layer and link the driver against libpwgpu.so
    (rpath $ORIGIN/../pwgpu).  Returns the driver path, or None when no
    Fortran compiler is installed."""
    fc = flang()
    if fc is None:
        if verbose:
            sys.stderr.write("no Fortran compiler found; Fortran host layer not built\n")
        return None
    bdir = os.path.join(FDIR, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for src in FSOURCES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmd = [fc, "-O2", "-c", os.path.join(FDIR, src), "-o", obj, "-module-dir", bdir, "-I", bdir]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed on {src}:\n" + r.stdout.decode(errors="replace")[-6000:])
        objs.append(obj)
    tmp = FDRIVER + ".tmp"
    cmd = [fc, "-o", tmp] + objs + ["-L", HERE, "-lpwgpu", "-Wl,-rpath,$ORIGIN/../pwgpu"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("Fortran link failed:\n" + r.stdout.decode(errors="replace")[-6000:])
    os.replace(tmp, FDRIVER)
    return FDRIVER


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
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    print(build())
