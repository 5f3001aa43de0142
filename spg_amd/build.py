"""Build the native libraries in-tree.

  spg_amd/lib/libspg_vbd.so   HIP kernels + C-ABI (include/spg_vbd.h, spg_vbd_learned.h, spg_xpbd.h, spg_implicit.h, spg_rigid.h), hipcc for gfx950
  spg_amd/lib/libspg_host.so  C++ mirror of spg's BaseSolver/ParticleGroup/Energy API (g++), over libspg_vbd.so

Run:  python -m spg_amd.build     (or spg_amd.build.build() from Python)
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIBDIR = os.path.join(PKG, "lib")
CSRC = os.path.join(PKG, "csrc")
CPP = os.path.join(PKG, "cpp")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions stay off: they are formed by the compiler from plain float code and
# some of their forms are unreliable on this hardware.
# FP32 division and square root use the hardware reciprocal / square root (about 1 ulp) instead of
# the correctly rounded sequences, and FP32 denormals flush to zero: the solver's FP32 magnitudes
# stay above 1e-30 and the FP64 paths are unaffected.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero",
             "-Wall", "-Wno-unused-result"]

VBD_SOURCES = ["vbd_kernels.hip", "vbd_context.hip", "mlp_kernels.hip", "mlp_context.hip", "warmstart_kernels.hip",
               "xpbd_kernels.hip", "xpbd_context.hip", "implicit_kernels.hip", "implicit_context.hip", "rigid_vbd.hip", "barrier_bench.hip", "coloring.cpp"]
HOST_SOURCES = ["particleGroup.cpp", "energies.cpp", "rigidBodyGroup.cpp", "meshIo.cpp", "flatten.cpp", "vbdGpu.cpp", "xpbdGpu.cpp",
                "implicitEulerGpu.cpp",
                "host_capi.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _atomic_link(cmd, out):
    tmp = out + ".tmp"
    _run(cmd + ["-o", tmp])
    os.replace(tmp, out)


def build_vbd():
    """Compiles every source to an object in parallel (each one is a separate hipcc process), then links."""
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in VBD_SOURCES]
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", src, "-o", obj])

    with ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 1, 8)) as pool:
        list(pool.map(compile_one, zip(srcs, objs)))
    out = os.path.join(LIBDIR, "libspg_vbd.so")
    _atomic_link([HIPCC, *HIP_FLAGS, "-shared", *objs], out)
    return out


def build_host():
    srcs = [os.path.join(CPP, "src", s) for s in HOST_SOURCES if os.path.exists(os.path.join(CPP, "src", s))]
    if not srcs:
        return None
    cxx = shutil.which("g++") or "g++"
    out = os.path.join(LIBDIR, "libspg_host.so")
    _atomic_link([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", f"-I{INCLUDE}",
                  f"-I{os.path.join(CPP, 'include')}", *srcs, f"-L{LIBDIR}", "-lspg_vbd",
                  f"-Wl,-rpath,$ORIGIN"], out)
    return out


def build_apps():
    """C++ example that drives VBDGpu exactly like a spg app drives spg::solver::VBD."""
    src = os.path.join(CPP, "apps", "clothVbdGpu.cpp")
    if not os.path.exists(src):
        return None
    cxx = shutil.which("g++") or "g++"
    out = os.path.join(LIBDIR, "cloth_vbd_gpu")
    _atomic_link([cxx, "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{os.path.join(CPP, 'include')}", src,
                  f"-L{LIBDIR}", "-lspg_host", "-lspg_vbd", f"-Wl,-rpath,{LIBDIR}"], out)
    return out


def build():
    build_vbd()
    build_host()
    build_apps()


if __name__ == "__main__":
    build()
    sys.exit(0)
