"""Build the gfem libraries in-tree (fem-g_amd/gfem/).

  libgfem.so                HIP kernels + C ABI (include/gfem.h), gfx950
  libgfem_host.so           C++ host API (gmshfem::problem::Formulation<T>)
  _gfem_host*.so            Python binding of the host API

Run:  python fem-g_amd/build.py   (hipcc cross-compiles; no GPU needed)
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "gfem")
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GFEM_ARCH", "gfx950")
# Packed FP32 instructions stay disabled in every HIP compilation.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", INC, "-I", CSRC]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I", INC, "-I", CSRC,
             "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]


def run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    import pybind11
    py_inc = sysconfig.get_paths()["include"]

    dev_objs = [os.path.join(OBJ, "kernels.o"), os.path.join(OBJ, "gfem_api.o"), os.path.join(OBJ, "jit.o"),
                os.path.join(OBJ, "solver.o"), os.path.join(OBJ, "batch.o"), os.path.join(OBJ, "setup.o"),
                os.path.join(OBJ, "elemmm.o"), os.path.join(OBJ, "mg.o")]
    host_srcs = ["model.cpp", "msh.cpp", "gmshfem.cpp", "serve.cpp"]
    host_objs = [os.path.join(OBJ, s.replace(".cpp", ".o")) for s in host_srcs]
    py_obj = os.path.join(OBJ, "pybind.o")

    jobs = [
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "kernels.hip"), "-o", dev_objs[0]],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "gfem_api.cpp"), "-o", dev_objs[1]],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "jit.cpp"), "-o", dev_objs[2]],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "solver.hip"), "-o", dev_objs[3]],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "batch.hip"), "-o", dev_objs[4]],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "setup.hip"), "-o", dev_objs[5]],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "elemmm.hip"), "-o", dev_objs[6]],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "mg.hip"), "-o", dev_objs[7]],
    ]
    jobs += [["g++", *CXX_FLAGS, "-c", os.path.join(HOST, s), "-o", o] for s, o in zip(host_srcs, host_objs)]
    jobs.append(["g++", *CXX_FLAGS, "-fvisibility=hidden", "-I", pybind11.get_include(), "-I", py_inc,
                 "-c", os.path.join(HOST, "pybind.cpp"), "-o", py_obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(run, jobs):
            if verbose and out.strip():
                print(out)

    libgfem = os.path.join(OUT, "libgfem.so")
    run([HIPCC, *HIP_FLAGS, "-shared", "-o", libgfem, *dev_objs, "-L/opt/rocm/lib", "-lhiprtc",
         "-Wl,-rpath,/opt/rocm/lib"])
    libhost = os.path.join(OUT, "libgfem_host.so")
    run(["g++", "-shared", "-o", libhost, *host_objs, "-L", OUT, "-lgfem", "-Wl,-rpath,$ORIGIN"])
    run(["g++", "-shared", "-o", os.path.join(OUT, "_gfem_host" + ext), py_obj,
         "-L", OUT, "-lgfem_host", "-lgfem", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN",
         "-Wl,-rpath,/opt/rocm/lib"])


def build_oracle():
    """Compile the CPU oracle (test infrastructure) into oracle/build/."""
    run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


if __name__ == "__main__":
    build(verbose=True)
    build_oracle()
    print("built", sorted(os.listdir(OUT)))
