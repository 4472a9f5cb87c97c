"""Build step of the device library: compiles csrc/*.hip for gfx950 into libcipc_gpu.so.

Contact, broad-phase, CCD, energy and sparse translation units are compiled with
-ffp-contract=off so their expressions round as written (bit-exact pair sets and step
bounds); the elastic kernels keep fused multiply-adds.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcipc_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-lambda-capture",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
NO_CONTRACT = {"contact.hip", "broadphase.hip", "ccd.hip", "energy.hip", "sparse.hip", "kappa.hip", "solver.hip"}
SOURCES = ["context.hip", "elastic.hip", "cloth.hip", "contact.hip", "sparse.hip", "broadphase.hip", "ccd.hip", "energy.hip",
           "pcg.hip", "kappa.hip", "learned.hip", "solver.hip"]


def _compile(src):
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    flags = list(COMMON)
    if src in NO_CONTRACT:
        flags.append("-ffp-contract=off")
    cmd = [HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msg}")
    return obj, msg


HOST = os.path.join(HERE, "host")
HOST_SOURCES = ["mesh.cpp", "abi.cpp"]
HEADLESS = os.path.join(HERE, "cipc_headless")


def _compile_host(src):
    """Host-only C++ (scene set-up and its C ABI): g++, rounding as written."""
    obj = os.path.join(BUILD, "host_" + src.replace(".cpp", ".o"))
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off", "-c", os.path.join(HOST, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{r.stdout}{r.stderr}")
    return obj, r.stdout + r.stderr


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for obj, msg in results:
            if msg.strip():
                print(msg)
    results += [_compile_host(src) for src in HOST_SOURCES]
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + objs + ["-lrocsolver", "-lrocblas"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    # the headless driver (the reference's apps/cipc_headless), linked to the library
    tmp = HEADLESS + ".tmp"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-ffp-contract=off", os.path.join(HOST, "headless.cpp"), "-o", tmp,
                        "-L" + HERE, "-lcipc_gpu", "-Wl,-rpath,$ORIGIN"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("headless link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, HEADLESS)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
