"""Build recipe for librerdmft_gpu.so (gfx950) and the test oracle.

`python -m rerdmft_amd.build` compiles in-tree: rerdmft_amd/librerdmft_gpu.so.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "librerdmft_gpu.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("rr_lib.hip", "rr_sigma.hip", "rr_transform.hip", "rr_rkb_chol.hip",
                                                       "rr_batch.hip", "rr_mlp.hip", "rr_diis.hip", "rr_cas.hip", "rr_gemm_wide.hip")]
# translation units whose matrix-core accumulators stay in the AGPR form (wide tiles)
AGPR_FORM = {"rr_gemm_wide.hip"}
HEADERS = [os.path.join(HERE, "csrc", f) for f in ("rr_device.hpp", "rr_gemm.hpp", "rr_gemm_rows.hpp", "rr_core.hpp", "rr_validate.hpp")] + [
    os.path.join(ROOT, "include", "rerdmft_gpu.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RR_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # packed fp32 VALU forms disabled on purpose (device code only; the host pass warns, harmlessly)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    # keep matrix-core accumulators in the VGPR form: with the AGPR form the allocator copied the
    # accumulators between the two register files on every k-stage of the fp64 GEMM loops
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
]
LIBS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def build_library(verbose=False):
    """Compiles each translation unit in parallel, then links the shared library."""
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(obj_dir, os.path.splitext(os.path.basename(src))[0] + ".o")
        flags = compile_flags
        if os.path.basename(src) in AGPR_FORM:
            i = compile_flags.index("-amdgpu-mfma-vgpr-form=1")
            flags = compile_flags[:i - 1] + compile_flags[i + 1:]  # drop the "-mllvm" pair
        cmd = [HIPCC] + flags + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), cmd))
        objs.append(obj)
    failed = [cmd for p, cmd in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}"] + objs + ["-o", LIB + ".tmp"] + LIBS
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_adapter_test(verbose=False):
    """C++ drop-in test of the adapter header (tests/cpp/test_adapter.cpp) linked to the library."""
    out_dir = os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "test_adapter")
    cmd = ["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cpp", "test_adapter.cpp"), "-o", exe,
           "-L" + HERE, "-lrerdmft_gpu", "-Wl,-rpath," + HERE, "-ldl"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return exe


VALIDATE_DIR = os.path.join(HERE, "_build", "validate")


def build_validation(verbose=False):
    """Validation build of the library (-DRR_VALIDATE: launch sites check pointer liveness and GEMM
    extents instead of launching) linked against the host-memory HIP stand-in of tests/validate, and
    the driver that replays the benchmark's call sequence through it. Test infrastructure: nothing
    here runs on or needs a GPU. Returns the path of the replay executable."""
    os.makedirs(VALIDATE_DIR, exist_ok=True)
    vdir = os.path.join(ROOT, "tests", "validate")
    mock = os.path.join(VALIDATE_DIR, "libhipmock.so")
    cmds = [["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
             os.path.join(vdir, "hip_mock.cpp"), "-o", mock]]
    flags = ["-O1", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-DRR_VALIDATE", "-Wno-unused-function",
             "-Wno-unused-variable", "-Wno-unused-parameter"]
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(VALIDATE_DIR, os.path.splitext(os.path.basename(src))[0] + ".o")
        cmd = [HIPCC] + flags + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stderr=subprocess.DEVNULL if not verbose else None), cmd))
        objs.append(obj)
    for cmd in cmds:
        subprocess.run(cmd, check=True)
    failed = [cmd for p, cmd in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    vlib = os.path.join(VALIDATE_DIR, "librerdmft_validate.so")
    subprocess.run(["g++", "-shared", "-fPIC"] + objs + ["-o", vlib, "-L" + VALIDATE_DIR, "-lhipmock",
                    "-Wl,-rpath," + VALIDATE_DIR], check=True)
    exe = os.path.join(VALIDATE_DIR, "replay")
    subprocess.run(["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    os.path.join(vdir, "replay.cpp"), "-o", exe, "-L" + VALIDATE_DIR, "-lrerdmft_validate",
                    "-lhipmock", "-Wl,-rpath," + VALIDATE_DIR], check=True)
    return exe


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def main():
    build_library(verbose=True)
    build_adapter_test(verbose=True)
    build_oracle()
    print("built", LIB)


if __name__ == "__main__":
    sys.exit(main())
