"""Build libsoi_hip.so in-tree: HIP kernels (hipcc, gfx950) + C host layer (gcc).

    python -m soi_fft_amd.build          # or soi_fft_amd.build.build()
    python -m soi_fft_amd.build --incremental   # recompile only sources newer than their objects

Outputs soi_fft_amd/libsoi_hip.so; objects go to soi_fft_amd/_build/.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsoi_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("SOI_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["soi_kernels.hip", "soi_filter.hip", "soi_fft.hip"]
C_SOURCES = ["pfft.c", "compress_host.c", "harness.c"]
DRIVER = os.path.join(HERE, "soi_test")   # command-line driver (soi_test.c), links the library

# Packed FP32 (v_pk_*_f32) is disabled for all device code.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
CC_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function", "-fopenmp",
            "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "unknown target feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def _stale(src, obj):
    """True when obj is missing or older than src or any header under csrc/ and include/."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith((".h", ".hpp"))]
    inc = os.path.join(os.path.dirname(HERE), "include")
    deps += [os.path.join(inc, h) for h in os.listdir(inc) if h.endswith(".h")]
    return any(os.path.getmtime(dp) > t for dp in deps)


def build(verbose=False, incremental=False):
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    cc = shutil.which("gcc") or "cc"
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:  # the HIP translation units compile in parallel
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        if incremental and not _stale(path, obj):
            continue
        cmd = [hipcc] + HIPCC_FLAGS + ["-c", path, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
    for cmd, p in procs:
        out = p.communicate()[0]
        if p.returncode != 0:
            sys.stderr.write("\n".join(l for l in out.splitlines() if "packed-fp32" not in l) + "\n")
            raise RuntimeError("build step failed: " + " ".join(cmd))
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([cc] + CC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs +
         [f"-L{ROCM}/lib", "-lamdhip64", "-lgomp", "-lm"])
    os.replace(tmp, LIB)
    _run([cc, "-O2", "-std=gnu11", "-Wall", "-fopenmp", "-D__HIP_PLATFORM_AMD__",
          f"-I{ROCM}/include", os.path.join(CSRC, "soi_test.c"), "-o", DRIVER + ".tmp",
          f"-L{HERE}", "-lsoi_hip", "-Wl,-rpath,$ORIGIN", f"-L{ROCM}/lib", "-lrocfft",
          "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib", "-lm"])
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print("built", LIB, "and", DRIVER)
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv[1:])
