"""Builds the package's native libraries in-tree.

    libgasal2_amd.so    HIP kernels (gfx950) + the host runtime of the C/C++ API
                        (kernels.hip: the alignment pipeline; band.hip: learned band;
                        noalloc.hip: the entry point that allocates nothing;
                        stream.cpp: FASTA/FASTQ reader and the multi-storage driver,
                        sam.cpp: CIGAR text and SAM records)
    libgasal2_synth.so  seeded synthetic pair generator (host C++/OpenMP)
    test_prog           the reference harness re-implemented on the API
    c_host              a C99 host of the API (the public headers' C view)

Run as ``python -m gasal2_amd.build``.  Every call recompiles.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GASAL2_ARCH", "gfx950")
# Packed FP32 VALU forms are kept out of generated code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-I", INC, "-I", CSRC]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_synth():
    out = os.path.join(HERE, "libgasal2_synth.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-o", out,
          os.path.join(CSRC, "synth.cpp")])
    return out


def build_lib():
    out = os.path.join(HERE, "libgasal2_amd.so")
    objs = []
    for src in ("kernels.hip", "band.hip", "noalloc.hip", "runtime.cpp", "stream.cpp", "sam.cpp"):
        obj = os.path.join(CSRC, "_build", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out] + objs)
    return out


def build_test_prog():
    src = os.path.join(CSRC, "test_prog.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(HERE, "test_prog")
    _run([HIPCC, "-O2", "-std=c++17", "-fopenmp", "-I", INC, src, "-o", out,
          "-L", HERE, "-lgasal2_amd", f"-Wl,-rpath,{HERE}", "-Wno-unused-command-line-argument"])
    return out


def build_c_host():
    """c_host: the API's C view compiled by a C compiler (gcc -std=c99) and
    linked against the library -- HIP's platform flags as `hipconfig
    --cpp_config` gives them."""
    out = os.path.join(HERE, "c_host")
    _run(["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
          "-I", "/opt/rocm/include", "-I", INC, os.path.join(CSRC, "c_host.c"), "-o", out,
          "-L", HERE, "-lgasal2_amd", f"-Wl,-rpath,{HERE}"])
    return out


def build_all():
    build_synth()
    build_lib()
    build_test_prog()
    build_c_host()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_lib, "synth": build_synth, "test_prog": build_test_prog,
         "c_host": build_c_host}[w]()
