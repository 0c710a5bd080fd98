"""Builds libamber_mi355x.so (HIP for gfx950 + host C) and the drop-in
``amber`` / ``amber_v2`` / ``amber_maps`` executables, in-tree, with hipcc
and gcc.

    python -m amber_amd.build            # library + drivers
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libamber_mi355x.so")
BIN = os.path.join(HERE, "bin")

HIP_SOURCES = ["amb_fft.hip", "amb_contract.hip", "amb_api.hip", "amb_wave.hip", "amb_post.hip", "amb_maps.hip",
               "amb_grid.hip", "amb_tune.hip", "amb_serve.hip", "amb_lagft.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed-FP32 instructions are kept out of every HIP object
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-value",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + INC, "-I" + CSRC]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    objs = []
    host_o = os.path.join(OBJ, "amb_host.o")
    _run(["gcc", "-O2", "-fPIC", "-fopenmp", "-std=gnu11", "-Wall", "-I" + INC, "-c",
          os.path.join(CSRC, "amb_host.c"), "-o", host_o])
    objs.append(host_o)
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        out = _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", o])
        if verbose and out.strip():
            print(out)
        objs.append(o)
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", LIB] + objs + ["-lgomp", "-lm"])
    for name, src, extra in (("amber", "amber_main.c", []), ("amber_v2", "amber_main.c", ["-DAMB_V2"]),
                             ("amber_maps", "amber_maps.c", [])):
        _run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I" + INC] + extra +
             [os.path.join(CSRC, src), "-o", os.path.join(BIN, name),
              "-L" + HERE, "-lamber_mi355x", "-Wl,-rpath,$ORIGIN/.."])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
