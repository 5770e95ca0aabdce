"""Build libvqc_hip.so (gfx950) in-tree from vqc.jl_amd/csrc/*.hip.

Each translation unit is compiled by hipcc in parallel, then linked into a
shared library next to this file: vqc.jl_amd/libvqc_hip.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvqc_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no packed FP32 instructions are generated (see the
# package README); FP64 code is unaffected.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    objs = []
    failed = False
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for src, obj, rc, out in ex.map(_compile, srcs):
            objs.append(obj)
            text = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
            if rc != 0:
                failed = True
                sys.stderr.write(f"[vqc build] FAILED {os.path.basename(src)}\n{out}\n")
            elif verbose and text.strip():
                sys.stderr.write(f"[vqc build] {os.path.basename(src)}:\n{text}\n")
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    # rocSOLVER / rocBLAS: the Hermitian eigensolver and complex product of the
    # density-matrix fidelity (vqc_fidelity.hip)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
