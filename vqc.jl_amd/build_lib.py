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


# Host-code sanitizer build (CPU tests only: tests/test_sanitizers.py). The
# translation units holding host logic -- the planner (vqc_circuit), the batch
# planner and work queue set-up (vqc_batch), the tuning tables (vqc_core) and the
# serving scheduler (vqc_serve) -- are compiled with AddressSanitizer and
# UndefinedBehaviorSanitizer on their host side; their device code is built at
# -O1 (it is never launched from this library: -O0 overflows the kernels' stack
# frames) and the other units are linked as in the regular build.
SAN_DIR = os.path.join(OBJ, "san")
SAN_LIB = os.path.join(SAN_DIR, "libvqc_hip_san.so")
SAN_UNITS = ("vqc_batch", "vqc_circuit", "vqc_core", "vqc_serve")
SAN_FLAGS = ["--offload-arch=gfx950", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Xarch_host", "-O1", "-Xarch_host", "-g", "-Xarch_host", "-fno-omit-frame-pointer",
             "-Xarch_host", "-fsanitize=address,undefined", "-Xarch_host", "-fno-sanitize-recover=undefined",
             "-Xarch_device", "-O1"]


# ThreadSanitizer build of the same units (host code only). The image has no
# clang TSan runtime; the process preloads gcc's libtsan, whose interface lacks
# the __tsan_mem* entry points recent clang emits for memory intrinsics, so those
# stay plain calls (which the runtime intercepts).
TSAN_DIR = os.path.join(OBJ, "tsan")
TSAN_LIB = os.path.join(TSAN_DIR, "libvqc_hip_tsan.so")
TSAN_FLAGS = ["--offload-arch=gfx950", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Xarch_host", "-O1", "-Xarch_host", "-g", "-Xarch_host", "-fno-omit-frame-pointer",
              "-Xarch_host", "-fsanitize=thread", "-mllvm", "-tsan-instrument-memintrinsics=false",
              "-Xarch_device", "-O1"]


def _compile_with(flags, outdir, src):
    obj = os.path.join(outdir, os.path.basename(src)[:-4] + ".o")
    r = subprocess.run([HIPCC] + flags + ["-c", src, "-o", obj], capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def _compile_units(flags, outdir, what):
    os.makedirs(outdir, exist_ok=True)
    srcs = [os.path.join(CSRC, u + ".hip") for u in SAN_UNITS]
    objs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        for src, obj, rc, out in ex.map(lambda s: _compile_with(flags, outdir, s), srcs):
            if rc != 0:
                raise RuntimeError(f"{what} build of {os.path.basename(src)} failed:\n{out}")
            objs.append(obj)
    return objs


def compile_san():
    """The sanitized objects (independent of build(), so both can run at once)."""
    return _compile_units(SAN_FLAGS, SAN_DIR, "sanitizer")


def compile_tsan():
    return _compile_units(TSAN_FLAGS, TSAN_DIR, "thread-sanitizer")


def link_san(san_objs):
    """libvqc_hip_san.so from the sanitized objects and the regular ones (after build())."""
    objs = [os.path.join(OBJ, os.path.basename(s)[:-4] + ".o") for s in sources() if os.path.basename(s)[:-4] not in SAN_UNITS]
    objs += san_objs
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-fsanitize=address,undefined", "-shared-libasan",
           "-o", SAN_LIB] + objs + ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("sanitizer link failed:\n" + r.stdout + r.stderr)
    return SAN_LIB


def link_tsan(tsan_objs):
    """libvqc_hip_tsan.so: the regular objects and the thread-sanitized ones, linked
    without a runtime (the process preloads it)."""
    objs = [os.path.join(OBJ, os.path.basename(s)[:-4] + ".o") for s in sources() if os.path.basename(s)[:-4] not in SAN_UNITS]
    objs += tsan_objs
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", TSAN_LIB] + objs + \
          ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("thread-sanitizer link failed:\n" + r.stdout + r.stderr)
    return TSAN_LIB


def build_all():
    """build() and the two sanitizer builds, compiled side by side."""
    with concurrent.futures.ThreadPoolExecutor(max_workers=2) as ex:
        san = ex.submit(compile_san)
        tsan = ex.submit(compile_tsan)
        lib = build()
        link_san(san.result())
        link_tsan(tsan.result())
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
