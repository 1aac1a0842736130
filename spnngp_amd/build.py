"""Build libnngp_hip.so (gfx950) in-tree with hipcc.

    python -m spnngp_amd.build              # spnngp_amd/libnngp_hip.so
    python -m spnngp_amd.build --sanitize   # spnngp_amd/_san/libnngp_hip.so

Each .hip file of spnngp_amd/csrc is compiled to an object in parallel,
then linked into spnngp_amd/libnngp_hip.so.  No GPU is needed to build.

--sanitize builds a second copy whose host code (only: the device code is
compiled as usual) carries AddressSanitizer and UndefinedBehaviorSanitizer,
for running the CPU tests against it (tests/run_sanitized.sh).  It needs
the clang ASan runtime preloaded, and it is never the library the package
loads by default.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libnngp_hip.so")
SAN_OBJ = os.path.join(HERE, "_obj_san")
SAN_LIB = os.path.join(HERE, "_san", "libnngp_hip.so")
ARCH = os.environ.get("NNGP_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are disabled (they have given wrong results on
# these machines); FP64 code is unaffected.
HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function",
]

# host-side sanitizers; each flag behind -Xarch_host so the device code is
# built exactly as in the plain library
SANFLAGS = [f for x in ("-fsanitize=address", "-fsanitize=undefined",
                        "-fno-sanitize-recover=undefined",
                        "-fno-omit-frame-pointer", "-g")
            for f in ("-Xarch_host", x)]


def _hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: install ROCm or put hipcc on PATH")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip"))


def _stale(src, obj):
    """An object is stale when it is missing or older than its source or
    any header of csrc/ or include/."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, f) for f in os.listdir(CSRC)
                    if f.endswith(".hpp")]
    deps.append(os.path.join(ROOT, "include", "nngp.h"))
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose: bool = True, jobs: int | None = None,
          incremental: bool = False, sanitize: bool = False) -> str:
    """Compile every source (incremental: only stale objects) and link."""
    hipcc = _hipcc()
    obj_dir, lib = (SAN_OBJ, SAN_LIB) if sanitize else (OBJ, LIB)
    extra = SANFLAGS if sanitize else []
    os.makedirs(obj_dir, exist_ok=True)
    os.makedirs(os.path.dirname(lib), exist_ok=True)
    srcs = sources()
    objs = [os.path.join(obj_dir, os.path.basename(s)[:-4] + ".o")
            for s in srcs]
    only = [a[len("--only="):] for a in sys.argv[1:]
            if a.startswith("--only=")] if __name__ == "__main__" else []
    todo = [(s, o) for s, o in zip(srcs, objs)
            if (not incremental or _stale(s, o)) and
            (not only or os.path.basename(s)[:-4] in only)]

    def compile_one(pair):
        src, obj = pair
        cmd = [hipcc, *HIPFLAGS, *extra, "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
        return src

    jobs = jobs or max(1, min(len(todo), os.cpu_count() or 4, 8))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src in ex.map(compile_one, todo):
            if verbose:
                print(f"[nngp build] compiled {os.path.basename(src)}",
                      flush=True)
    tmp = lib + ".tmp"
    link_san = (["-Xarch_host", "-fsanitize=address", "-Xarch_host",
                 "-fsanitize=undefined", "-Xarch_host", "-shared-libasan"]
                if sanitize else [])
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread",
           *link_san, "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"[nngp build] wrote {lib}", flush=True)
    return lib


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv[1:],
          sanitize="--sanitize" in sys.argv[1:])
    sys.exit(0)
