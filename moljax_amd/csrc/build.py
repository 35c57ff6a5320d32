"""Build libmoljax_hip.so (gfx950) in-tree with hipcc.

Every ``*.hip`` file in this directory is compiled to an object in parallel,
then linked into ``moljax_amd/libmoljax_hip.so``.  Always recompiles (the
build check must exercise the compiler); objects go to ``_build/`` next to
this file.
"""
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
OUT = PKG / "libmoljax_hip.so"
BUILD = HERE / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MJX_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code (the compiler
# would otherwise form v_pk_* from plain float code).
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{PKG.parent / 'include'}",
]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(HERE.glob("*.hip"))
    if not sources:
        raise RuntimeError("no HIP sources found")
    workers = min(len(sources), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} from {len(sources)} sources", file=sys.stderr)
    return OUT


if __name__ == "__main__":
    build()


def build_host_check(device: bool, verbose: bool = False) -> Path:
    """Build host_check (host_check.cpp against every *.hip source) with
    AddressSanitizer on host code only (each -fsanitize= behind
    -Xarch_host; GPU code is never built with a sanitizer).  device=False
    compiles host code only (--cuda-host-only: the argument checks run on a
    machine without a GPU, the kernels are absent); device=True also builds
    the gfx950 code so that the GPU workloads run.  Returns the executable
    under csrc/_asan/."""
    out_dir = HERE / "_asan"
    obj_dir = out_dir / ("obj_gpu" if device else "obj_cpu")
    obj_dir.mkdir(parents=True, exist_ok=True)
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
    # host-only objects carry an empty device binary (no kernels to launch)
    mode = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"] if device \
        else ["--cuda-host-only", "-Xclang", "-fcuda-include-gpubinary", "-Xclang", os.devnull]
    flags = [*mode, "-O1", "-g", "-std=c++17", "-fPIC", "-Wno-unused-function", *san,
             f"-I{PKG.parent / 'include'}"]
    sources = sorted(HERE.glob("*.hip"))

    def comp(src: Path) -> Path:
        obj = obj_dir / (src.stem + ".o")
        cmd = [HIPCC, *flags, "-c", str(src), "-o", str(obj)]
        if src.suffix == ".cpp":
            cmd.insert(1, "-xhip")
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
        return obj

    workers = min(len(sources) + 1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(comp, sources + [HERE / "host_check.cpp"]))
    exe = out_dir / ("host_check_gpu" if device else "host_check_cpu")
    link = [HIPCC, *( [f"--offload-arch={ARCH}"] if device else []), *san, "-o", str(exe), *map(str, objs)]
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {exe}", file=sys.stderr)
    return exe
