"""Build libmofft.so for gfx950 in-tree.

Steps: run the kernel generator (mofft_compiler.emit) and the wisdom compiler
into csrc/generated, derive the 64-hex-digit build ID from the content of every
source that goes into the library (the role of the reference's content-hash
build ID, CMakeLists.txt:49-97), compile the C runtime with gcc and the HIP
sources with hipcc --offload-arch=gfx950, and link mofft_amd/lib/libmofft.so.

Usage: python -m mofft_amd.build [--force] [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
GEN = CSRC / "generated"
BUILD = PKG / "_build"
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libmofft.so"
WISDOM_JSON = PKG / "wisdom" / "gfx950.json"
PASSES_JSON = PKG / "wisdom" / "gfx950_passes.json"   # mofft_tune --pass-costs

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
CXX = os.environ.get("CXX", "g++")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wextra",
           "-Wno-unused-parameter"]


def _generate() -> None:
    sys.path.insert(0, str(PKG / "compiler"))
    from mofft_compiler import emit, wisdom  # noqa: E402
    emit.emit(GEN, passes_json=PASSES_JSON if PASSES_JSON.exists() else None)
    text = wisdom.emit_c(WISDOM_JSON)
    path = GEN / "mofft_wisdom.c"
    if not path.exists() or path.read_text() != text:
        path.write_text(text)


def _sources() -> tuple[list[Path], list[Path], list[Path]]:
    c_srcs = [CSRC / "mofft.c", CSRC / "mofft_batch.c", CSRC / "mofft_server.cpp",
              GEN / "mofft_registry.c",
              GEN / "mofft_wisdom.c",
              GEN / "mofft_costs.c",
              GEN / "mofft_launch_table.cpp"]
    hip_srcs = [CSRC / "hip" / "device.hip"]
    hip_srcs += sorted(GEN.glob("mofft_kernels_p*.hip"))
    hip_srcs += sorted(GEN.glob("mofft_batch_f*.hip"))
    headers = sorted(CSRC.glob("*.h")) + sorted((CSRC / "hip").glob("*.h")) + \
        sorted(GEN.glob("*.h")) + sorted((ROOT / "include").glob("*.h"))
    return c_srcs, hip_srcs, headers


def build_id() -> str:
    c_srcs, hip_srcs, headers = _sources()
    h = hashlib.sha256()
    extra = [WISDOM_JSON] + ([PASSES_JSON] if PASSES_JSON.exists() else [])
    for p in sorted(set(c_srcs + hip_srcs + headers + extra)):
        h.update(str(p.relative_to(ROOT)).encode())
        h.update(b"\0")
        h.update(p.read_bytes())
    h.update(" ".join(HIP_FLAGS + C_FLAGS).encode())
    return h.hexdigest()


def _compile(cmd: list[str], src: Path, obj: Path, deps_digest: str,
             force: bool) -> str:
    stamp = obj.with_suffix(obj.suffix + ".stamp")
    key = hashlib.sha256((" ".join(cmd) + deps_digest).encode() +
                         src.read_bytes()).hexdigest()
    if not force and obj.exists() and stamp.exists() and stamp.read_text() == key:
        return f"up to date {obj.name}"
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    stamp.write_text(key)
    return f"built {obj.name}"


def build(force: bool = False, jobs: int | None = None, verbose: bool = False,
          defines: tuple[str, ...] = (), out_dir: Path | None = None) -> Path:
    """Compile and link.  `defines` are extra -D options for the HIP sources
    (the kernel options documented in csrc/hip/fft_pass.h); `out_dir`
    redirects objects and the library (default: _build/ and lib/)."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    _generate()
    build_dir = Path(out_dir) / "obj" if out_dir else BUILD
    lib_dir = Path(out_dir) if out_dir else LIB_DIR
    lib = lib_dir / "libmofft.so"
    build_dir.mkdir(parents=True, exist_ok=True)
    lib_dir.mkdir(parents=True, exist_ok=True)
    c_srcs, hip_srcs, headers = _sources()
    hdr = hashlib.sha256()
    for p in headers:
        hdr.update(p.read_bytes())
    deps = hdr.hexdigest()
    # device code includes only the ABI header, the kernel headers and the
    # generated codelets: a change to the public or runtime headers does not
    # recompile the kernels
    hip_hdr = hashlib.sha256()
    for p in [CSRC / "mofft_abi.h"] + sorted((CSRC / "hip").glob("*.h")) + \
            sorted(GEN.glob("*.h")):
        hip_hdr.update(p.read_bytes())
    hip_deps = hip_hdr.hexdigest()
    bid = build_id()
    tasks = []
    objs = []
    for src in c_srcs:
        obj = build_dir / (src.stem + src.suffix + ".o")
        if src.suffix == ".cpp":
            cmd = [CXX, "-O2", "-fPIC", "-std=c++17", "-I", str(CSRC),
                   "-c", str(src), "-o", str(obj)]
        else:
            cmd = [CC, *C_FLAGS, f'-DMOFFT_BINARY_ID="{bid}"', "-I", str(CSRC),
                   "-c", str(src), "-o", str(obj)]
        tasks.append((cmd, src, obj))
        objs.append(obj)
    for src in hip_srcs:
        obj = build_dir / (src.stem + ".hip.o")
        cmd = [HIPCC, *HIP_FLAGS, *[f"-D{d}" for d in defines], "-I", str(GEN),
               "-I", str(CSRC / "hip"), "-c", str(src), "-o", str(obj)]
        tasks.append((cmd, src, obj))
        objs.append(obj)
    jobs = jobs or min(16, os.cpu_count() or 4)
    # longest first: the kernel files dominate the build
    tasks.sort(key=lambda t: -t[1].stat().st_size)
    with cf.ThreadPoolExecutor(max_workers=jobs) as pool:
        futs = [pool.submit(_compile, cmd, src, obj,
                            hip_deps if src.suffix == ".hip" else deps, force)
                for cmd, src, obj in tasks]
        for f in futs:
            msg = f.result()
            if verbose:
                print(msg)
    tmp = lib.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
           *[str(o) for o in objs]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, lib)
    (lib_dir / "BUILD_ID").write_text(bid + "\n")
    return lib


def build_c_tests(lib: Path = LIB) -> list[Path]:
    """Compile the drop-in C tests (tests/c/*.c) against include/ and the
    built library; binaries go to tests/c/build/."""
    out = []
    src_dir = ROOT / "tests" / "c"
    bin_dir = src_dir / "build"
    bin_dir.mkdir(parents=True, exist_ok=True)
    for src in sorted(src_dir.glob("*.c")):
        exe = bin_dir / src.stem
        cmd = [CC, "-O2", "-std=gnu11", "-Wall", "-I", str(ROOT / "include"),
               str(src), "-o", str(exe), "-L", str(lib.parent), "-lmofft",
               f"-Wl,-rpath,{lib.parent}", "-lm"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"C test build failed: {' '.join(cmd)}\n{res.stderr}")
        out.append(exe)
    return out


def build_tools(lib: Path = LIB) -> list[Path]:
    """Compile the autotuner (tune/mofft_tune.cpp) against the library."""
    out_dir = PKG / "tune" / "build"
    out_dir.mkdir(parents=True, exist_ok=True)
    exe = out_dir / "mofft_tune"
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wno-unused-command-line-argument",
           "-I", str(ROOT / "include"), str(PKG / "tune" / "mofft_tune.cpp"),
           "-o", str(exe), "-L", str(lib.parent), "-lmofft",
           f"-Wl,-rpath,{lib.parent}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"tool build failed: {' '.join(cmd)}\n{res.stderr}")
    return [exe]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--define", action="append", default=[],
                    help="extra NAME=VALUE for the HIP sources")
    ap.add_argument("--out-dir", type=Path, default=None)
    args = ap.parse_args()
    lib = build(force=args.force, jobs=args.jobs, verbose=True,
                defines=tuple(args.define), out_dir=args.out_dir)
    if args.out_dir is None:
        build_c_tests(lib)
        build_tools(lib)
    print(lib)


if __name__ == "__main__":
    main()
