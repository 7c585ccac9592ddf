"""Build recipe for libportblas_amd.so (gfx950 only).

Every HIP/C++ source under ``portblas_amd/csrc`` is compiled by ``hipcc`` into
an object under ``portblas_amd/_build`` (in parallel), then linked into
``portblas_amd/libportblas_amd.so``.  The C++ drop-in test program under
``tests/cpp`` is linked against it when present.  Run ``python -m
portblas_amd.build`` or call :func:`build`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libportblas_amd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops is disabled on every device pass: packed FP32 VALU forms are
# not trusted on this hardware pool, so the compiler must not synthesise them.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> tuple[Path, str]:
    OBJ.mkdir(exist_ok=True)
    obj = OBJ / (src.stem + ("_hip" if src.suffix == ".hip" else "_cpp") + ".o")
    cmd = [HIPCC, *DEVICE_FLAGS, "-I", str(ROOT / "include"), "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd[1:1] = ["-x", "hip"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so.
    noise = "'-packed-fp32-ops' is not a recognized feature"
    msg = "\n".join(l for l in (proc.stdout + proc.stderr).splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{msg}\n$ {' '.join(cmd)}")
    return obj, msg


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile every source and link the shared library; returns its path."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    srcs = _sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    objs: list[Path] = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, msg in ex.map(_compile, srcs):
            objs.append(obj)
            if verbose:
                print(f"[build] {obj.name}", flush=True)
                if msg.strip():
                    print(msg, flush=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-Wl,-soname,libportblas_amd.so"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}{proc.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[build] linked {LIB}", flush=True)
    _build_cpp_tests(verbose)
    return LIB


def _build_cpp_tests(verbose: bool) -> None:
    """Link the C++ drop-in API test (tests/cpp/*.cpp) against the library."""
    tdir = ROOT / "tests" / "cpp"
    for src in sorted(tdir.glob("*.cpp")) if tdir.exists() else []:
        exe = tdir / src.stem
        cmd = [HIPCC, "-O2", "-std=c++17", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
               "-L", str(PKG), "-lportblas_amd", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../portblas_amd"]
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError(f"C++ test build failed for {src.name}:\n{proc.stdout}{proc.stderr}")
        if verbose:
            print(f"[build] {exe.relative_to(ROOT)}", flush=True)


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
