"""Build the in-tree render library psycles_amd/libpsx.so.

HIP kernels are compiled for gfx950 with hipcc; host C++ with g++ against the
HIP runtime headers.  Floating-point flags are part of the render contract:
no contraction, correctly rounded division and square root, no fast math,
and packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG.parent / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("PSX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = [
    "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    f"-I{ROCM / 'include'}", "-D__HIP_PLATFORM_AMD__", "-Wall",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(map(str, cmd)) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return r


# device headers that generated surface programs include (psx_rtc.cpp
# compiles them with hipRTC at run time)
EMBEDDED_HEADERS = ["psx_device.h", "psx_path.h", "psx_wavefront_kernels.h"]


def write_embedded_headers() -> Path:
    """Write _build/psx_embedded.cpp: the headers as string constants."""
    delim = "PSXSRC"
    parts = []
    for name in EMBEDDED_HEADERS:
        text = (CSRC / name).read_text()
        if f"){delim}\"" in text:
            raise RuntimeError(f"{name} contains the raw-string delimiter")
        parts.append(f'R"{delim}({text}){delim}"')
    out = BUILD / "psx_embedded.cpp"
    body = ("// written by psycles_amd/build.py from psycles_amd/csrc; do not edit\n"
            "namespace psx {\n"
            f"extern const unsigned kEmbeddedCount = {len(EMBEDDED_HEADERS)};\n"
            "extern const char *const kEmbeddedNames[] = {"
            + ", ".join(f'"{n}"' for n in EMBEDDED_HEADERS) + "};\n"
            "extern const char *const kEmbeddedSources[] = {\n" + ",\n".join(parts) + "};\n"
            "}  // namespace psx\n")
    out.write_text(body)
    return out


def build_library(verbose: bool = False) -> Path:
    """Compile every source and link libpsx.so; returns its path."""
    hipcc = shutil.which("hipcc") or str(ROCM / "bin" / "hipcc")
    BUILD.mkdir(exist_ok=True)
    jobs = []
    objects = []
    embedded = write_embedded_headers()
    jobs.append(["g++", *HOST_FLAGS, "-c", str(embedded), "-o", str(BUILD / "psx_embedded.o")])
    objects.append(BUILD / "psx_embedded.o")
    for src in sorted(CSRC.glob("*.hip")):
        obj = BUILD / (src.stem + ".hip.o")
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        objects.append(obj)
    for src in sorted(CSRC.glob("*.cpp")):
        obj = BUILD / (src.stem + ".o")
        jobs.append(["g++", *HOST_FLAGS, "-c", str(src), "-o", str(obj)])
        objects.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    out = PKG / "libpsx.so"
    tmp = PKG / "libpsx.so.tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objects), f"-L{ROCM / 'lib'}", "-lamdhip64", "-lhiprtc"])
    os.replace(tmp, out)
    return out


def build_contract_driver() -> Path:
    """Compile tests/contract_driver.cpp (a host program on the C++ contract
    mirror include/psx_contract.hpp) against libpsx.so into _build/."""
    BUILD.mkdir(exist_ok=True)
    out = BUILD / "contract_driver"
    src = PKG.parent / "tests" / "contract_driver.cpp"
    _run(["g++", "-std=c++20", "-O1", "-Wall", f"-I{ROCM / 'include'}", "-D__HIP_PLATFORM_AMD__", str(src),
          "-o", str(out), f"-L{PKG}", "-lpsx", "-Wl,-rpath,$ORIGIN/../psycles_amd", f"-Wl,-rpath,{ROCM / 'lib'}"])
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
