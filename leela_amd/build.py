"""Build recipe of the HIP backend: compiles every source under csrc/ for
gfx950 with hipcc and links leela_amd/libleela_amd.so in-tree.

Run as ``python -m leela_amd.build`` (or through ``__graft_entry__.build``).
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
OBJ = ROOT / "_build" / "leela_amd"
LIB = PKG / "libleela_amd.so"

ARCH = "gfx950"
# -packed-fp32-ops: keep the compiler from forming packed FP32 VALU ops.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc",
                 shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


# Sources with no device code and no HIP runtime calls (the board engine,
# encoders, playouts and their C ABI): built by the host C++ compiler.
HOST_ONLY = {"go_board.cpp", "input_planes.cpp", "playout.cpp",
             "lz_host_abi.cpp", "lz_enc_record.cpp", "mc_policy.cpp"}
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-mpopcnt", "-Wall",
              f"-I{ROOT / 'include'}", f"-I{CSRC}"]


# Per-source flags: the pattern playouts' scores must round as the host's
# (no fused multiply-adds).
EXTRA_FLAGS = {"mc_playout.hip": ["-ffp-contract=off"]}


def host_cxx() -> str:
    for cand in (os.environ.get("CXX"), shutil.which("g++"),
                 "/opt/rocm/llvm/bin/clang++"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("no host C++ compiler")


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    obj = OBJ / (src.name + ".o")
    if src.name in HOST_ONLY:
        cmd = [host_cxx(), *HOST_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [hipcc(), *HIP_FLAGS, *EXTRA_FLAGS.get(src.name, []), "-c",
               str(src), "-o", str(obj)]
        if src.suffix == ".cpp":
            cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed on {src.name}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC",
           *map(str, objs), "-o", str(tmp), "-lpthread", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
