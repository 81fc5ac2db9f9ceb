"""Builds superlinear_amd/libsus2amd.so for gfx950 with hipcc (in-tree)."""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "csrc"
OUT = HERE / "libsus2amd.so"
CLI = HERE / "mlp-sus2"
ROCM_LIB = Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "lib"
OBJ = HERE / "_build"
ARCH = os.environ.get("SUS2_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-result"]


def sources():
    return sorted(list(SRC.glob("*.hip")) + list(SRC.glob("*.cpp")))


def _compile(src):
    obj = OBJ / (src.name + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-x", "c++", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=None, verbose=False):
    OBJ.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT.with_suffix(f".so.tmp{os.getpid()}")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    # the C++ calc-efs / calc-errors driver over the C ABI
    exe = CLI.with_name(CLI.name + f".tmp{os.getpid()}")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", str(SRC / "cli" / "mlp_sus2.cpp"), f"-L{HERE}", "-lsus2amd",
           "-Wl,-rpath,$ORIGIN", f"-L{ROCM_LIB}", f"-Wl,-rpath-link,{ROCM_LIB}", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(exe, CLI)
    if verbose:
        print(f"built {OUT} and {CLI}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
