"""Builds superlinear_amd/libsus2amd.so for gfx950 with hipcc (in-tree)."""
import concurrent.futures as cf
import ctypes
import hashlib
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


# Model shapes compiled ahead into the shape-library cache (_jit/) besides the
# prebuilt instances: (L, K, radial basis size, scaling_map mode).  The GPU
# tests of runtime-compiled shapes use these, so no compiler runs on a GPU box.
EXTRA_SHAPES = [(2, 3, 8, 3)]


def sources():
    return sorted(list(SRC.glob("*.hip")) + list(SRC.glob("*.cpp")))


def source_hash():
    """Hash of every kernel source and header: the tag of shape libraries
    compiled against this build."""
    h = hashlib.sha1()
    for p in sorted(list(SRC.glob("*.hip")) + list(SRC.glob("*.hpp")) + list(SRC.glob("*.cpp")) +
                    [HERE.parent / "include" / "sus2amd.h"]):
        h.update(p.name.encode())
        h.update(p.read_bytes())
    return h.hexdigest()[:12]


def _compile(src, defs=()):
    obj = OBJ / (src.name + ".o")
    cmd = [HIPCC, *FLAGS, *defs, "-c", str(src), "-o", str(obj)]
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
    defs = (f'-DSUS2_SRC_HASH="{source_hash()}"', f'-DSUS2_OFFLOAD_ARCH="{ARCH}"')
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda p: _compile(p, defs), srcs))
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
    built = build_shapes(EXTRA_SHAPES)
    if verbose:
        print(f"built {OUT} and {CLI}" + "".join(f"\n  shape library {p}" for p in built))
    return OUT


def build_shapes(shapes):
    """Compile (once per source hash) the shape libraries of `shapes` through
    the library's own sus2_shape_build; returns their paths.  Stale shape
    libraries of other source hashes are removed."""
    lib = ctypes.CDLL(str(OUT))
    lib.sus2_shape_build.argtypes = [ctypes.c_int] * 4 + [ctypes.c_char_p, ctypes.c_int]
    lib.sus2_last_error.restype = ctypes.c_char_p
    tag = source_hash()
    jit = Path(os.environ.get("SUS2_JIT_DIR") or HERE / "_jit")
    if jit.is_dir():
        for p in jit.glob("sus2shape_*.so"):
            if not p.stem.endswith("_" + tag):
                p.unlink()
    out = []
    for L, K, B, SM in shapes:
        buf = ctypes.create_string_buffer(4096)
        if lib.sus2_shape_build(L, K, B, SM, buf, len(buf)) != 0:
            raise RuntimeError(f"shape library ({L}, {K}, {B}, {SM}): {lib.sus2_last_error().decode()}")
        out.append(buf.value.decode())
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
