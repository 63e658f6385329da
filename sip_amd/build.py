"""Builds libsip_fft.so (the C-ABI FFT library) for gfx950 with hipcc, in-tree.

Steps: generate the Stockham leaf-kernel instances (csrc/kernels/gen_instances.py) into
sip_amd/_gen/, compile every .hip / .cpp to an object in parallel, link sip_amd/libsip_fft.so.
Run `python -m sip_amd.build` (or call build()). Objects are cached by a hash of their inputs
unless force=True.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import hashlib
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(HERE, "_gen")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsip_fft.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
CLANGXX = os.path.join(ROCM, "llvm", "bin", "clang++")
ARCH = "gfx950"
NPARTS = 16

INCLUDES = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(CSRC, "kernels"), "-I" + GEN,
            "-I" + os.path.join(CSRC, "host")]
# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops", "-Wno-unused-command-line-argument"] + INCLUDES
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
              "-Wall", "-Wno-unused-parameter"] + INCLUDES


_INC_DIRS = [os.path.join(ROOT, "include"), os.path.join(CSRC, "kernels"), GEN, os.path.join(CSRC, "host")]


def _headers(src, seen=None):
    """Project headers reached from src through #include "..." lines (recursively), so that an
    object is rebuilt only when a header it really includes changes."""
    seen = set() if seen is None else seen
    here = os.path.dirname(src)
    with open(src, errors="replace") as fh:
        for line in fh:
            line = line.strip()
            if not line.startswith("#include") or '"' not in line:
                continue
            name = line.split('"')[1]
            for d in [here] + _INC_DIRS:
                cand = os.path.normpath(os.path.join(d, name))
                if os.path.exists(cand):
                    if cand not in seen:
                        seen.add(cand)
                        _headers(cand, seen)
                    break
    return sorted(seen)


def _digest(src, flags, headers):
    h = hashlib.sha256()
    for p in [src] + headers:
        with open(p, "rb") as fh:
            h.update(p.encode())
            h.update(fh.read())
    h.update(" ".join(flags).encode())
    return h.hexdigest()


def _compile(job):
    src, obj, cmd, dig, force = job
    stamp = obj + ".sha"
    if not force and os.path.exists(obj) and os.path.exists(stamp) and open(stamp).read() == dig:
        return src, None
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return src, r.stderr[-4000:]
    with open(stamp, "w") as fh:
        fh.write(dig)
    return src, None


def build(force: bool = True, jobs: int | None = None, verbose: bool = False, extra_device_flags=(),
          out: str | None = None, obj_dir: str | None = None) -> str:
    """Compiles and links the library. extra_device_flags / out / obj_dir build a variant (for
    tuning experiments) next to the default one."""
    out = out or LIB
    obj_root = obj_dir or OBJ
    os.makedirs(GEN, exist_ok=True)
    os.makedirs(obj_root, exist_ok=True)
    dflags = DEVICE_FLAGS + list(extra_device_flags)
    subprocess.run([sys.executable, os.path.join(CSRC, "kernels", "gen_instances.py"), GEN, str(NPARTS)],
                   check=True, capture_output=not verbose)
    work = []
    dev_srcs = sorted(glob.glob(os.path.join(CSRC, "kernels", "*.hip"))) + sorted(glob.glob(os.path.join(GEN, "*.hip")))
    host_srcs = sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp")))
    for src in dev_srcs:
        obj = os.path.join(obj_root, os.path.basename(src) + ".o")
        cmd = [HIPCC] + dflags + ["-c", src, "-o", obj]
        work.append((src, obj, cmd, _digest(src, dflags, _headers(src)), force))
    for src in host_srcs:
        obj = os.path.join(obj_root, os.path.basename(src) + ".o")
        cmd = [CLANGXX] + HOST_FLAGS + ["-c", src, "-o", obj]
        work.append((src, obj, cmd, _digest(src, HOST_FLAGS, _headers(src)), force))
    jobs = jobs or min(8, os.cpu_count() or 4)
    errors = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, err in ex.map(_compile, work):
            if err:
                errors.append((src, err))
    if errors:
        msg = "\n".join(f"--- {s}\n{e}" for s, e in errors)
        raise RuntimeError("libsip_fft build failed:\n" + msg)
    objs = [w[1] for w in work]
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    t = build(force="--force" in sys.argv, verbose=True)
    print(t)
