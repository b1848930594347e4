"""Build the in-tree HIP library ``madflow_amd/libmadflow_amd.so`` for gfx950.

Each translation unit is compiled with hipcc in parallel, then linked into one
shared library whose C ABI is declared in ``include/madflow_amd.h``.
``python -m madflow_amd.build [--only gg_ttx,gg_ttxg]`` builds from the command line.
"""

import argparse
import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmadflow_amd.so")

ARCH = os.environ.get("MFX_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # keep packed-FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-pass-failed",
]

UNITS = [
    "mfx_api.hip",
    "mfx_phasespace.hip",
    "mfx_integrand.hip",
    "mfx_flow.hip",
    "procs/gg_ttx.hip",
    "procs/gg_ttxg.hip",
    "procs/gg_ttxgg.hip",
    "procs/gg_ttxgg_fused.hip",
    "procs/gg_ttxggg.hip",
    "procs/uux_ttx.hip",
    "procs/uux_ttxg.hip",
    "procs/uux_ttxgg.hip",
    "procs/ug_ttxu.hip",
    "procs/uxg_ttxux.hip",
]


def _deps_hash(src):
    h = hashlib.sha1()
    for root, _, files in os.walk(CSRC):
        for f in sorted(files):
            if f.endswith((".h", ".inc")):
                with open(os.path.join(root, f), "rb") as fh:
                    h.update(fh.read())
    with open(src, "rb") as fh:
        h.update(fh.read())
    h.update(" ".join(CFLAGS).encode())
    return h.hexdigest()[:16]


def _compile(unit, force):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(OBJ, unit.replace("/", "_") + ".o")
    stamp = obj + ".hash"
    digest = _deps_hash(src)
    if not force and os.path.exists(obj) and os.path.exists(stamp) and open(stamp).read() == digest:
        return unit, obj, False, ""
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (unit, " ".join(cmd), r.stderr[-6000:]))
    with open(stamp, "w") as f:
        f.write(digest)
    return unit, obj, True, r.stderr


def build(force=False, jobs=None, verbose=True, exclude=()):
    """Compile every translation unit (cached by content hash) and link the library."""
    os.makedirs(OBJ, exist_ok=True)
    exclude = set(exclude) | set(filter(None, os.environ.get("MFX_BUILD_EXCLUDE", "").split(",")))
    units = [u for u in UNITS if os.path.exists(os.path.join(CSRC, u)) and u not in exclude]
    jobs = jobs or min(len(units), max(1, (os.cpu_count() or 4)))
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for unit, obj, rebuilt, log in ex.map(lambda u: _compile(u, force), units):
            objs.append(obj)
            if verbose:
                print("[madflow_amd.build] %s %s" % ("compiled" if rebuilt else "cached  ", unit), flush=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, LIB)
    if verbose:
        print("[madflow_amd.build] linked", LIB, flush=True)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    a = ap.parse_args(argv)
    build(force=a.force, jobs=a.jobs)


if __name__ == "__main__":
    main(sys.argv[1:])
