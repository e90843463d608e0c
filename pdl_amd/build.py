"""Builds pdl_amd/libpdlhip.so for gfx950 with hipcc.

Kernel sources are compiled once per element type (the type switch of each
launcher lives in its own translation unit), in parallel, then linked with
the host planner into one shared library exporting the C ABI of
include/pdlhip.h.

    python -m pdl_amd.build            # or: python pdl_amd/build.py
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpdlhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# No FMA contraction (PDL's a*b+a rounds twice), wrapping signed ints (PDL
# is built with -fwrapv), no packed-FP32 instructions.
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fwrapv",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument", "-I" + CSRC]

NTYPES = 11


def units():
    u = []
    for t in range(NTYPES):
        u.append(("k_elementwise.hip", "ew_t%d" % t, ["-DPDLHIP_EW_TYPES_IN_THIS_TU=PDLHIP_EW_TYPE(%d)" % t]))
        u.append(("k_reduce.hip", "red_t%d" % t, ["-DPDLHIP_RED_TYPES_IN_THIS_TU=PDLHIP_RED_TYPE(%d)" % t]))
        u.append(("k_matmult.hip", "mm_t%d" % t, ["-DPDLHIP_MM_TYPES_IN_THIS_TU=PDLHIP_MM_TYPE(%d)" % t]))
        u.append(("k_mixed.hip", "mx_t%d" % t, ["-DPDLHIP_MIXED_TYPES_IN_THIS_TU=PDLHIP_MIXED_TYPE(%d)" % t]))
    u.append(("k_misc.hip", "misc", []))
    u.append(("k_badops.hip", "badops", []))
    u.append(("k_rng.hip", "rng", []))
    u.append(("k_index.hip", "index", []))
    u.append(("k_complex.hip", "complex", []))
    u.append(("k_mfma.hip", "mfma", []))
    u.append(("k_batch.hip", "batch", []))
    u.append(("k_ragged.hip", "ragged", []))
    u.append(("api.cpp", "api", ["-x", "hip"]))
    return u


def _fresh(out, src):
    """True when out is newer than its source and every shared header."""
    if not os.path.exists(out):
        return False
    t = os.path.getmtime(out)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    return all(os.path.getmtime(d) <= t for d in deps)


def compile_one(src, name, extra, force=True):
    out = os.path.join(OBJ, name + ".o")
    if not force and _fresh(out, src):
        return out
    cmd = [HIPCC] + FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s (%s):\n%s\n%s" % (src, name, " ".join(cmd), r.stderr[-6000:]))
    return out


def build(jobs=None, verbose=False, force=True):
    """Compiles every unit (force=True) or only the stale ones, then links."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(compile_one, *u, force=force) for u in units()]
        for f in futs:
            objs.append(f.result())
            if verbose:
                print("compiled", os.path.basename(objs[-1]), flush=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-6000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, force="--changed" not in sys.argv))
