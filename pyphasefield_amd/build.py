"""Build libppf_hip.so (gfx950) in-tree with hipcc.

    python -m pyphasefield_amd.build        # or pyphasefield_amd.build.build()

Each translation unit is compiled separately (in parallel); the FP64 unit is
compiled with -ffp-contract=off so that FP64 runs keep the operation order of
the reference formulas.  Packed-FP32 instructions are disabled for all device
code.
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libppf_hip.so")
ARCH = os.environ.get("PPF_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

UNITS = {
    "ppf_abi.hip": [],
    # FP32 physics is checked at stated tolerances: hardware-rate 1-ulp
    # division and square root instead of the correctly rounded expansions
    "ppf_f32.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt"],
    "ppf_f64.hip": ["-ffp-contract=off"],
    "ppf_mlp.hip": [],
    "ppf_implicit.hip": [],
    # same arithmetic flags as the single-instance FP32 kernels it must match
    "ppf_batched.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt"],
}


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src, extra, objdir=OBJ):
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc()] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False, defines=(), lib=LIB, objdir=OBJ):
    """Compile every translation unit and link `lib`.  `defines`: extra
    -D options (tile and occupancy knobs of the kernels) for a variant build
    into another `lib` / `objdir`."""
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(os.path.dirname(os.path.abspath(lib)), exist_ok=True)
    missing = [k for k in UNITS if not os.path.exists(os.path.join(CSRC, k))]
    if missing:
        raise FileNotFoundError("missing HIP sources: %s" % ", ".join(missing))
    units = dict(UNITS)
    # any further translation unit in csrc/ builds with the common flags
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") and f not in units:
            units[f] = []
    dflags = ["-D" + d for d in defines]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(units))) as ex:
        objs = list(ex.map(lambda kv: _compile(kv[0], kv[1] + dflags, objdir), units.items()))
    tmp = lib + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
