"""Build libwake_hip.so for gfx950 in-tree.

Each source compiles to its own object under wake_amd/_build/ (in parallel;
an object is rebuilt when its source or any header in csrc/ or include/ is
newer), then everything links into wake_amd/libwake_hip.so.

    python wake_amd/build.py [--force] [-j N]
    WAKE_EXTRA_DEVICE_FLAGS=-DNAME=1 python wake_amd/build.py --variant V

A variant build (code-generation switches such as WAKE_MRT_MFMA, passed in
WAKE_EXTRA_DEVICE_FLAGS) goes to wake_amd/_build_V/ and
wake_amd/_variants/lib_V.so, and is loaded instead of the default library
when WAKE_HIP_LIB names it.

Run it as a script: `python -m wake_amd.build` would import the package,
which refuses to load until the library exists.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwake_hip.so")
ARCH = os.environ.get("WAKE_OFFLOAD_ARCH", "gfx950")

# Device code: gfx950 only, FP32 kept unpacked (no v_pk_*_f32).
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
] + os.environ.get("WAKE_EXTRA_DEVICE_FLAGS", "").split()
# Host code: the flag and link builders must round exactly as written
# (bit-exact cell types and BFL fractions), so no FP contraction.
HOST_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
              "-Wno-unused-command-line-argument"]

SOURCES = [
    "step_d3q19.hip", "step_d3q27.hip", "step_d2q9.hip", "step_kernels.hip", "thermal.hip", "analytics.hip",
    "ispc_compat.hip", "batch.hip", "domain.cpp", "capi.cpp", "lattice_grid.cpp", "runner.cpp", "batch_host.cpp",
    "host_api.cpp", "geometry.hip", "export.cpp", "thermal_solver.cpp", "hbm_probe.hip",
]
# The voxelizer's overlap test and the link builder's wall fractions must
# round every product and sum on its own, with IEEE division and sqrt.
EXTRA_FLAGS = {"geometry.hip": ["-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt"]}


def _headers_mtime():
    hs = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(ROOT, "include", "**", "*.h*"),
                                                               recursive=True)
    return max((os.path.getmtime(h) for h in hs), default=0.0)


def _compile(src, force, hdr_mtime, build_dir=BUILD):
    path = os.path.join(CSRC, src)
    obj = os.path.join(build_dir, src + ".o")
    if not force and os.path.exists(obj):
        om = os.path.getmtime(obj)
        if om >= os.path.getmtime(path) and om >= hdr_mtime:
            return obj, "up to date"
    flags = (DEVICE_FLAGS if src.endswith(".hip") else HOST_FLAGS) + EXTRA_FLAGS.get(src, [])
    cmd = ["hipcc", *flags, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, "built"


def build(force=False, jobs=None, verbose=True, variant=None):
    build_dir, lib = BUILD, LIB
    if variant:
        build_dir = os.path.join(HERE, "_build_" + variant)
        lib = os.path.join(HERE, "_variants", f"lib_{variant}.so")
        os.makedirs(os.path.dirname(lib), exist_ok=True)
    os.makedirs(build_dir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    hdr = _headers_mtime()
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 16)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, force, hdr, build_dir): s for s in srcs}
        for f in cf.as_completed(futs):
            obj, what = f.result()
            objs.append(obj)
            if verbose:
                print(f"[wake_amd.build] {futs[f]}: {what}", flush=True)
    objs.sort()
    newest = max(os.path.getmtime(o) for o in objs)
    if force or not os.path.exists(lib) or os.path.getmtime(lib) < newest:
        cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib, *objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose:
            print(f"[wake_amd.build] linked {lib}", flush=True)
    if not variant:
        build_api_check(force=force, verbose=verbose)
    return lib


API_CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "grid_api_check.cpp")
API_CHECK_BIN = os.path.join(ROOT, "tests", "cpp", "grid_api_check")


def build_api_check(force=False, verbose=True):
    """Compile the C++ API check program against the in-tree library."""
    if not os.path.exists(API_CHECK_SRC):
        return None
    if (not force and os.path.exists(API_CHECK_BIN)
            and os.path.getmtime(API_CHECK_BIN) >= max(os.path.getmtime(API_CHECK_SRC), os.path.getmtime(LIB))):
        return API_CHECK_BIN
    cmd = ["hipcc", "-O2", "-std=c++20", "-I", os.path.join(ROOT, "include"), API_CHECK_SRC, "-o", API_CHECK_BIN,
           "-L", HERE, "-lwake_hip", f"-Wl,-rpath,$ORIGIN/../../wake_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"api check build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"[wake_amd.build] built {API_CHECK_BIN}", flush=True)
    return API_CHECK_BIN


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--variant", default=None, help="build into _build_V/ and _variants/lib_V.so")
    a = ap.parse_args()
    try:
        build(force=a.force, jobs=a.j, variant=a.variant)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
