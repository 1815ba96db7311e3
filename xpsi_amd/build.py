"""Builds libxpsi_hip.so (gfx950) in-tree with hipcc: one object per source,
compiled in parallel, then linked."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["xh_mesh.hip", "xh_rays.hip", "xh_integrate.hip", "xh_moment.hip", "xh_lanefold.hip", "xh_elsewhere.hip", "xh_signal.hip", "xh_capi.hip",
           "xh_serve.hip", "xh_noalloc.hip", "xh_prior.hip", "xh_atmplugin.hip", "xh_wavefold.hip",
           "xh_general.hip", "xh_stokes.hip"]
LIB = os.path.join(HERE, "libxpsi_hip.so")
OBJ = os.path.join(HERE, "_build")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         # packed FP32 instructions stay off (wrong results seen on this hardware)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False, jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    inc = "-I" + os.path.join(ROOT, "include")

    def compile_one(src):
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        cmd = [hipcc, *FLAGS, inc, "-c", os.path.join(HERE, "csrc", src), "-o", obj]
        return src, obj, subprocess.run(cmd, capture_output=True, text=True)

    jobs = jobs or min(len(SOURCES), max(1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, SOURCES))
    for src, _, res in results:
        if res.returncode != 0:
            sys.stderr.write(res.stderr)
            raise RuntimeError("hipcc failed on %s" % src)
        if verbose:
            sys.stderr.write(res.stderr)
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *[o for _, o, _ in results], "-o", LIB + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stderr)
        raise RuntimeError("hipcc failed linking libxpsi_hip.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build())
