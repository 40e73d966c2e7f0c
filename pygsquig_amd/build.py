"""In-tree build of libgsqg.so for gfx950 (MI355X).

``python -m pygsquig_amd.build`` or ``pygsquig_amd.build.build()``.  The
kernel translation unit is split by grid size so the instantiations compile
in parallel; everything links into one shared library next to this file.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "libgsqg.so")
ARCH = os.environ.get("GSQG_OFFLOAD_ARCH", "gfx950")
SIZES = (16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192)

# -packed-fp32-ops: hipcc must not emit packed FP32 VALU ops (v_pk_*_f32).
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", os.path.join(ROOT, "include")]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False, jobs: int | None = None) -> str:
    hipcc = _hipcc()
    tmp = tempfile.mkdtemp(prefix="gsqg_build_")
    try:
        units = []
        for n in SIZES:
            units.append((os.path.join(CSRC, "gsqg_kernels.hip"), os.path.join(tmp, f"k{n}.o"),
                          [f"-DGSQG_INSTANTIATE_N={n}"]))
        units.append((os.path.join(CSRC, "gsqg_kernels.hip"), os.path.join(tmp, "kcommon.o"),
                      ["-DGSQG_INSTANTIATE_COMMON=1"]))
        units.append((os.path.join(CSRC, "gsqg_kernels.hip"), os.path.join(tmp, "kbatch.o"),
                      ["-DGSQG_INSTANTIATE_BATCH=1"]))
        units.append((os.path.join(CSRC, "gsqg_api.hip"), os.path.join(tmp, "api.o"), []))
        units.append((os.path.join(CSRC, "closure_kernels.hip"), os.path.join(tmp, "closure.o"), []))
        jobs = jobs or min(len(units), max(1, (os.cpu_count() or 4)), 16)

        def compile_unit(u):
            src, obj, extra = u
            _run([hipcc, *HIP_FLAGS, *extra, "-c", src, "-o", obj], verbose)
            return obj

        with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
            objs = list(ex.map(compile_unit, units))
        tmp_out = OUT + ".tmp"
        _run([hipcc, *HIP_FLAGS, "-shared", "-o", tmp_out, *objs], verbose)
        os.replace(tmp_out, OUT)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
