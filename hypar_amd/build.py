"""In-tree build of the native parts (run by __graft_entry__.build()).

* libhypar_mi355x.so   hand-written gfx950 HIP kernels + C-ABI (hipcc)
* libhypar_host.so /   C host layer: HyPar-style solver object, input decks,
  hypar_mi355x           NS3D plugin slots and the deck driver (gcc)
* oracle/libns3d_oracle.so   CPU restatement used by the tests (gcc)
* oracle/_ref/HyPar    the reference itself, when its sources are present
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hypar_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
REFERENCE = "/root/reference"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # no packed FP32 instructions anywhere in device code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_SOURCES = ["hm_api.hip", "hm_kernels.hip", "hm_pass.hip", "hm_slots.hip", "hm_adjoint.hip",
               "hm_noalloc.hip", "hm_bind.hip"]
LIB = os.path.join(PKG, "libhypar_mi355x.so")


def _run(cmd, cwd=ROOT):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout[-4000:] + r.stderr[-8000:])
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return r


def build_hip():
    """Compile the HIP sources in parallel (one hipcc per source), then link."""
    import concurrent.futures
    import tempfile
    with tempfile.TemporaryDirectory(prefix="hm_build_") as tdir:
        objs = [os.path.join(tdir, os.path.splitext(s)[0] + ".o") for s in HIP_SOURCES]
        cmds = [[HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(HIP_SOURCES, objs)]
        with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as ex:
            for f in [ex.submit(_run, c) for c in cmds]:
                f.result()
        tmp = LIB + ".tmp"
        _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_host():
    mk = os.path.join(HOST, "Makefile")
    if os.path.exists(mk):
        _run(["make", "-s", "-C", HOST, "-j8"])


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_reference():
    """Compile the reference from its sources (only where they are present)."""
    if not os.path.isdir(os.path.join(REFERENCE, "src")):
        return None
    _run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "Makefile.ref"),
          f"REF={REFERENCE}", "-j8"])
    return os.path.join(ROOT, "oracle", "_ref", "HyPar")


def build_all():
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found")
    build_hip()
    build_host()
    build_oracle()
    build_reference()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
