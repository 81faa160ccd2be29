"""Builds the MI355X library cppqed_amd/libqtraj.so (HIP for gfx950 + host C++).

Usage: python -m cppqed_amd.build   (or cppqed_amd.build.build() from Python)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libqtraj.so")
ARCH = os.environ.get("QT_OFFLOAD_ARCH", "gfx950")

HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            # packed FP32 instructions stay off (they are not trusted on this hardware pool)
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC"]

SOURCES = [("qt_system.cpp", "cxx"), ("qt_capi.cpp", "hip"), ("qt_mcwf.hip", "hip"), ("qt_subgroup.hip", "hip"),
           ("qt_master.hip", "hip")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for src, kind in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        if kind == "cxx":
            cmd = ["g++"] + CXXFLAGS + ["-c", path, "-o", obj]
        else:
            cmd = ["hipcc"] + HIPFLAGS + ["-c", path, "-o", obj]
        jobs.append((cmd, obj))
    with ThreadPoolExecutor(max_workers=4) as ex:
        for out in ex.map(lambda j: _run(j[0]), jobs):
            if verbose and out.strip():
                print(out)
    tmp = OUT + ".tmp"
    _run(["hipcc", "-shared", f"--offload-arch={ARCH}", "-o", tmp] + [o for _, o in jobs])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
