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

# heaviest first: the kernel templates are instantiated in one translation unit per (picture,
# controller, precision) table, compiled in parallel
_TABLES = [f"qt_mixed_{v}_{p}.hip" for v in ("td_gsl", "gsl", "td", "plain") for p in ("f32", "f64")] + \
          [f"qt_batch_{v}_{p}.hip" for v in ("gsl_td", "gsl", "td", "sch") for p in ("f32", "f64")]
SOURCES = [(f, "hip") for f in _TABLES] + \
          [("qt_plain.hip", "hip"), ("qt_master.hip", "hip"), ("qt_subgroup.hip", "hip"), ("qt_capi.cpp", "hip"),
           ("qt_aux.hip", "hip"), ("qt_batch.hip", "hip"), ("qt_mixed.hip", "hip"), ("qt_master_sparse.hip", "hip"),
           ("qt_system.cpp", "cxx")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _newest_header():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    hs.append(os.path.join(HERE, "..", "include", "qtraj.h"))
    return max(os.path.getmtime(h) for h in hs if os.path.exists(h))


def build(verbose=False, incremental=False):
    """Compiles every source (incremental=True: only objects older than their source or a header)
    and links the library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    hdr = _newest_header()
    objs = []
    for src, kind in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > max(os.path.getmtime(path), hdr):
            continue
        if kind == "cxx":
            cmd = ["g++"] + CXXFLAGS + ["-c", path, "-o", obj]
        else:
            cmd = ["hipcc"] + HIPFLAGS + ["-c", path, "-o", obj]
        jobs.append((cmd, obj))
    with ThreadPoolExecutor(max_workers=max(1, min(8, os.cpu_count() or 1))) as ex:
        for out in ex.map(lambda j: _run(j[0]), jobs):
            if verbose and out.strip():
                print(out)
    tmp = OUT + ".tmp"
    _run(["hipcc", "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
