"""Build recipe for libdsscf.so (gfx950).

  python -m dollarstorescf_amd.build [--maxl L] [--jobs J]

1. gen_eri.py writes the per-class device code into csrc/_gen/, and one
   small translation unit per shell-quartet class (cls_NN_<class>.hip);
   gen_grad.py writes the s/p gradient classes (grad_classes.inc).
2. hipcc compiles the class units, the dispatcher and the C-ABI for gfx950
   (packed FP32 instructions disabled) in parallel; g++ compiles the host SCF
   layer with OpenMP.
3. hipcc links everything into dollarstorescf_amd/libdsscf.so (in-tree).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "_gen")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdsscf.so")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + GEN, "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-I" + os.path.join(ROOT, "include")]

HIP_SOURCES = ["eri_kernels.hip", "dsscf_abi.cpp", "scf_device.hip", "grad_kernels.hip"]
HOST_SOURCES = ["host/chem.cpp", "host/linalg.cpp", "host/scf.cpp", "host/multi.cpp", "host/serve.cpp",
                "host/job_api.cpp", "host/molden.cpp", "host/gradient.cpp"]

DEFAULT_MAXL = 2


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def class_units(maxl: int):
    """One translation unit per shell-quartet class: it instantiates that
    class's kernels and launchers (ClassLaunch in eri_class.hpp), so the
    classes compile in parallel."""
    sys.path.insert(0, CSRC)
    try:
        import gen_eri
    finally:
        sys.path.pop(0)
    _, cls_list = gen_eri.classes(maxl)
    units = []
    for k, c in enumerate(cls_list):
        name = "".join("spdfg"[l] for l in c)
        path = os.path.join(GEN, "cls_%02d_%s.hip" % (k, name))
        text = ("// Generated by build.py -- do not edit.\n#include \"eri_class.hpp\"\n"
                "template struct ClassLaunch<%d, %d, %d, %d, %d>;\n" % (*c, 1 if gen_eri.is_wave(c) else 0))
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as f:
                f.write(text)
        units.append(path)
    for f in os.listdir(GEN):  # units of classes no longer built
        if f.startswith("cls_") and os.path.join(GEN, f) not in units:
            os.remove(os.path.join(GEN, f))
    return units


def _newest(paths):
    return max((os.path.getmtime(p) for p in paths if os.path.exists(p)), default=0.0)


def _deps(src: str, seen=None):
    """The unit and every project file it includes, followed through
    #include "..." lines (searched beside the includer, then in csrc/_gen,
    csrc and include/)."""
    import re
    seen = set() if seen is None else seen
    if src in seen or not os.path.exists(src):
        return seen
    seen.add(src)
    with open(src, errors="replace") as f:
        names = re.findall(r'^\s*#\s*include\s*"([^"]+)"', f.read(), flags=re.M)
    for name in names:
        for folder in (os.path.dirname(src), GEN, CSRC, os.path.join(ROOT, "include")):
            cand = os.path.normpath(os.path.join(folder, name))
            if os.path.exists(cand):
                _deps(cand, seen)
                break
    return seen


def build(maxl: int = DEFAULT_MAXL, jobs: int = 8, verbose: bool = False, incremental: bool = False) -> str:
    os.makedirs(GEN, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    for gen, inc, extra in (("gen_eri.py", "eri_classes.inc", [str(maxl)]), ("gen_grad.py", "grad_classes.inc", [])):
        out = os.path.join(GEN, inc)
        _run([sys.executable, os.path.join(CSRC, gen), out + ".new", *extra])
        # replace only on change, so that unchanged tables keep their time stamp
        if os.path.exists(out) and open(out).read() == open(out + ".new").read():
            os.remove(out + ".new")
        else:
            os.replace(out + ".new", out)
    cmds = []
    objs = []
    # the largest classes (last in class order) first, so that they start at once
    hip_units = class_units(maxl)[::-1] + [os.path.join(CSRC, s) for s in HIP_SOURCES]
    def add(src, o, cmd):
        objs.append(o)
        # incremental (a development aid): skip units whose object is newer
        # than every file they may include
        if incremental and os.path.exists(o) and os.path.getmtime(o) > _newest(_deps(src)):
            return
        cmds.append(cmd)

    for src in hip_units:
        o = os.path.join(OBJ, os.path.basename(src) + ".o")
        add(src, o, [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", src, "-o", o])
    for s in HOST_SOURCES:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        src = os.path.join(CSRC, s)
        add(src, o, ["g++", *CXX_FLAGS, "-c", src, "-o", o])
    # the AVX2/FMA build of the host eigensolver and products (chosen at run time)
    o = os.path.join(OBJ, "linalg_avx2.cpp.o")
    src = os.path.join(CSRC, "host/linalg.cpp")
    add(src, o, ["g++", *CXX_FLAGS, "-DLA_AVX2", "-mavx2", "-mfma", "-c", src, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for r in ex.map(_run, cmds):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lgomp",
          "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--maxl", type=int, default=DEFAULT_MAXL)
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only units older than a file they may include")
    a = ap.parse_args()
    print(build(a.maxl, a.jobs, verbose=True, incremental=a.incremental))


if __name__ == "__main__":
    main()
