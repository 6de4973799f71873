"""Builds libtcscf_hip.so (gfx950) in-tree with hipcc.

    python -m tcscf_amd.build            # or tcscf_amd.build.build()

Every HIP translation unit is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtcscf_hip.so")
SOURCES = ["kernels.hip", "scf.hip", "tc_scf.hip", "noalloc.hip", "host_grids.cpp", "dense.cpp", "batch.cpp"]
ARCH = os.environ.get("TCSCF_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=None):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    common = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
              "-Wno-unused-lambda-capture"]
    # code object v5: loadable by the HIP runtime torch bundles as well as /opt/rocm's
    hip_flags = [f"--offload-arch={ARCH}", "-mcode-object-version=5", "-Xclang", "-target-feature", "-Xclang",
                 "-packed-fp32-ops", "-munsafe-fp-atomics"]
    # the dense FP64 solvers: complex products and quotients inline (the
    # textbook product, Smith's quotient) instead of the C99 runtime routines
    # with their inf / NaN recovery, which the finite SCF matrices never need
    per_file = {"dense.cpp": ["-fcomplex-arithmetic=improved"]}
    objs, procs = [], []
    for s in srcs:
        obj = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(obj)
        cmd = ([hipcc()] + common + per_file.get(os.path.basename(s), []) +
               (hip_flags + ["-x", "hip"] if s.endswith(".hip") else []) + ["-c", s, "-o", obj])
        if verbose:
            print(" ".join(cmd))
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = False
    for s, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {s}\n{out}\n")
        elif verbose and out.strip():
            sys.stderr.write(f"--- {s}\n{out}\n")
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-mcode-object-version=5", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_facade_test(verbose)
    return LIB


def build_facade_test(verbose=False):
    """The C++ facade driver (tests/cpp/facade_test.cpp over include/tcscf.hpp),
    linked against the in-tree library; rpath $ORIGIN/.. finds it there."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "facade_test.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(HERE, "_build", "facade_test")
    cmd = [hipcc(), "-O2", "-std=c++17", "-Wall", "-I", os.path.join(root, "include"), src, "-o", out + ".tmp",
           "-L", HERE, "-ltcscf_hip", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
