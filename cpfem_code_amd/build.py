"""Build the HIP library libcpfem.so for gfx950 (in-tree, no JIT cache).

    python -m cpfem_code_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcpfem.so")

SOURCES = ["constitutive.hip", "mesh.hip", "calib.hip"]

HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            sys.stderr.write(text)
            raise RuntimeError("hipcc failed: " + " ".join(cmd))
        if verbose and text.strip():
            print(text)
    tmp = LIB + ".tmp"
    subprocess.check_call([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_host()
    build_test_helpers()
    return LIB


TEST_NATIVE = os.path.join(os.path.dirname(HERE), "tests", "native")


def build_test_helpers() -> str:
    """Host build of the constitutive adjoint's per-point step for the CPU tests
    (tests/native/libadjoint_host.so); no product path loads it."""
    src = os.path.join(TEST_NATIVE, "adjoint_host.hip")
    out = os.path.join(TEST_NATIVE, "libadjoint_host.so")
    subprocess.check_call([hipcc()] + HIPFLAGS + ["-shared", "-o", out, src],
                          stderr=subprocess.DEVNULL)
    return out


FLANG = "/opt/rocm/lib/llvm/bin/flang"
HOST_EXE = os.path.join(LIBDIR, "fasolt_gpu")


def build_host() -> str:
    """The Fortran driver (host/fasolt_gpu.f90) linked against libcpfem.so."""
    if not os.path.exists(FLANG):
        raise RuntimeError("flang not found: the Fortran host needs /opt/rocm/lib/llvm/bin/flang")
    hostdir = os.path.join(HERE, "host")
    tmp = os.path.join(LIBDIR, "_host_build")
    os.makedirs(tmp, exist_ok=True)
    objs = []
    for src in ("cpfem_capi.f90", "fasolt_gpu.f90"):
        obj = os.path.join(tmp, src.replace(".f90", ".o"))
        subprocess.check_call([FLANG, "-O2", "-module-dir", tmp, "-c", os.path.join(hostdir, src), "-o", obj])
        objs.append(obj)
    subprocess.check_call([FLANG, "-o", HOST_EXE] + objs + ["-L" + LIBDIR, "-lcpfem", "-Wl,-rpath,$ORIGIN"])
    for f in os.listdir(tmp):
        os.remove(os.path.join(tmp, f))
    os.rmdir(tmp)
    return HOST_EXE


if __name__ == "__main__":
    print(build(verbose=True))
