"""In-tree build of the gfx950 library and the pybind11 module.

    python -m guanaco_amd._build

compiles guanaco_amd/libguanaco_hip.so (hipcc, gfx950) and
guanaco_amd/guanaco_ext*.so (g++, pybind11, linked against the former).
"""
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")

HIP_SOURCES = ["gx_bp.hip", "gx_capi.hip", "gx_correct.hip", "gx_noalloc.hip",
               "gx_serve.hip", "gx_synth.hip"]
HOST_SOURCES = ["host/guanaco_host.cpp", "ext/guanaco_ext.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GUANACO_ARCH", "gfx950")

# No FMA contraction: the FP32 path reproduces the reference's x86-64
# operation order. Packed FP32 instructions are disabled.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
    "-packed-fp32-ops", "-I" + INCLUDE, "-I" + CSRC,
]

HIP_LIB = os.path.join(PKG, "libguanaco_hip.so")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s%s" % (
            " ".join(cmd), r.stdout, r.stderr))


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in HIP_SOURCES]
    cmds = [[HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o]
            for s, o in zip(HIP_SOURCES, objs)]
    with cf.ThreadPoolExecutor(len(cmds)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp]
         + objs, verbose)
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def ext_path():
    return os.path.join(PKG, "guanaco_ext" + sysconfig.get_config_var(
        "EXT_SUFFIX"))


def build_ext(verbose=False):
    import pybind11
    out = ext_path()
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return None
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-fvisibility=hidden",
           "-I" + INCLUDE, "-I" + CSRC, "-I" + pybind11.get_include(),
           "-I" + sysconfig.get_paths()["include"]] + srcs + [
           "-L" + PKG, "-lguanaco_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build_all(verbose=False):
    build_hip(verbose)
    build_ext(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", HIP_LIB)
