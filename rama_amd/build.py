"""Build librama_hip.so (HIP, gfx950) and the rama_py extension in-tree.

    python -m rama_amd.build            # build everything
    python -m rama_amd.build --lib      # only the HIP library

hipcc compiles each .hip translation unit for gfx950 in parallel and links
them into rama_amd/librama_hip.so.  rama_py (pybind11) is host-only C++ that
calls the C ABI, built with g++ into rama_amd/rama_py<EXT_SUFFIX>; the
command-line front end rama_amd/rama_text_input likewise.
"""
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librama_hip.so")
ARCH = os.environ.get("RAMA_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["profiler", "pool", "prims", "dcoo", "matching", "cycles", "mp", "dual", "forest", "solver", "capi", "batch", "lockstep", "tuning"]

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    # float expressions are not fused into FMAs: the message-passing sweep
    # rounds exactly like its CPU restatement
    "-ffp-contract=off",
    # no packed-FP32 instructions are generated
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(line for line in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in line)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


def build_lib(jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()

    def one(name):
        src = os.path.join(CSRC, name + ".hip")
        obj = os.path.join(BUILD, name + ".o")
        _run([cc] + HIPCC_FLAGS + ["-c", src, "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    tmp = LIB + ".tmp%d" % os.getpid()
    _run([cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_py():
    """pybind11 module rama_py on top of the C ABI (host-only, g++)."""
    import pybind11

    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "rama_py" + ext)
    src = os.path.join(CSRC, "rama_py.cpp")
    tmp = out + ".tmp%d" % os.getpid()
    cmd = [
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
        "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"], "-I" + os.path.join(ROOT, "include"),
        src, "-o", tmp, "-L" + HERE, "-lrama_hip", "-Wl,-rpath,$ORIGIN",
    ]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_cli():
    """rama_text_input: the command-line front end over include/rama_cuda.h
    (host-only C++, g++, linked against librama_hip.so)."""
    out = os.path.join(HERE, "rama_text_input")
    tmp = out + ".tmp%d" % os.getpid()
    cmd = [
        "g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "rama_text_input.cpp"),
        "-o", tmp, "-L" + HERE, "-lrama_hip", "-Wl,-rpath,$ORIGIN",
    ]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build(lib=True, py=True, cli=True):
    built = []
    if lib:
        built.append(build_lib())
    if py and os.path.exists(os.path.join(CSRC, "rama_py.cpp")):
        built.append(build_py())
    if cli:
        built.append(build_cli())
    return built


if __name__ == "__main__":
    only_lib = "--lib" in sys.argv
    for p in build(lib=True, py=not only_lib, cli=not only_lib):
        print(p)
