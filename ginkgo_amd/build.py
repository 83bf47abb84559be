"""Build the native libraries of ginkgo_amd in-tree.

  * ginkgo_amd/lib/libgkomi.so       -- the gfx950 HIP kernel library (C ABI, include/gkomi.h)
  * ginkgo_amd/lib/libginkgo_amd.so  -- the C++ host library (gko:: API) over the C ABI

Run ``python -m ginkgo_amd.build`` (or ``__graft_entry__.build()``).  Every
HIP translation unit is compiled for gfx950 only, with packed FP32 ops
disabled in the device code generator.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ginkgo_amd")
CSRC = os.path.join(PKG, "csrc")
CORE = os.path.join(PKG, "core")
LIBDIR = os.path.join(PKG, "lib")
BUILDDIR = os.path.join(ROOT, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE,
    "-I", CSRC,
    "-Wall",
    "-Wno-unused-function",
]


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _newer(src_list, target):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(s) > t for s in src_list)


def _headers(d, exts=(".hpp", ".h")):
    out = []
    for base, _, files in os.walk(d):
        out += [os.path.join(base, f) for f in files if f.endswith(exts)]
    return out


def build_kernels(force=False, jobs=None):
    os.makedirs(BUILDDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    sources = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    hdrs = _headers(CSRC) + _headers(INCLUDE)
    objs = []
    todo = []
    for s in sources:
        o = os.path.join(BUILDDIR, os.path.basename(s) + ".o")
        objs.append(o)
        if force or _newer([s] + hdrs, o):
            todo.append((s, o))
    jobs = jobs or min(8, max(1, len(todo)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_run, [HIPCC] + HIP_FLAGS + ["-c", s, "-o", o]) for s, o in todo]
        for f in futs:
            f.result()
    lib = os.path.join(LIBDIR, "libgkomi.so")
    if force or todo or _newer(objs, lib):
        tmp = lib + ".tmp"
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
        os.replace(tmp, lib)
    return lib


def build_host(force=False, jobs=None):
    """C++ host library (gko:: API) compiled with g++, linked against libgkomi.so."""
    if not os.path.isdir(CORE):
        return None
    os.makedirs(BUILDDIR, exist_ok=True)
    srcs = sorted(os.path.join(CORE, "src", f) for f in os.listdir(os.path.join(CORE, "src"))
                  if f.endswith(".cpp"))
    hdrs = _headers(os.path.join(CORE, "include")) + _headers(INCLUDE)
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
             "-I", INCLUDE, "-I", os.path.join(CORE, "include")]
    objs, todo = [], []
    for s in srcs:
        o = os.path.join(BUILDDIR, "host_" + os.path.basename(s) + ".o")
        objs.append(o)
        if force or _newer([s] + hdrs, o):
            todo.append((s, o))
    jobs = jobs or min(8, max(1, len(todo)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_run, [CXX] + flags + ["-c", s, "-o", o]) for s, o in todo]
        for f in futs:
            f.result()
    lib = os.path.join(LIBDIR, "libginkgo_amd.so")
    if force or todo or _newer(objs + [os.path.join(LIBDIR, "libgkomi.so")], lib):
        tmp = lib + ".tmp"
        _run([CXX, "-shared", "-o", tmp] + objs +
             ["-L", LIBDIR, "-lgkomi", "-ldl", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, lib)
    return lib


def build_python(force=False):
    """pybind11 module ginkgo_amd/_gko*.so over libginkgo_amd.so."""
    import sysconfig
    import pybind11
    src = os.path.join(CORE, "python", "bindings.cpp")
    if not os.path.exists(src):
        return None
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(PKG, "_gko" + suffix)
    hdrs = _headers(os.path.join(CORE, "include")) + _headers(INCLUDE)
    host = os.path.join(LIBDIR, "libginkgo_amd.so")
    if force or _newer([src, host] + hdrs, out):
        tmp = out + ".tmp"
        _run([CXX, "-O1", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
              "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
              "-I", INCLUDE, "-I", os.path.join(CORE, "include"), src, "-o", tmp,
              "-L", LIBDIR, "-lginkgo_amd", "-lgkomi", "-Wl,-rpath,$ORIGIN/lib"])
        os.replace(tmp, out)
    return out


def build_examples(force=False):
    """C++ programs under examples/ linked against the host library."""
    exdir = os.path.join(ROOT, "examples")
    if not os.path.isdir(exdir):
        return []
    outs = []
    hdrs = _headers(os.path.join(CORE, "include")) + _headers(INCLUDE)
    for f in sorted(os.listdir(exdir)):
        if not f.endswith(".cpp"):
            continue
        src = os.path.join(exdir, f)
        out = os.path.join(LIBDIR, os.path.splitext(f)[0])
        if force or _newer([src, os.path.join(LIBDIR, "libginkgo_amd.so")] + hdrs, out):
            _run([CXX, "-O2", "-std=c++17", "-I", INCLUDE, "-I", os.path.join(CORE, "include"),
                  src, "-o", out, "-L", LIBDIR, "-lginkgo_amd", "-lgkomi",
                  "-Wl,-rpath,$ORIGIN"])
        outs.append(out)
    return outs


def build_all(force=False):
    k = build_kernels(force)
    h = build_host(force)
    p = build_python(force)
    e = build_examples(force)
    return [k, h, p] + e


if __name__ == "__main__":
    force = "--force" in sys.argv
    for p in build_all(force):
        if p:
            print(p)
