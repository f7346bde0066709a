"""Build the gfx950 multislice library and the stem3-compatible driver in-tree.

    python -m qstem_amd.build            # everything
    python -m qstem_amd.build --lib      # only libqstem_hip.so

Objects go to qstem_amd/_build/, products to qstem_amd/ (libqstem_hip.so)
and qstem_amd/bin/ (qstem3 driver).  Translation units compile in parallel.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(CSRC, "host")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqstem_hip.so")
BIN = os.path.join(HERE, "bin")
ROOT = os.path.dirname(HERE)

SIZES = (64, 128, 256, 512, 1024, 2048, 4096)


def mix_sizes():
    """Lengths with a compiled mixed-radix plan: the QS_MIX_SIZES list of qs_internal.hpp."""
    import re
    text = open(os.path.join(CSRC, "qs_internal.hpp")).read()
    m = re.search(r"#define QS_MIX_SIZES\(X\)((?:[ \t]*\\\n|[ \t]*X\(\d+\))+)", text)
    if not m:
        raise RuntimeError("QS_MIX_SIZES not found in qs_internal.hpp")
    return tuple(int(n) for n in re.findall(r"X\((\d+)\)", m.group(1)))
ARCH = os.environ.get("QS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions are disabled on purpose: some of their forms have
# given silently wrong results on MI355X, so every FP32 operation is a plain
# one-per-lane instruction (the FP32 vector rate is then half the packed figure)
HIPFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + HOST]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def _newer(target, deps):
    if not os.path.exists(target):
        return False
    t = os.path.getmtime(target)
    return all(os.path.getmtime(d) <= t for d in deps)


def build_lib(force=True, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    headers.append(os.path.join(ROOT, "include", "qstem.h"))
    jobs_list = []
    for n in SIZES:
        src = os.path.join(CSRC, "qs_sizes.hip")
        obj = os.path.join(OBJ, "qs_sizes_%d.o" % n)
        jobs_list.append((obj, [HIPCC] + HIPFLAGS + ["-DQS_N=%d" % n, "-c", src, "-o", obj], [src] + headers))
    for n in mix_sizes():
        src = os.path.join(CSRC, "qs_mix.hip")
        obj = os.path.join(OBJ, "qs_mix_%d.o" % n)
        jobs_list.append((obj, [HIPCC] + HIPFLAGS + ["-DQS_MIX_N=%d" % n, "-c", src, "-o", obj], [src] + headers))
    for name in ("qs_setup.hip", "qs_capi.cpp", "qs_potential.hip", "qs_adjoint.hip", "qs_generic.hip"):
        src = os.path.join(CSRC, name)
        if not os.path.exists(src):
            continue
        obj = os.path.join(OBJ, os.path.splitext(name)[0] + ".o")
        lang = ["-x", "hip"] if name.endswith(".cpp") else []
        jobs_list.append((obj, [HIPCC] + HIPFLAGS + lang + ["-c", src, "-o", obj], [src] + headers))
    todo = [j for j in jobs_list if force or not _newer(j[0], j[2])]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, cmd) for _, cmd, _ in todo]:
            f.result()
    objs = [o for o, _, _ in jobs_list]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_host(force=True):
    """stem3-compatible driver: C++ host code over the C ABI."""
    if not os.path.isdir(HOST):
        return None
    os.makedirs(BIN, exist_ok=True)
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    exe = os.path.join(BIN, "qstem3")
    cmd = ["g++"] + CXXFLAGS + ["-o", exe] + srcs + ["-L" + HERE, "-lqstem_hip",
                                                   "-Wl,-rpath,$ORIGIN/..", "-lm"]
    _run(cmd)
    return exe


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("--incremental", action="store_true", help="skip up-to-date objects")
    ap.add_argument("-j", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args(argv)
    print(build_lib(force=not a.incremental, jobs=a.j))
    if not a.lib:
        exe = build_host()
        if exe:
            print(exe)


if __name__ == "__main__":
    sys.exit(main())
