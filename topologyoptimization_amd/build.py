"""Build libtopo_hip.so (HIP, gfx950) and the Node-API addon in-tree.

    python -m topologyoptimization_amd.build          # everything
    python -m topologyoptimization_amd.build --hip    # the HIP library only
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtopo_hip.so")
ADDON_DIR = os.path.join(HERE, "node")
ADDON = os.path.join(ADDON_DIR, "topo_addon.node")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 VALU ops are kept out of the code objects
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HIP_SOURCES = ["solver.hip", "host_setup.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if out.strip():
        print(out)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s" % " ".join(cmd))


def build_hip():
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-o", tmp] + srcs + ["-lrocblas"])
    os.replace(tmp, LIB)
    return LIB


def build_addon():
    """N-API addon (C++, no GPU code) linking libtopo_hip.so by rpath."""
    src = os.path.join(ADDON_DIR, "topo_addon.cpp")
    if not os.path.exists(src):
        return None
    node_inc = None
    for cand in ("/usr/include/node", "/usr/local/include/node"):
        if os.path.exists(os.path.join(cand, "node_api.h")):
            node_inc = cand
            break
    if node_inc is None:
        print("node_api.h not found: addon not built")
        return None
    cxx = shutil.which("g++") or "g++"
    tmp = ADDON + ".tmp"
    _run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + node_inc, "-I" + os.path.join(ROOT, "include"),
          "-o", tmp, src, "-L" + HERE, "-ltopo_hip", "-Wl,-rpath,$ORIGIN/..", "-Wl,--allow-shlib-undefined"])
    os.replace(tmp, ADDON)
    return ADDON


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_hip()
    if "--hip" not in argv:
        build_addon()


if __name__ == "__main__":
    main()
