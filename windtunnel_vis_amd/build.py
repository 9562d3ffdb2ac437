"""Build recipe for the HIP library (libwt.so) and the Node N-API addon.

Everything is compiled in-tree: the shared objects land in windtunnel_vis_amd/_lib/ and travel
with the source tree. gfx950 (MI355X) is the only target.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIB_DIR, "libwt.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

SOURCES = ["kernels_lbm.hip", "kernels_vox.hip", "kernels_mlp.hip", "kernels_fields.hip", "wt_api.hip", "mlp_api.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


# the voxelizer is compiled without contraction so the CPU restatement reproduces it exactly
EXTRA_FLAGS = {"kernels_vox.hip": ["-ffp-contract=off"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_lib(verbose=False):
    os.makedirs(os.path.join(LIB_DIR, "obj"), exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(LIB_DIR, "obj", src.replace(".hip", ".o"))
        objs.append(obj)
        extra = EXTRA_FLAGS.get(src, [])
        cmds.append([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_node_addon(verbose=False):
    """N-API addon (plain g++, Node's own headers) linking libwt.so with an $ORIGIN rpath."""
    node_inc = "/usr/include/node"
    if not os.path.isdir(node_inc) or shutil.which("g++") is None:
        return None
    src = os.path.join(HERE, "node", "wt_addon.cc")
    if not os.path.exists(src):
        return None
    out = os.path.join(LIB_DIR, "wt_addon.node")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-DNODE_GYP_MODULE_NAME=wt_addon", "-I", node_inc,
          "-I", os.path.join(HERE, "..", "include"), src, "-o", out + ".tmp", "-L", LIB_DIR, "-lwt",
          "-Wl,-rpath,$ORIGIN", "-Wl,--unresolved-symbols=ignore-in-object-files"])
    os.replace(out + ".tmp", out)
    return out


def build(verbose=False):
    lib = build_lib(verbose)
    addon = build_node_addon(verbose)
    return lib, addon


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
