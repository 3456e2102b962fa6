"""Builds libuppe.so (all HIP kernels + the C ABI) in-tree for gfx950.

    python -m project_amd.build        (or project_amd.build.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "csrc", f) for f in ("uppe3d.hip", "uppe1d.hip", "rf3d.hip")
       if os.path.exists(os.path.join(HERE, "csrc", f))]
OUT = os.path.join(HERE, "libuppe.so")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed-FP32 instructions (see the package notes on v_pk_* FP32 forms)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose=False):
    """Compiles every source to an object in parallel (no cross-file device calls, so no
    relocatable device code), then links the shared library."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.basename(f) + ".o") for f in SRC]
    compile_flags = [f for f in FLAGS if f != "-shared"]
    cmds = [[hipcc] + compile_flags + ["-c", src, "-o", obj] for src, obj in zip(SRC, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        results = list(ex.map(lambda c: _run(c, verbose), cmds))
    for r in results:
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed compiling a libuppe.so source")
    tmp = OUT + ".tmp"
    r = _run([hipcc] + FLAGS + objs + ["-o", tmp], verbose)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libuppe.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
