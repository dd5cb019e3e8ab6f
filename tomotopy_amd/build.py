"""Build the gfx950 HIP library of the LDA engine in-tree.

``python -m tomotopy_amd.build`` compiles ``csrc/*.hip`` and ``csrc/*.cpp``
with hipcc into ``tomotopy_amd/_lib/libtomotopy_amd.so``.  The library is
the product's only compute path; nothing in the package falls back to a CPU
sampler when it is missing.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB_NAME = "libtomotopy_amd.so"

# -ffp-contract=off: the weight expression must not be fused (tomoto evaluates
# it with separate IEEE mul/add/div); explicit fma calls stay fused.
# -packed-fp32-ops: no v_pk_*_f32 generation on this platform.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc_path():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the tomotopy_amd HIP library cannot be built")


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    hipcc = hipcc_path()
    srcs = [os.path.join(CSRC, f) for f in ("lda_device.hip", "lda_bigk.hip", "lda_capi.cpp", "lda_serve.cpp")]
    objs = [os.path.join(OUT_DIR, os.path.basename(src) + ".o") for src in srcs]
    cmds = [[hipcc] + HIP_FLAGS + ["-c", src, "-o", obj] for src, obj in zip(srcs, objs)]
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    # the translation units compile side by side
    procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for cmd in cmds]
    errs = []
    for src, pr in zip(srcs, procs):
        _, err = pr.communicate()
        if pr.returncode != 0:
            errs.append("hipcc failed for %s:\n%s" % (src, err))
    if errs:
        raise RuntimeError("\n".join(errs))
    out = os.path.join(OUT_DIR, LIB_NAME)
    tmp = out + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc link failed:\n%s" % r.stderr)
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
