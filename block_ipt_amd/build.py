"""Build libipt_hip.so (HIP for gfx950) in-tree with hipcc.

    python -m block_ipt_amd.build
"""
import os
import signal
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libipt_hip.so")
OBJ = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

SOURCES = [
    "ipt_context.cpp",
    "ipt_prep.cpp",
    "ipt_plain.cpp",
    "ipt_mixed.cpp",
    "ipt_block.cpp",
    "ipt_kernels.hip",
    "ipt_gemm_f64.hip",
    "ipt_block_kernels.hip",
    "ipt_block_w1.hip",
    "ipt_block_w2.hip",
    "ipt_block_w3.hip",
    "ipt_block_w4.hip",
    "ipt_block_w5.hip",
    "ipt_block_w6.hip",
    "ipt_block_class.hip", "ipt_block_wide.hip",
    "ipt_gemm_lp.hip",
    "ipt_gen.hip",
    "ipt_gen_sparse.hip",
    "ipt_sparse.hip",
    "ipt_sparse_lp.hip",
    "ipt_csr_tiles.hip",
    "ipt_acx.hip",
    "ipt_mlp.hip",
    "ipt_profile.cpp",
    "ipt_fci.cpp",
    "ipt_eig.hip",
    "ipt_chol.hip",
]

# Packed-FP32 VALU ops are disabled for all device code.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-fopenmp",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def build(verbose=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [hipcc, "-x", "hip"] + HIP_FLAGS + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        # own session, so a timed-out compile can be ended with its children
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            start_new_session=True)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    link = [hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-fopenmp", "-o", OUT + ".tmp"] + objs + [
        "-L" + os.path.join(ROCM, "lib"), "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    if verbose:
        print(" ".join(link))
    subprocess.run(link, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


def _drain(procs):
    failed = []
    for src, p in procs:
        try:
            # a runaway optimisation must fail the build, not hang it
            out, _ = p.communicate(timeout=900)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            out, _ = p.communicate()
            failed.append((src, "hipcc timed out after 900 s\n" + out.decode(errors="replace")))
            continue
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif text.strip():
            lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if lines:
                print("\n".join(lines))
    procs.clear()
    if failed:
        for src, text in failed:
            print("=== %s ===\n%s" % (src, text), file=sys.stderr)
        raise RuntimeError("hipcc failed for: " + ", ".join(s for s, _ in failed))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
