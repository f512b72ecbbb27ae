"""In-tree build of liblacmo_amd.so for gfx950 (hipcc, no JIT cache).

-ffp-contract=off keeps the reference's rounding (no fused a*b+c);
-packed-fp32-ops is disabled for all HIP code.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "liblacmo_amd.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc():
    for cand in [os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"]:
        if cand and (os.path.exists(cand) or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    """Compile csrc/lacmo_kernels.hip into lacmopy_amd/liblacmo_amd.so."""
    src = os.path.join(CSRC, "lacmo_kernels.hip")
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + [src, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
