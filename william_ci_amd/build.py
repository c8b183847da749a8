"""Builds the engine's shared library william_ci_amd/libtcg.so for gfx950 with hipcc.

    python -m william_ci_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtcg.so")
SOURCES = ["tcg_eri.hip", "tcg_rows.hip", "tcg_sort.hip", "tcg_oneel.hip", "tcg_mp2.hip", "tcg_mlp.hip", "tcg_learned.hip", "tcg_setup.hip", "tcg_df.hip", "tcg_noalloc.hip", "tcg_scf.hip", "tcg_screen.hip", "tcg_host.cpp", "tcg_mlp_host.cpp", "tcg_df.cpp", "tcg_scf.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # packed FP32 instructions stay off (they are not used by this FP64 code either)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # matrix-core accumulators in ordinary VGPRs (no AGPR copies around every MFMA)
         "-Xarch_device", "-mllvm=-amdgpu-mfma-vgpr-form=1",
         "-I", os.path.join(HERE, "..", "include"), "-pthread"]


def build(verbose: bool = False) -> str:
    cmd = [HIPCC] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", OUT + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{msg}")
    if verbose and msg.strip():
        print(msg)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
