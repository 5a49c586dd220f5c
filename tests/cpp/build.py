"""Build the C++ drop-in header tests against the in-tree library.

    python tests/cpp/build.py        # (also run by __graft_entry__.build())
"""
import glob
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "_bin")


def build():
    os.makedirs(BIN, exist_ok=True)
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    for src in sorted(glob.glob(os.path.join(HERE, "*.cpp"))):
        out = os.path.join(BIN, os.path.basename(src)[:-4])
        cmd = [hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I" + os.path.join(ROOT, "include"), src,
               "-L" + os.path.join(ROOT, "hashinator_amd"), "-lhashinator_amd",
               "-Wl,-rpath,$ORIGIN/../../../hashinator_amd", "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {src} failed:\n{r.stderr}")


if __name__ == "__main__":
    build()
