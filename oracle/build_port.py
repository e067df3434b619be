"""Compile the C restatement oracle/fxport.c into oracle/libfxport.so (gcc, host only)."""
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build() -> Path:
    out = HERE / "libfxport.so"
    cmd = ["gcc", "-O3", "-mavx2", "-mfma", "-fPIC", "-shared", "-std=gnu11",
           str(HERE / "fxport.c"), "-o", str(out), "-lpthread", "-ldl", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("build of libfxport.so failed")
    return out


if __name__ == "__main__":
    print(build())
