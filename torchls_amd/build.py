"""Build recipe for ``libtorchls_hip.so`` (hand-written HIP for gfx950).

Run ``python -m torchls_amd.build`` (or ``__graft_entry__.build()``).  The
library is built in-tree next to this file so that it travels with the source
tree; it links only the HIP runtime and exports the C ABI declared in
``include/torchls_hip.h``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
INCLUDE = ROOT / "include"
LIBRARY = PACKAGE / "libtorchls_hip.so"
ARCH = os.environ.get("TORCHLS_HIP_ARCH", "gfx950")

# Packed-FP32 VALU ops are disabled (the compiler would otherwise form
# v_pk_{add,mul,fma}_f32 from ordinary float code); code object v5 keeps the
# library loadable by the HIP runtime bundled with PyTorch.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-mcode-object-version=5",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-ffp-contract=fast",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = False) -> Path:
    """Compile every HIP source into the in-tree shared library (always rebuilds)."""
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", *map(str, sources()), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(result.stdout + result.stderr)
        raise RuntimeError(f"hipcc failed with exit code {result.returncode}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
