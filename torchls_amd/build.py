"""Build recipe for ``libtorchls_hip.so`` (hand-written HIP for gfx950).

Run ``python -m torchls_amd.build`` (or ``__graft_entry__.build()``).  The
library is built in-tree next to this file so that it travels with the source
tree; it links only the HIP runtime and exports the C ABI declared in
``include/torchls_hip.h``.  Each source compiles to its own object (in
parallel, under ``torchls_amd/_build``), then one link step writes the library.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
INCLUDE = ROOT / "include"
OBJDIR = PACKAGE / "_build"
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


def _run(cmd, verbose):
    if verbose:
        print(" ".join(map(str, cmd)), flush=True)
    result = subprocess.run(list(map(str, cmd)), capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(result.stdout + result.stderr)
        raise RuntimeError(f"hipcc failed with exit code {result.returncode}: {cmd[-3]}")


def _stale(obj: Path, src: Path, headers: list[Path]) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(p.stat().st_mtime > t for p in [src, *headers, Path(__file__)])


def build(verbose: bool = False, incremental: bool = False) -> Path:
    """Compile every HIP source into the in-tree shared library.

    Every call recompiles every source unless ``incremental`` is set (then only
    sources newer than their object, or than any header, are recompiled)."""
    OBJDIR.mkdir(exist_ok=True)
    headers = [*CSRC.glob("*.hpp"), *INCLUDE.glob("*.h")]
    compiler = hipcc()
    jobs = []
    objs = []
    for src in sources():
        obj = OBJDIR / (src.stem + ".o")
        objs.append(obj)
        if not incremental or _stale(obj, src, headers):
            jobs.append([compiler, *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", src, "-o", obj])
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, j, verbose) for j in jobs]:
            f.result()
    tmp = LIBRARY.with_suffix(".so.tmp")
    _run([compiler, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv[1:]))
