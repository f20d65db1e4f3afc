"""MI355X-native (gfx950) block-sparse normal-equation path of torchls.

Hand-written HIP kernels behind a C ABI (include/torchls_hip.h), driven from
PyTorch-ROCm host code that mirrors torchls's solver interface with
``backend="hip"``.
"""

__version__ = "0.1.0"
