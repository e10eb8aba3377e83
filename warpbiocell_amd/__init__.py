"""MI355X-native (gfx950 HIP) implementation of WarpBioCell's per-step hot path.

The host API mirrors the reference's modules (``cells``, ``spatial``, ``fields``,
``simulation``); arrays are torch tensors on a ROCm device and every kernel runs in the
hand-written HIP library ``libwbc_hip.so`` (C ABI in ``include/wbc.h``).
"""

__version__ = "0.1.0"
