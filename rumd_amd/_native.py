"""Loads the in-tree native libraries and fails loudly when they are missing:
there is no CPU fallback for any computation in this package."""
from __future__ import annotations

import ctypes
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIBRUMDX = _PKG / "librumdx.so"

if not LIBRUMDX.exists():
    raise ImportError(f"{LIBRUMDX} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
# load the kernel library first so the extension resolves it from this tree
_librumdx = ctypes.CDLL(str(LIBRUMDX), mode=ctypes.RTLD_GLOBAL)

try:
    from . import _rumd  # noqa: E402
except ImportError as e:  # pragma: no cover - reported to the user
    raise ImportError(f"the rumd_amd host extension is not built ({e}); run __graft_entry__.build()") from e



def librumdx() -> ctypes.CDLL:
    """The loaded C-ABI kernel library (include/rumdx.h)."""
    return _librumdx


__all__ = ["_rumd", "LIBRUMDX", "librumdx"]
