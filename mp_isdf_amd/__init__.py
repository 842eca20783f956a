"""mp_isdf_amd: MI355X-native ISDF anchor-selection hot path.

``mp_isdf_amd.mp_isdf`` is the compiled pybind11 module with the reference
extension's surface (LibcintContext, AnchorSelection, enums, eval_ao*).  The
package directory is also put on ``sys.path`` so ``import mp_isdf`` resolves
to the same module, as the reference driver expects (python/isdf.py:24-25).
Importing fails loudly when the HIP extension has not been built: there is no
CPU fallback.
"""
from __future__ import annotations

import sys
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
LIBRARY = PACKAGE_DIR / "libmp_isdf_amd.so"

if str(PACKAGE_DIR) not in sys.path:
    sys.path.insert(0, str(PACKAGE_DIR))

# torch bundles its own HIP runtime under the same SONAME (libamdhip64.so.7)
# as the system ROCm the extension links.  Whichever loads first serves the
# whole process, and torch cannot see the GPU when it arrives second, so
# torch (plumbing for device buffers and autograd) is imported first.
try:
    import torch  # noqa: F401
except ImportError:  # the extension itself does not need torch
    pass


def __getattr__(name):
    if name == "mp_isdf":
        if not LIBRARY.exists():
            raise ImportError(f"{LIBRARY} is missing; run __graft_entry__.build() (hipcc, gfx950)")
        import mp_isdf as module  # the compiled extension next to this file
        return module
    raise AttributeError(name)
