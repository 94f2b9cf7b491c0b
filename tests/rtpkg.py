"""Imports the package module (its directory name is not a Python identifier)."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vulkan-restir-pt_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import restirpt as rt  # noqa: E402

__all__ = ["rt", "ROOT", "PKG"]
