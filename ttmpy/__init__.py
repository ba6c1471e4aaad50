"""``import ttmpy`` -- the reference module's name, served by ttm_amd (MI355X).

Re-exports :func:`ttm` and :func:`ttms` of :mod:`ttm_amd.ttmpy`.
"""
import os as _os
import sys as _sys

_root = _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__)))
if _root not in _sys.path:
    _sys.path.insert(0, _root)

from ttm_amd.ttmpy import ttm, ttms  # noqa: E402

__all__ = ["ttm", "ttms"]
