"""Registers the built `_pathtracer` extension under its top-level name.

The reference's Python package starts with ``import _pathtracer``
(nectar_render/python/__init__.py:1); importing this module first makes that
import resolve to the MI355X build, which must exist (no fallback).
"""
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
if "_pathtracer" not in sys.modules:
    from . import _pathtracer as _module   # raises ImportError if the module was not built
    sys.modules["_pathtracer"] = _module
module = sys.modules["_pathtracer"]
