"""Classes representing various methods of interpolating BC1-3 blocks (quicktex/s3tc/interpolator/__init__.py)."""
from ... import _export
from .. import _interpolator as _mod

__all__ = _export(_mod, globals())
