"""Classes for encoding/decoding BC1 textures (quicktex/s3tc/bc1/__init__.py)."""
from ... import _export
from .. import _bc1 as _mod

__all__ = _export(_mod, globals())
