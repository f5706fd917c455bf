"""Classes for encoding/decoding BC4 textures (quicktex/s3tc/bc4/__init__.py)."""
from ... import _export
from .. import _bc4 as _mod

__all__ = _export(_mod, globals())
