"""Classes for encoding/decoding BC3 textures (quicktex/s3tc/bc3/__init__.py)."""
from ... import _export
from .. import _bc3 as _mod

__all__ = _export(_mod, globals())
