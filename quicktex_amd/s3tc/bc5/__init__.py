"""Classes for encoding/decoding BC5 textures (quicktex/s3tc/bc5/__init__.py)."""
from ... import _export
from .. import _bc5 as _mod

__all__ = _export(_mod, globals())
