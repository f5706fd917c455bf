"""S3TC (BC1, BC3, BC4, BC5) encoders and decoders on gfx950 (quicktex/s3tc/__init__.py)."""
from .. import _native

_s3tc = _native()._s3tc
_interpolator, _bc1, _bc3, _bc4, _bc5 = _s3tc._interpolator, _s3tc._bc1, _s3tc._bc3, _s3tc._bc4, _s3tc._bc5
