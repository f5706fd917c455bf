"""S3TC encoders and decoders (quicktex/s3tc/__init__.py), from quicktex_amd.s3tc."""
from quicktex_amd.s3tc import _bc1, _bc3, _bc4, _bc5, _interpolator  # noqa: F401
