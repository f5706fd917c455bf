"""quicktex.s3tc.bc4 (quicktex/s3tc/bc4/__init__.py), from quicktex_amd.s3tc.bc4."""
from quicktex_amd.s3tc.bc4 import *  # noqa: F401,F403
from quicktex_amd.s3tc.bc4 import __all__  # noqa: F401
