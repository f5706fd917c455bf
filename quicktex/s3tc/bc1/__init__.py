"""quicktex.s3tc.bc1 (quicktex/s3tc/bc1/__init__.py), from quicktex_amd.s3tc.bc1."""
from quicktex_amd.s3tc.bc1 import *  # noqa: F401,F403
from quicktex_amd.s3tc.bc1 import __all__  # noqa: F401
