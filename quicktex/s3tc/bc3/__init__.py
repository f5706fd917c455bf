"""quicktex.s3tc.bc3 (quicktex/s3tc/bc3/__init__.py), from quicktex_amd.s3tc.bc3."""
from quicktex_amd.s3tc.bc3 import *  # noqa: F401,F403
from quicktex_amd.s3tc.bc3 import __all__  # noqa: F401
