"""quicktex.s3tc.bc5 (quicktex/s3tc/bc5/__init__.py), from quicktex_amd.s3tc.bc5."""
from quicktex_amd.s3tc.bc5 import *  # noqa: F401,F403
from quicktex_amd.s3tc.bc5 import __all__  # noqa: F401
