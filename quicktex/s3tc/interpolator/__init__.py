"""quicktex.s3tc.interpolator (quicktex/s3tc/interpolator/__init__.py), from quicktex_amd.s3tc.interpolator."""
from quicktex_amd.s3tc.interpolator import *  # noqa: F401,F403
from quicktex_amd.s3tc.interpolator import __all__  # noqa: F401
