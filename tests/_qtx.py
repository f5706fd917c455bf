"""The product's C ABI (include/qtx.h) as the tests call it: through ctypes,
exactly as an FFI client would.  See quicktex_amd/capi.py."""
from quicktex_amd.capi import LIB_PATH, Qtx, QtxError, declared_symbols  # noqa: F401
