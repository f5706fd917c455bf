"""quicktex_amd: quicktex's S3TC codec API on MI355X (gfx950).

Mirrors the reference package layout (quicktex/__init__.py:1-8):

    quicktex_amd.RawTexture, quicktex_amd.Texture, quicktex_amd.__version__
    quicktex_amd.s3tc.bc1 / bc3 / bc4 / bc5 / interpolator

All encode and decode work runs on the GPU through the C ABI in include/qtx.h;
there is no CPU codec.  Build the native parts with ``python -m quicktex_amd.build``.
"""
_import_error = None
try:
    from ._quicktex import RawTexture, Texture, __version__  # noqa: F401
except ImportError as e:  # not built yet: fail when the codec is used, not on import
    _import_error = e


def __getattr__(name):
    if name in ('RawTexture', 'Texture', '__version__') and _import_error is not None:
        raise ImportError('quicktex_amd native module is not built (python -m quicktex_amd.build): %s'
                          % _import_error)
    raise AttributeError(name)


def _native():
    if _import_error is not None:
        raise ImportError('quicktex_amd native module is not built (python -m quicktex_amd.build): %s'
                          % _import_error)
    from . import _quicktex
    return _quicktex


def _export(module, namespace):
    """`from module import *` for a pybind11 submodule; returns the exported names."""
    names = [n for n in dir(module) if not n.startswith('_')]
    for n in names:
        namespace[n] = getattr(module, n)
    return names
