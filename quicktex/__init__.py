"""``quicktex`` as the reference's callers import it (quicktex/__init__.py:1-8), served
by this package's gfx950 codec: ``quicktex.RawTexture``, ``quicktex.Texture``,
``quicktex.__version__`` and the submodules ``quicktex.s3tc.{bc1,bc3,bc4,bc5,
interpolator}``, ``quicktex.dds`` and ``quicktex.image_utils`` are the objects of
``quicktex_amd`` under the reference's names."""
import quicktex_amd as _impl

__all__ = ['RawTexture', 'Texture', '__version__']


def __getattr__(name):
    if name in __all__:
        return getattr(_impl, name)  # raises ImportError when the native module is not built
    raise AttributeError(name)
