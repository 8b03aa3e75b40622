"""``import rednoise``: the reference binding's import name for this library.

Scripts written against the reference's Python binding
(bindings/python/rednoise/__init__.py: ``RenderMode``, ``Scene``, ``Image``,
``version``) run unchanged: every public name of :mod:`rednoise_amd` is
re-exported here, the four names above first among them. The torch
integration is :mod:`rednoise.torch` (``PathReplayRender``, ``render``).
"""
from rednoise_amd import *  # noqa: F401,F403
from rednoise_amd import __all__ as _amd_all
from rednoise_amd import Image, RenderMode, Scene, __version__, version  # noqa: F401

__all__ = ["RenderMode", "Scene", "Image", "version"] + [n for n in _amd_all
                                                         if n not in ("RenderMode", "Scene", "Image", "version")]
