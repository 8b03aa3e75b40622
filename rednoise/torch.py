"""``rednoise.torch``: the differentiable render of :mod:`rednoise`.

``PathReplayRender`` is a ``torch.autograd.Function`` whose forward pass is
the physical estimator on the GPU (``rn_render_ex_device``) and whose backward
pass is the path-replay adjoint (``rn_adjoint_device``); ``render`` applies it.
Both are those of :mod:`rednoise_amd.differentiable`.
"""
from rednoise_amd.differentiable import PathReplayRender, render  # noqa: F401

__all__ = ["PathReplayRender", "render"]
