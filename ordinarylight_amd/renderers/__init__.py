"""Renderer implementations: ``renderers.hip`` (MI355X) and the contracts."""

from .base import (
    ProductRendererProtocol, RendererImplementation, RendererImplementationInfo,
    RendererProtocol, ResidentSceneRendererProtocol,
)


def __getattr__(name):
    # The HIP implementations load libolhip.so lazily on first use.
    if name in {"HipPathTracer", "HipRendererConfig", "HipReferenceRenderer", "HipReferenceConfig"}:
        from . import hip
        return getattr(hip, name)
    raise AttributeError(name)


__all__ = [
    "HipPathTracer", "HipReferenceConfig", "HipReferenceRenderer", "HipRendererConfig",
    "ProductRendererProtocol", "RendererImplementation", "RendererImplementationInfo",
    "RendererProtocol", "ResidentSceneRendererProtocol",
]
