"""Renderer capability records (``ordinarylight/capabilities/_core.py:10-99``)."""

from __future__ import annotations

from dataclasses import dataclass, field, replace
from types import MappingProxyType
from typing import Any, Mapping


@dataclass(frozen=True)
class RendererCapabilities:
    """Immutable description of one initialised renderer implementation."""

    renderer: str
    outputs: tuple = ("color",)
    features: frozenset = frozenset()
    limits: Mapping = field(default_factory=dict)
    device: Any | None = None
    selection: Mapping = field(default_factory=dict)

    def __post_init__(self):
        if not isinstance(self.renderer, str) or not self.renderer:
            raise TypeError("renderer must be a non-empty string")
        outputs = tuple(self.outputs)
        if not outputs or any(not isinstance(n, str) or not n for n in outputs):
            raise TypeError("outputs must contain non-empty strings")
        features = frozenset(self.features)
        if any(not isinstance(n, str) or not n for n in features):
            raise TypeError("features must contain non-empty strings")
        limits = dict(self.limits)
        if any(not isinstance(v, (int, float)) for v in limits.values()):
            raise TypeError("limit values must be numeric")
        object.__setattr__(self, "outputs", outputs)
        object.__setattr__(self, "features", features)
        object.__setattr__(self, "limits", MappingProxyType(limits))
        object.__setattr__(self, "selection", MappingProxyType(dict(self.selection)))

    def supports(self, feature: str) -> bool:
        return feature in self.features

    def supports_output(self, output: str) -> bool:
        return output in self.outputs

    def require(self, *features: str) -> None:
        missing = tuple(f for f in features if f not in self.features)
        if missing:
            raise RuntimeError(f"renderer {self.renderer!r} does not support {missing}")

    def as_dict(self) -> dict:
        return {
            "renderer": self.renderer, "outputs": self.outputs,
            "features": tuple(sorted(self.features)), "limits": dict(self.limits),
            "device": getattr(self.device, "name", self.device),
            "selection": dict(self.selection),
        }


def capabilities_from_renderer(renderer) -> RendererCapabilities:
    """Normalise an optional capability declaration (dict or record)."""
    declared = getattr(renderer, "capabilities", None)
    if isinstance(declared, RendererCapabilities):
        return declared
    values = dict(declared or {})
    return RendererCapabilities(
        renderer=values.pop("renderer", type(renderer).__name__),
        outputs=tuple(values.pop("outputs", getattr(renderer, "available_outputs", ("color",)))),
        features=frozenset(values.pop("features", ())),
        limits=values.pop("limits", {}),
        device=values.pop("device", getattr(renderer, "device", None)),
        selection=values.pop("selection", getattr(renderer, "renderer_selection", {})),
    )


__all__ = ["RendererCapabilities", "capabilities_from_renderer", "replace"]
