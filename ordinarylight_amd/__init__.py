"""MI355X-native wavefront path tracer behind OrdinaryLight's renderer API.

``Renderer`` mirrors ``ordinarylight.Renderer``; ``renderers.hip`` holds the
drop-in implementations (GI estimator and reference estimator) that run on
``libolhip.so``.
"""

from .cameras import CAMERA_TYPES, OrthographicCamera, PanoramicCamera, PerspectiveCamera
from .capabilities import RendererCapabilities, capabilities_from_renderer
from .lights import DirectionalLight, EnvironmentLight, PointLight, SpotLight
from .renderer import RenderFrame, RenderJob, RenderStatistics, Renderer
from .scene import Instance, Material, Mesh, MeshResource, Scene, Transform

__version__ = "0.1.0"

__all__ = [
    "CAMERA_TYPES", "DirectionalLight", "EnvironmentLight", "Instance", "Material", "Mesh",
    "MeshResource", "OrthographicCamera", "PanoramicCamera", "PerspectiveCamera", "PointLight",
    "RenderFrame", "RenderJob", "RenderStatistics", "Renderer", "RendererCapabilities", "Scene",
    "SpotLight", "Transform", "capabilities_from_renderer",
]
