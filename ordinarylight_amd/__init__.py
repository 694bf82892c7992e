"""MI355X-native wavefront path tracer behind OrdinaryLight's renderer API.

``renderers.hip`` holds the drop-in implementations of ``RendererProtocol``
(the GI estimator and the reference estimator) that run on ``libolhip.so``;
they plug into the reference's own ``ol.Renderer``.  The scene, camera and
light records here carry the reference's field names so the workloads and
tests can build scenes without the reference installed.
"""

from .cameras import CAMERA_TYPES, OrthographicCamera, PanoramicCamera, PerspectiveCamera
from .lights import DirectionalLight, EnvironmentLight, PointLight, SpotLight
from .scene import Instance, Material, Mesh, MeshResource, Scene, Transform

__version__ = "0.2.0"

__all__ = [
    "CAMERA_TYPES", "DirectionalLight", "EnvironmentLight", "Instance", "Material", "Mesh",
    "MeshResource", "OrthographicCamera", "PanoramicCamera", "PerspectiveCamera", "PointLight",
    "Scene", "SpotLight", "Transform",
]
