"""chatcarlo_amd: the chatcarlo photon-transport path on one MI355X (gfx950).

Public API mirrors the reference (chatcarlo/transport.py, tally.py,
geometry.py, scene.py): run_transport, transport_photons, VoxelGrid,
Geometry, load_scene / validate_scene.
"""
__version__ = "0.1.0"

from .geometry import Geometry  # noqa: E402,F401
from .scene import Scene, SceneError, load_scene, validate_scene  # noqa: E402,F401
from .tally import ScalarMoments, VoxelGrid  # noqa: E402,F401


def __getattr__(name):
    # transport needs the native library; import it lazily so that scene
    # validation and the data pack work on hosts without the built library
    if name in ("run_transport", "transport_photons", "TransportResult", "BatchResult", "Engine"):
        from . import transport
        return getattr(transport, name)
    raise AttributeError(name)
