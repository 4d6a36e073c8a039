"""MI355X-native AnisoMPM step (drop-in for anisompm_py/src/anisompm.py's
``AnisoMPM``), backed by hand-written gfx950 HIP kernels in libpmpm.so."""
from .boundaries import HalfspaceCollider, SphereCollider, VelocityBC, halfspace_collider, sphere_collider
from .solver import AnisoMPM

__all__ = ["AnisoMPM", "sphere_collider", "halfspace_collider", "SphereCollider", "HalfspaceCollider",
           "VelocityBC"]
