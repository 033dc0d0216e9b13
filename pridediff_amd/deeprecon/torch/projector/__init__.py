"""Projector API: same public names as deeprecon/deeprecon/torch/projector."""
from .projector import projector2d, projector3d, projectorsp  # noqa: F401
