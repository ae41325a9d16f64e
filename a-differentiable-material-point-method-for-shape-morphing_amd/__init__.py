"""MI355X-native differentiable MLS-MPM step (DiffMPMLib3D hot path)."""
from . import scenes  # noqa: F401
