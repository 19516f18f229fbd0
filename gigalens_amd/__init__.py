"""gigalens_amd: the GIGA-Lens forward lens model and its gradient on MI355X.

The API mirrors the reference's (``gigalens.simulator``, ``gigalens.model``,
``gigalens.tf.simulator``, ``gigalens.tf.model``, ``gigalens.tf.inference``,
``gigalens.tf.profiles``); the compute path is the gfx950 HIP library in
``gigalens_amd/_lib`` reached through the C ABI of ``include/gigalens_amd.h``.
"""
from .simulator import SimulatorConfig, LensSimulator  # noqa: F401
from .model import PhysicalModel, ForwardProbModel, BackwardProbModel  # noqa: F401

__all__ = ["SimulatorConfig", "LensSimulator", "PhysicalModel", "ForwardProbModel", "BackwardProbModel"]
