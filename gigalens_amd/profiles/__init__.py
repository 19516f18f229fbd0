"""Profile descriptors (mirrors ``gigalens.tf.profiles``)."""
from . import light, mass  # noqa: F401
