"""Mass profiles (mirrors ``gigalens.tf.profiles.mass``)."""
from . import epl, shear, sis  # noqa: F401
