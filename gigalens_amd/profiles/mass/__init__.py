"""Mass profiles (mirrors ``gigalens.tf.profiles.mass``)."""
from . import epl, shear, sie, sis, tnfw  # noqa: F401
