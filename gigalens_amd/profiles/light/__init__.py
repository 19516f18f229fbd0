"""Light profiles (mirrors ``gigalens.tf.profiles.light``)."""
from . import sersic, shapelets  # noqa: F401
