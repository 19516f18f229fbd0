"""External shear, ``src/gigalens/tf/profiles/mass/shear.py:6-14``."""
from ... import _native as nat
from ...profile import MassProfile


class Shear(MassProfile):
    _name = "SHEAR"
    _params = ["gamma1", "gamma2"]
    _kind = nat.GL_SHEAR
