"""Truncated NFW halo (Baltz, Marshall & Oguri 2009 truncation),
``src/gigalens/tf/profiles/mass/tnfw.py:9-58``."""
from ... import _native as nat
from ...profile import MassProfile


class TNFW(MassProfile):
    _name = "TNFW"
    _params = ["Rs", "alpha_Rs", "r_trunc", "center_x", "center_y"]
    _kind = nat.GL_TNFW
