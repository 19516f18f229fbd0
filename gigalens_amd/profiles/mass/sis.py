"""Singular isothermal sphere, ``src/gigalens/tf/profiles/mass/sis.py:9-18``."""
from ... import _native as nat
from ...profile import MassProfile


class SIS(MassProfile):
    _name = "SIS"
    _params = ["theta_E", "center_x", "center_y"]
    _kind = nat.GL_SIS
