"""Singular isothermal ellipsoid, ``src/gigalens/tf/profiles/mass/sie.py:9-43``.

The deflection is the reference's (s_scale = 0); the kernels evaluate it in a
form that stays finite at |e| = 0, where the reference's expression is 0/0.
"""
from ... import _native as nat
from ...profile import MassProfile


class SIE(MassProfile):
    _name = "SIE"
    s_scale = 1e-4
    _params = ["theta_E", "e1", "e2", "center_x", "center_y"]
    _kind = nat.GL_SIE
