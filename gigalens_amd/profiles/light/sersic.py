"""Sersic light profiles, ``src/gigalens/tf/profiles/light/sersic.py:9-121``."""
from ... import _native as nat
from ...profile import LightProfile


class Sersic(LightProfile):
    """Circular Sersic, I = Ie exp(-b_n ((R/R_s)^(1/n) - 1)), b_n = 1.9992 n - 0.3271."""

    _name = "SERSIC"
    _params = ["R_sersic", "n_sersic", "center_x", "center_y"]
    _kind = nat.GL_SERSIC


class SersicEllipse(Sersic):
    """Elliptical Sersic with complex ellipticity (e1, e2) (sersic.py:62-72)."""

    _name = "SERSIC_ELLIPSE"
    _params = ["R_sersic", "n_sersic", "e1", "e2", "center_x", "center_y"]
    _kind = nat.GL_SERSIC_ELLIPSE


class CoreSersic(Sersic):
    """Core-Sersic as the reference writes it (sersic.py:75-121):
    I = Ie (1 + (Rb/R)^alpha)^(gamma/alpha) exp(-b_n (R^alpha + Rb^alpha) / (R_s^alpha alpha n) - 1)
    on the elliptical distance R."""

    _name = "CORE_SERSIC"
    _params = ["R_sersic", "n_sersic", "Rb", "alpha", "gamma", "e1", "e2", "center_x", "center_y"]
    _kind = nat.GL_CORE_SERSIC
