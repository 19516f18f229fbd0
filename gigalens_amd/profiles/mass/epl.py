"""Elliptical power law, ``src/gigalens/tf/profiles/mass/epl.py:9-68``."""
from ... import _native as nat
from ...profile import MassProfile


class EPL(MassProfile):
    """EPL deflection by the Tessore & Metcalf (2015) series.

    ``niter`` caps the series length (epl.py:19).  The series length of each
    particle follows the reference's rule ``n < log(1e-12)/log(f) + 2``
    (epl.py:41) evaluated on that particle's own ``f``.
    """

    _name = "EPL"
    _params = ["theta_E", "gamma", "e1", "e2", "center_x", "center_y"]
    _kind = nat.GL_EPL

    def __init__(self, niter=50):
        super().__init__()
        if not 0 <= int(niter) <= nat.GL_MAX_EPL_ITER:
            raise ValueError(f"niter must be in [0, {nat.GL_MAX_EPL_ITER}]")
        self.niter = int(niter)
