"""Shapelet light, ``src/gigalens/tf/profiles/light/shapelets.py:11-85``."""
from ... import _native as nat
from ...profile import LightProfile


class Shapelets(LightProfile):
    """Cartesian shapelets up to order ``n_max``; ``interpolate`` selects the
    6000-point table of the reference (zero outside |u| > 5) instead of the
    Hermite recurrence.  Amplitudes are named ``amp{i:0Nd}`` (shapelets.py:32-36)."""

    _name = "SHAPELETS"
    _params = ["beta", "center_x", "center_y"]
    _kind = nat.GL_SHAPELETS

    def __init__(self, n_max, use_lstsq=False, interpolate=True):
        super().__init__(use_lstsq=use_lstsq)
        if use_lstsq:
            self.params.pop()
        self.n_layers = int((n_max + 1) * (n_max + 2) / 2)
        self.n_max = int(n_max)
        self.interpolate = bool(interpolate)
        width = len(str(self.n_layers))
        self._amp_names = [f"amp{str(i).zfill(width)}" for i in range(self.n_layers)]
        self.params.extend(self._amp_names)
        self.depth = self.n_layers
        n1 = n2 = 0
        self.N1, self.N2 = [], []
        for _ in range(self.n_layers):
            self.N1.append(n1)
            self.N2.append(n2)
            if n1 == 0:
                n1, n2 = n2 + 1, 0
            else:
                n1, n2 = n1 - 1, n2 + 1
