"""Profile descriptors with the reference's interface (src/gigalens/profile.py:5-77).

In this package a profile is a *descriptor*: it names the profile kind and
its parameters; no per-profile Python arithmetic runs on the hot path.  A
``PhysicalModel`` of descriptors is compiled into a device plan.
"""
from __future__ import annotations

from abc import ABC
from typing import List


class Parameterized(ABC):
    """``gigalens.profile.Parameterized`` (profile.py:5-21)."""

    _name: str
    _params: List[str]
    _kind: int = 0

    def __init__(self, *args, **kwargs):
        self.name = self._name
        self.params = list(self._params)

    def __str__(self):
        return self.name


class LightProfile(Parameterized):
    """``gigalens.profile.LightProfile`` (profile.py:24-58): ``use_lstsq`` adds
    the linear ``amp`` parameter solved by least squares."""

    def __init__(self, use_lstsq=False, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._use_lstsq = use_lstsq
        self.depth = 1
        if self._use_lstsq:
            self.params.append("amp")

    @property
    def use_lstsq(self):
        return self._use_lstsq

    @use_lstsq.setter
    def use_lstsq(self, use_lstsq: bool):
        if use_lstsq and not self._use_lstsq:
            self.params.append("amp")
        elif not use_lstsq and self._use_lstsq:
            self.params.pop()
        self._use_lstsq = use_lstsq


class MassProfile(Parameterized):
    """``gigalens.profile.MassProfile`` (profile.py:61-77)."""
