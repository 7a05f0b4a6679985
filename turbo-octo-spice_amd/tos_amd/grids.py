"""Binning grids: the three AbsTwoPointGrid geometries and the
"[min:max]*n" binning grammar of likely::createBinning (h5healxi.cc:138-140).

A uniform binning "[a:b]*n" has width (b - a)/n, low edge a + i*width and
high edge a + (i+1)*width (likely's UniformBinning arithmetic is not vendored
in the reference; this is the textbook definition).
"""
from __future__ import annotations

import re
from dataclasses import dataclass

from . import _lib

POLAR, CARTESIAN, OBSERVING = _lib.POLAR, _lib.CARTESIAN, _lib.OBSERVING
_SPEC = re.compile(r"^\s*\[\s*([^:\]]+)\s*:\s*([^\]]+)\s*\]\s*\*\s*(\d+)\s*$")


def parse_binning(spec: str):
    m = _SPEC.match(spec)
    if not m:
        raise ValueError(f"not a uniform binning '[min:max]*n': {spec!r}")
    lo, hi, n = float(m.group(1)), float(m.group(2)), int(m.group(3))
    if n <= 0 or not hi > lo:
        raise ValueError(f"bad binning {spec!r}")
    width = (hi - lo) / n
    return n, lo, width, lo + n * width


@dataclass
class GridSpec:
    type: int
    nbins: tuple
    min: tuple
    width: tuple
    max: tuple
    specs: tuple = ()

    @classmethod
    def from_specs(cls, type_: int, axis1: str, axis2: str, axis3: str = "[3.556:3.735]*1"):
        parts = [parse_binning(s) for s in (axis1, axis2, axis3)]
        return cls(type_, tuple(p[0] for p in parts), tuple(p[1] for p in parts),
                   tuple(p[2] for p in parts), tuple(p[3] for p in parts), (axis1, axis2, axis3))

    @classmethod
    def cartesian(cls, rpar="[0:200]*50", rperp="[0:200]*50", z="[3.556:3.735]*1"):
        return cls.from_specs(CARTESIAN, rpar, rperp, z)

    @classmethod
    def polar(cls, r="[0:200]*50", mu="[0:1]*20", z="[3.556:3.735]*1"):
        return cls.from_specs(POLAR, r, mu, z)

    @classmethod
    def observing(cls, dloglam="[0:0.02]*20", theta="[0:0.06]*6", z="[3.556:3.735]*1"):
        return cls.from_specs(OBSERVING, dloglam, theta, z)

    @property
    def nbins_total(self) -> int:
        return int(self.nbins[0] * self.nbins[1] * self.nbins[2])

    def to_c(self) -> _lib.Grid:
        g = _lib.Grid()
        g.type = int(self.type)
        for k in range(3):
            g.nbins[k] = int(self.nbins[k])
            g.min[k] = float(self.min[k])
            g.width[k] = float(self.width[k])
            g.max[k] = float(self.max[k])
        return g

    def max_angular_scale(self, scale: float) -> float:
        """AbsTwoPointGrid::maxAngularScale (CartesianGrid.h:46, PolarGrid.h:46,
        QuasarGrid.h:45), via the library."""
        g = self.to_c()
        return float(_lib.lib().tos_grid_max_angular_scale(g, float(scale)))
