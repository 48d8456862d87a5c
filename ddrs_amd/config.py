"""The subset of the reference's routing configuration the path reads
(src/config.rs AttributeMinimums 426-446, ParameterRanges 449-479,
Params 580-672)."""
from __future__ import annotations

from dataclasses import dataclass, field

from ._native import DdrsParams

DT_SECONDS = 3600.0  # src/routing/mmc.rs:33


@dataclass
class AttributeMinimums:
    discharge: float = 1e-4
    slope: float = 1e-3
    velocity: float = 0.01
    depth: float = 0.01
    bottom_width: float = 0.01


@dataclass
class ParameterRanges:
    n: tuple = (0.015, 0.25)
    q_spatial: tuple = (0.0, 1.0)
    p_spatial: tuple = (1.0, 200.0)


@dataclass
class Params:
    attribute_minimums: AttributeMinimums = field(default_factory=AttributeMinimums)
    parameter_ranges: ParameterRanges = field(default_factory=ParameterRanges)
    log_space_parameters: list = field(default_factory=lambda: ["p_spatial"])
    defaults: dict = field(default_factory=lambda: {"p_spatial": 21.0})
    ddr_match: bool = False
    enforce_positivity: bool = False
    dt: float = DT_SECONDS

    def to_c(self) -> DdrsParams:
        m = self.attribute_minimums
        return DdrsParams(self.dt, m.discharge, m.depth, m.velocity, m.bottom_width, m.slope,
                          int(bool(self.ddr_match)), int(bool(self.enforce_positivity)))
