"""Small types shared by the mesh path (``ionization/core.py:122-129``)."""

import dataclasses
import enum


class Gauge(str, enum.Enum):
    LENGTH = "LEN"
    VELOCITY = "VEL"


class WarningRecord:
    """One recorded warning: the time index it happened at and its message (``core.py:122-124``)."""

    __slots__ = ("time_index", "message")

    def __init__(self, time_index, message):
        self.time_index = time_index
        self.message = message

    def __repr__(self):
        return f"WarningRecord(time_index={self.time_index}, message={self.message!r})"


@dataclasses.dataclass(frozen=True)
class SphericalHarmonic:
    """
    The key of one spherical harmonic Y_l^m. The reference takes this type from
    simulacra (``simulacra.math.SphericalHarmonic``, used at sims.py:1051-1053
    and data.py:441-443) and only needs it there as a hashable (l, m) label.
    """

    l: int = 0
    m: int = 0

    def __post_init__(self):
        if self.l < 0 or abs(self.m) > self.l:
            raise ValueError(f"invalid spherical harmonic (l={self.l}, m={self.m})")

    def __str__(self):
        return f"Y_({self.l},{self.m})"
