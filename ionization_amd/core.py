"""Small enums shared by the mesh path (``ionization/core.py:122-129``)."""

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
