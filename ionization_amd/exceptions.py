"""Exception types of the mesh path (names follow ``ionization/exceptions.py:1-40``)."""


class IonizationException(Exception):
    pass


class IllegalQuantumState(IonizationException):
    pass


class InvalidPotentialParameter(IonizationException):
    pass


class InvalidMaskParameter(IonizationException):
    pass


class InvalidWindowParameter(IonizationException):
    pass


class UnknownData(IonizationException):
    pass


class MissingDatastore(IonizationException):
    pass


class DuplicateDatastores(IonizationException):
    pass


class DeviceError(IonizationException):
    """Raised when the HIP library reports a failure or is missing."""
