"""
Host-side physics inputs of the evolve step: static potentials V(r), the
time-dependent field amplitude E(t), time windows and absorbing masks.

These run on the host only. The device path consumes them as per-simulation
tables (E sampled at every step, V and the mask sampled on the radial grid),
built once when a simulation is compiled for the device.
"""

import numpy as np

from .. import units as u
from .. import exceptions
from ..summables import Summand, Sum


def _require(ok, error_type, owner, message):
    """Raise ``error_type`` naming the class whose parameter failed (the reference's error types)."""
    if not ok:
        raise error_type(f"{owner}: {message}")


# ---------------------------------------------------------------- potentials


class PotentialEnergy(Summand):
    """A potential energy term; ``+`` builds a PotentialEnergySum (``potentials/potential.py:14-20``)."""

    @property
    def summation_class(self):
        return PotentialEnergySum


class PotentialEnergySum(Sum, PotentialEnergy):
    container_name = "potentials"

    def get_electric_field_amplitude(self, t):
        return sum(x.get_electric_field_amplitude(t) for x in self._container)


class NoPotentialEnergy(PotentialEnergy):
    def __call__(self, *args, **kwargs):
        return 0


class CoulombPotential(PotentialEnergy):
    """k q Q / r (``potentials/static.py:12-38``)."""

    def __init__(self, charge: float = 1 * u.proton_charge):
        self.charge = charge

    def __call__(self, *, r, test_charge, **kwargs):
        return u.coulomb_constant * self.charge * test_charge / r

    def __repr__(self):
        return f"CoulombPotential(charge={self.charge / u.proton_charge} e)"


class HarmonicOscillator(PotentialEnergy):
    """1/2 k (r - center)^2, constant beyond an optional cutoff (``potentials/static.py:113-233``)."""

    def __init__(self, spring_constant: float = 4.20521 * u.N / u.m, center: float = 0.0, cutoff_distance=None):
        self.spring_constant = spring_constant
        self.center = center
        self.cutoff_distance = cutoff_distance

    @classmethod
    def from_frequency_and_mass(cls, omega: float = 1.5192675e15 * u.Hz, mass: float = u.electron_mass, **kwargs):
        return cls(spring_constant=mass * (omega ** 2), **kwargs)

    @classmethod
    def from_ground_state_energy_and_mass(cls, ground_state_energy: float = 0.5 * u.eV, mass: float = u.electron_mass,
                                          **kwargs):
        return cls.from_frequency_and_mass(omega=2 * ground_state_energy / u.hbar, mass=mass, **kwargs)

    @classmethod
    def from_energy_spacing_and_mass(cls, energy_spacing: float = 1 * u.eV, mass: float = u.electron_mass, **kwargs):
        return cls.from_frequency_and_mass(omega=energy_spacing / u.hbar, mass=mass, **kwargs)

    def __call__(self, *, r, **kwargs):
        rr = np.asarray(r) - self.center
        inside = 0.5 * self.spring_constant * (rr ** 2)
        if self.cutoff_distance is None:
            return inside
        return np.where(np.abs(rr) <= self.cutoff_distance, inside, 0.5 * self.spring_constant * self.cutoff_distance ** 2)

    def omega(self, mass: float) -> float:
        return np.sqrt(self.spring_constant / mass)

    def __repr__(self):
        return f"HarmonicOscillator(spring_constant={self.spring_constant}, center={self.center})"


class SoftCoulombPotential(PotentialEnergy):
    """k q Q / sqrt(r^2 + s^2) (``potentials/static.py:54-91``)."""

    def __init__(self, charge: float = 1 * u.proton_charge, softening_distance: float = 0.05 * u.bohr_radius):
        self.charge = charge
        self.softening_distance = softening_distance

    def __call__(self, *, r, test_charge, **kwargs):
        numerator = u.coulomb_constant * self.charge * test_charge
        return numerator / np.sqrt(r ** 2 + self.softening_distance ** 2)


# ---------------------------------------------------------------- windows


class TimeWindow(Summand):
    @property
    def summation_class(self):
        return TimeWindowSum


class TimeWindowSum(Sum, TimeWindow):
    container_name = "windows"
    combine = staticmethod(lambda values: np.prod(np.broadcast_arrays(*values), axis=0))


class NoTimeWindow(TimeWindow):
    def __call__(self, t):
        return 1


class RectangularWindow(TimeWindow):
    """1 on [start, end], else 0 (``potentials/windows.py:36-70``)."""

    def __init__(self, start_time: float = 0 * u.asec, end_time: float = 50 * u.asec):
        self.start_time = start_time
        self.end_time = end_time

    def __call__(self, t):
        cond = np.greater_equal(t, self.start_time) * np.less_equal(t, self.end_time)
        return np.where(cond, 1, 0)


class LogisticWindow(TimeWindow):
    """Difference of two logistic steps (``potentials/windows.py:129-168``)."""

    def __init__(self, *, window_time: float, window_width: float, window_center: float = 0.0):
        self.window_time = window_time
        self.window_width = window_width
        self.window_center = window_center

    def __call__(self, t):
        tau = np.array(t) - self.window_center
        return (1 / (1 + np.exp(-(tau + self.window_time) / self.window_width))) - (
            1 / (1 + np.exp(-(tau - self.window_time) / self.window_width))
        )


class SmoothedTrapezoidalWindow(TimeWindow):
    """sin^2 ramp up over ``time_front``, plateau, cos^2 ramp down (``potentials/windows.py:186-225``)."""

    def __init__(self, *, time_front: float, time_plateau: float):
        self.time_front = time_front
        self.time_plateau = time_plateau

    def __call__(self, t):
        t = np.asarray(t, dtype=np.float64)
        front, flat_end = self.time_front, self.time_front + self.time_plateau
        # the rise has no lower bound: times before 0 follow sin^2 as well (windows.py:203-225)
        rise = np.sin(0.5 * u.pi * t / front) ** 2
        fall = np.cos(0.5 * u.pi * (t - flat_end) / front) ** 2
        return np.select([t < front, t <= flat_end, t <= flat_end + front], [rise, 1.0, fall], default=0.0)


# ---------------------------------------------------------------- fields


def simps_cumulative(y, t) -> np.ndarray:
    """
    The integral of samples y over t[0 .. i] for every i, each one by SciPy's
    ``integrate.simps`` as it stood before SciPy 1.11 (default even='avg'),
    the rule the reference's numeric vector potential is quoted with
    (pulses.py:59-78). An odd number of samples takes composite Simpson on the
    nonuniform grid (panels over samples (s, s + 1, s + 2) with
    h0 = t[s + 1] - t[s], h1 = t[s + 2] - t[s + 1]); an even number averages
    (Simpson over all but the last interval + trapezoid on it) and
    (trapezoid on the first interval + Simpson over the rest); a single
    sample integrates to 0. Two running sums of panels (those starting at
    even and at odd sample indices) give every prefix in O(n).
    """
    t = np.asarray(t, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64) * np.ones_like(t)
    out = np.zeros(len(t))
    if len(t) < 2:
        return out
    h = np.diff(t)
    h0, h1 = h[:-1], h[1:]
    hsum, hprod, h0divh1 = h0 + h1, h0 * h1, h0 / h1
    panel = hsum / 6.0 * (y[:-2] * (2 - 1.0 / h0divh1) + y[1:-1] * (hsum * hsum / hprod) + y[2:] * (2 - h0divh1))
    cum_e = np.concatenate([[0.0], np.cumsum(panel[0::2])])  # panels starting at samples 0, 2, 4, ...
    cum_o = np.concatenate([[0.0], np.cumsum(panel[1::2])])  # panels starting at samples 1, 3, 5, ...
    k = np.arange(1, len(t))  # last sample index of the prefix
    even_k = k % 2 == 0
    out[1:][even_k] = cum_e[k[even_k] // 2]
    ko = k[~even_k]
    q = (ko - 1) // 2
    val = 0.5 * (0.5 * h[ko - 1] * (y[ko] + y[ko - 1]) + 0.5 * h[0] * (y[1] + y[0]))
    out[1:][~even_k] = 0.5 * (cum_e[q] + cum_o[q]) + val
    return out


class ElectricPotential(PotentialEnergy):
    pass


class UniformLinearlyPolarizedElectricPotential(ElectricPotential):
    """
    A spatially uniform field along z. Its potential energy is -z q E(t)
    (``potentials/pulses.py:23-56``); subclasses give E(t) times the window.
    """

    def __init__(self, window: TimeWindow = None):
        self.window = NoTimeWindow() if window is None else window

    def get_electric_field_amplitude(self, t):
        return self.window(t)

    def __call__(self, *, t, z, test_charge, **kwargs):
        return -z * test_charge * self.get_electric_field_amplitude(t)

    def get_electric_field_integral_numeric(self, times, rule: str = "simps"):
        """Integral of E over ``times`` (pulses.py:59-72): rule ``'simps'`` is Simpson's rule as
        ``scipy.integrate.simps`` computed it before SciPy 1.11 (``simps_cumulative``), ``'trapz'``
        the trapezoid rule."""
        times = np.asarray(times, dtype=np.float64)
        e = np.asarray(self.get_electric_field_amplitude(times), dtype=np.float64) * np.ones(len(times))
        if rule == "simps":
            return simps_cumulative(e, times)[-1]
        if rule == "trapz":
            return float(np.sum(0.5 * (e[1:] + e[:-1]) * np.diff(times)))
        raise ValueError(f"unknown integration rule {rule!r} (expected 'simps' or 'trapz')")

    def get_vector_potential_amplitude_numeric(self, times, rule: str = "simps"):
        """The vector potential at ``times[-1]``: minus the integral of E (pulses.py:73-77)."""
        return -self.get_electric_field_integral_numeric(times, rule=rule)

    def get_electric_field_integral_numeric_cumulative(self, times):
        """Integral of E from ``times[0]`` to every time of ``times`` by the trapezoid rule, 0 first
        (``cumtrapz(..., initial=0)``, pulses.py:79-83)."""
        times = np.asarray(times, dtype=np.float64)
        e = np.asarray(self.get_electric_field_amplitude(times), dtype=np.float64) * np.ones(len(times))
        out = np.zeros(len(times))
        out[1:] = np.cumsum(0.5 * (e[1:] + e[:-1]) * np.diff(times))
        return out

    def get_vector_potential_amplitude_numeric_cumulative(self, times):
        """Minus the cumulative trapezoid integral of E (pulses.py:84-87)."""
        return -self.get_electric_field_integral_numeric_cumulative(times)

    def get_fluence_numeric(self, times, rule: str = "simps"):
        """eps0 c times the integral of E^2 over ``times`` (pulses.py:89-98)."""
        times = np.asarray(times, dtype=np.float64)
        e2 = np.abs(np.asarray(self.get_electric_field_amplitude(times), dtype=np.float64)) ** 2 * np.ones(len(times))
        if rule == "simps":
            return u.epsilon_0 * u.c * simps_cumulative(e2, times)[-1]
        if rule == "trapz":
            return u.epsilon_0 * u.c * float(np.sum(0.5 * (e2[1:] + e2[:-1]) * np.diff(times)))
        raise ValueError(f"unknown integration rule {rule!r} (expected 'simps' or 'trapz')")


class NoElectricPotential(UniformLinearlyPolarizedElectricPotential):
    def get_electric_field_amplitude(self, t):
        return np.zeros(np.shape(t)) * super().get_electric_field_amplitude(t)


class SineWave(UniformLinearlyPolarizedElectricPotential):
    """E(t) = A sin(w t + phase) * window (``potentials/pulses.py:215-246, 419-425``)."""

    def __init__(self, omega: float, amplitude: float = 1 * u.atomic_electric_field, phase=0, **kwargs):
        if omega <= 0:
            raise exceptions.InvalidPotentialParameter("omega must be positive")
        super().__init__(**kwargs)
        self.omega = omega
        self.phase = phase % u.twopi
        self.amplitude = amplitude

    @classmethod
    def from_photon_energy(cls, photon_energy: float, amplitude: float = 1 * u.atomic_electric_field, phase=0,
                           **kwargs):
        """(pulses.py:316-337)"""
        return cls(photon_energy / u.hbar, amplitude=amplitude, phase=phase, **kwargs)

    @property
    def period(self):
        return u.twopi / self.omega

    def get_electric_field_amplitude(self, t):
        return np.sin((self.omega * t) + self.phase) * self.amplitude * super().get_electric_field_amplitude(t)


DEFAULT_PULSE_WIDTH = 200 * u.asec
DEFAULT_FLUENCE = 1 * u.Jcm2
DEFAULT_PHASE = 0
DEFAULT_OMEGA_MIN = u.twopi * 30 * u.THz
DEFAULT_OMEGA_CARRIER = u.twopi * 2530 * u.THz
DEFAULT_PULSE_CENTER = 0 * u.asec


def sinc(x):
    """sin(x)/x (``potentials/pulses.py:594-596``)."""
    return np.sinc(x / u.pi)


class SincPulse(UniformLinearlyPolarizedElectricPotential):
    """
    E(t) = A sinc(dw tau / 2) cos(w_c tau + phase) * window, with dw = 2 pi / pulse_width,
    w_c = w_min + dw / 2 and A = sqrt(F dw / (pi eps0 c))
    (``potentials/pulses.py:646-695, 929-940``).
    """

    def __init__(
        self,
        pulse_width: float = DEFAULT_PULSE_WIDTH,
        fluence: float = DEFAULT_FLUENCE,
        phase: float = DEFAULT_PHASE,
        pulse_center: float = DEFAULT_PULSE_CENTER,
        omega_min: float = DEFAULT_OMEGA_MIN,
        **kwargs,
    ):
        bad = exceptions.InvalidPotentialParameter
        _require(pulse_width > 0, bad, "SincPulse", f"pulse_width must be > 0, got {pulse_width}")
        _require(fluence >= 0, bad, "SincPulse", f"fluence must be >= 0, got {fluence}")
        _require(omega_min > 0, bad, "SincPulse", f"omega_min must be > 0, got {omega_min}")
        super().__init__(**kwargs)
        self.omega_min = omega_min
        self.pulse_width = pulse_width
        self.phase = phase % u.twopi
        self.fluence = fluence
        self.pulse_center = pulse_center
        self.delta_omega = u.twopi / self.pulse_width
        self.omega_max = self.omega_min + self.delta_omega
        self.omega_carrier = (self.omega_min + self.omega_max) / 2
        self.amplitude = np.sqrt(self.fluence * self.delta_omega / (u.pi * u.epsilon_0 * u.c))

    def get_electric_field_envelope(self, t):
        tau = np.array(t) - self.pulse_center
        return sinc(self.delta_omega * tau / 2)

    def get_electric_field_amplitude(self, t):
        tau = np.array(t) - self.pulse_center
        amp = self.get_electric_field_envelope(t) * np.cos((self.omega_carrier * tau) + self.phase)
        return amp * self.amplitude * super().get_electric_field_amplitude(t)

    def __repr__(self):
        return (
            f"SincPulse(pulse_width={self.pulse_width / u.asec:.6g} as, "
            f"fluence={self.fluence / u.Jcm2:.6g} J/cm^2, phase={self.phase:.6g})"
        )


class GaussianPulse(UniformLinearlyPolarizedElectricPotential):
    """
    E(t) = A exp(-tau^2 / (2 pw^2)) cos(w_c tau + phase) * window, with
    A = sqrt(2 F / (sqrt(pi) eps0 c pw)) (``potentials/pulses.py:1016-1060, 1361-1372``).
    """

    def __init__(
        self,
        pulse_width=DEFAULT_PULSE_WIDTH,
        omega_carrier=DEFAULT_OMEGA_CARRIER,
        fluence=DEFAULT_FLUENCE,
        phase=DEFAULT_PHASE,
        pulse_center=DEFAULT_PULSE_CENTER,
        **kwargs,
    ):
        bad = exceptions.InvalidPotentialParameter
        _require(pulse_width > 0, bad, "GaussianPulse", f"pulse_width must be > 0, got {pulse_width}")
        _require(fluence >= 0, bad, "GaussianPulse", f"fluence must be >= 0, got {fluence}")
        _require(omega_carrier >= 0, bad, "GaussianPulse", f"omega_carrier must be >= 0, got {omega_carrier}")
        super().__init__(**kwargs)
        self.omega_carrier = omega_carrier
        self.pulse_width = pulse_width
        self.phase = phase % u.twopi
        self.fluence = fluence
        self.pulse_center = pulse_center
        self.delta_omega = 1 / pulse_width
        self.amplitude = np.sqrt(2 * self.fluence / (np.sqrt(u.pi) * u.epsilon_0 * u.c * self.pulse_width))

    @classmethod
    def fluence_from_pulse_width_and_amplitude(cls, pulse_width, amplitude):
        return np.sqrt(u.pi) * pulse_width * u.epsilon_0 * u.c * (amplitude ** 2) / 2

    @classmethod
    def from_omega_min(
        cls,
        pulse_width=DEFAULT_PULSE_WIDTH,
        omega_min=DEFAULT_OMEGA_MIN,
        fluence=DEFAULT_FLUENCE,
        phase=DEFAULT_PHASE,
        pulse_center=DEFAULT_PULSE_CENTER,
        **kwargs,
    ):
        """The carrier of a SincPulse with the same width and omega_min, w_min + pi / pulse_width
        (``pulses.py:1064-1100``)."""
        carrier = (omega_min + (omega_min + u.twopi / pulse_width)) / 2  # the sinc band's centre, as computed there
        pulse = cls(pulse_width=pulse_width, omega_carrier=carrier, fluence=fluence, phase=phase,
                    pulse_center=pulse_center, **kwargs)
        pulse.omega_min = omega_min
        return pulse

    @classmethod
    def from_amplitude(
        cls,
        pulse_width=DEFAULT_PULSE_WIDTH,
        omega_min=DEFAULT_OMEGA_MIN,
        amplitude=1 * u.atomic_electric_field,
        phase=DEFAULT_PHASE,
        pulse_center=DEFAULT_PULSE_CENTER,
        **kwargs,
    ):
        """Fluence derived from the amplitude, then the amplitude pinned (``pulses.py:1151-1175``)."""
        fluence = cls.fluence_from_pulse_width_and_amplitude(pulse_width, amplitude)
        pulse = cls.from_omega_min(pulse_width=pulse_width, omega_min=omega_min, fluence=fluence, phase=phase,
                                   pulse_center=pulse_center, **kwargs)
        # the amplitude recomputed from the fluence agrees to rounding; keep the caller's exact value
        if not np.isclose(amplitude / pulse.amplitude, 1):
            raise ValueError(f"GaussianPulse.from_amplitude: amplitude {amplitude} does not round-trip through "
                             f"the fluence (got {pulse.amplitude})")
        pulse.amplitude = amplitude
        return pulse

    def get_electric_field_envelope(self, t):
        tau = np.array(t) - self.pulse_center
        return np.exp(-0.5 * ((tau / self.pulse_width) ** 2))

    def get_electric_field_amplitude(self, t):
        tau = t - self.pulse_center
        amp = self.get_electric_field_envelope(t) * np.cos((self.omega_carrier * tau) + self.phase)
        return amp * self.amplitude * super().get_electric_field_amplitude(t)


# ---------------------------------------------------------------- masks


class Mask(Summand):
    @property
    def summation_class(self):
        return MaskSum


class MaskSum(Sum, Mask):
    """Masks multiply (``potentials/masks.py:19-28``)."""

    container_name = "masks"
    combine = staticmethod(lambda values: np.prod(np.broadcast_arrays(*values), axis=0))


class NoMask(Mask):
    def __call__(self, *args, **kwargs):
        return 1


class RadialCosineMask(Mask):
    """
    1 for r < r_in; |cos(pi/2 (r - r_in)/(r_out - r_in))|^(1/s) on [r_in, r_out); 0 beyond
    (``potentials/masks.py:38-89``).
    """

    def __init__(
        self,
        inner_radius: float = 50 * u.bohr_radius,
        outer_radius: float = 100 * u.bohr_radius,
        smoothness: float = 8,
    ):
        bad = exceptions.InvalidMaskParameter
        _require(min(inner_radius, outer_radius) >= 0, bad, "RadialCosineMask",
                 f"radii must be >= 0, got inner {inner_radius} and outer {outer_radius}")
        _require(outer_radius > inner_radius, bad, "RadialCosineMask",
                 f"outer_radius {outer_radius} must exceed inner_radius {inner_radius}")
        _require(smoothness >= 1, bad, "RadialCosineMask", f"smoothness must be >= 1, got {smoothness}")
        self.inner_radius = inner_radius
        self.outer_radius = outer_radius
        self.smoothness = smoothness

    def __call__(self, *, r, **kwargs):
        r = np.asarray(r, dtype=np.float64)
        width = abs(self.outer_radius - self.inner_radius)
        with np.errstate(invalid="ignore"):
            ramp = np.abs(np.cos(0.5 * u.pi * (r - self.inner_radius) / width)) ** (1 / self.smoothness)
        return np.select([r < self.inner_radius, r < self.outer_radius], [1.0, ramp], default=0.0)

    def __repr__(self):
        return (
            f"RadialCosineMask(inner_radius={self.inner_radius / u.bohr_radius:.6g} a0, "
            f"outer_radius={self.outer_radius / u.bohr_radius:.6g} a0, smoothness={self.smoothness})"
        )
