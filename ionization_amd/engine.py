"""
Device engine: compiles simulations into the device's table format and
drives them through libtdse_hip.so.

A ``DeviceContext`` owns one HIP stream and one device memory pool. Any
number of simulations of different kinds and sizes live in it at once; one
``run`` call advances all of them by a number of steps (one captured
hipGraph replay per step). This is the batched path; the "plain" path is the
same context holding a single simulation.

Device layout of every line (an l-line of a SphericalHarmonicMesh, or a
LineMesh): position p = i * P + k holds radial / z index j = k * m + i, with
P = 64 w chunk-threads of m rows and npad = P * m >= n. The helpers
``to_positions`` / ``from_positions`` convert between that layout and the
reference's natural order.
"""

import ctypes
import dataclasses
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from . import potentials
from . import units as u
from .exceptions import DeviceError

KIND_SH = 0
KIND_LINE = 1
KIND_CYL = 2
FLAG_NONFINITE = _lib.FLAG_NONFINITE
FLAG_NORM_GROWTH = _lib.FLAG_NORM_GROWTH

SH_NTAB = 6
SH_RQ = 4  # row of the interaction's radial factor
LINE_NTAB = 5


# ----------------------------------------------------------------- layout


# Measured SH layouts by radial size class (ionization_amd.autotune): consulted
# by choose_layout for kind "sh" when no rows are forced.
_SH_TUNED: dict = {}


def sh_size_class(n: int) -> int:
    """Radial sizes share a tuned layout in classes of 128 points (the class is its upper end)."""
    return 128 * -(-int(n) // 128)


def set_sh_layout_table(table) -> None:
    """Install {size class: (P, m)} measured by the autotuner (an empty table restores the rule)."""
    _SH_TUNED.clear()
    for k, (P, m) in dict(table).items():
        _SH_TUNED[int(k)] = (int(P), int(m))


def sh_layout_table() -> dict:
    return dict(_SH_TUNED)


def check_layout(n: int, P: int, m: int, kind: str = "sh") -> None:
    """Raise unless (P, m) is a layout the kernels support for a line of n points."""
    limits = {"sh": {8: 512, 16: 512}, "cylz": {8: 512, 16: 256}, "line": {8: 512, 16: 512}}[kind]
    if m < 2 or m > 16 or m % 2 or P < 64 or P % 64 or P * m < n or P > (limits[8] if m <= 8 else limits[16]):
        raise DeviceError(f"unsupported {kind} layout P={P}, m={m} for a line of {n} points")


def choose_layout(n: int, kind: str = "sh", rows: Optional[int] = None) -> Tuple[int, int]:
    """
    Pick (P, m) for a line of n points: P a multiple of 64, m even, P m >= n,
    within the kernel limits (P <= 512, m <= 16: lines up to 8192 points; the
    cylinder z-sweep: m <= 8 -> P <= 512, m <= 16 -> P <= 256).
    SH radial lines: P a power of two where one fits (on the 256-simulation
    scan 3-wave workgroups of P = 192 made the R-phase slower: 1.65 -> 1.59
    ms/step without them), then least padding, then the most rows per thread
    (fewer, longer chunks: a shorter reduced system and more work per thread);
    a table measured by the autotuner takes precedence when installed.
    Line and cylindrical z-lines: least padding, then m = 8, then the most rows.
    """
    if kind == "sh" and rows is None and _SH_TUNED:
        tuned = _SH_TUNED.get(sh_size_class(n))
        if tuned is not None and tuned[0] * tuned[1] >= n:
            return tuned
    limits = {"sh": {8: 512, 16: 512}, "cylz": {8: 512, 16: 256}, "line": {8: 512, 16: 512}}[kind]
    cands = [rows] if rows is not None else [16, 14, 12, 10, 8, 6, 4, 2]
    best = None
    for m in cands:
        if m < 2 or m % 2:
            continue
        cap = limits[8] if m <= 8 else limits[16]
        P = 64 * -(-n // (64 * m))
        if kind == "sh":
            p2 = 64
            while p2 < P:
                p2 *= 2
            if p2 <= cap:
                P = p2
        if P > cap:
            continue
        key = (P * m - n, -m) if kind == "sh" else (P * m - n, 0 if m == 8 else 1, -m)
        if best is None or key < best[0]:
            best = (key, P, m)
    if best is None:
        raise DeviceError(f"no device layout for a line of {n} points")
    return best[1], best[2]


def position_index(n: int, P: int, m: int) -> np.ndarray:
    """j for every position p (j >= n marks padding)."""
    p = np.arange(P * m)
    return (p % P) * m + p // P


def to_positions(a: np.ndarray, P: int, m: int) -> np.ndarray:
    """Natural order (last axis, length n) -> position order (length P m, zero padded)."""
    n = a.shape[-1]
    j = position_index(n, P, m)
    out = np.zeros(a.shape[:-1] + (P * m,), dtype=a.dtype)
    valid = j < n
    out[..., valid] = a[..., j[valid]]
    return out


def from_positions(a: np.ndarray, n: int, P: int, m: int) -> np.ndarray:
    """Position order -> natural order of length n."""
    j = position_index(n, P, m)
    out = np.zeros(a.shape[:-1] + (n,), dtype=a.dtype)
    valid = j < n
    out[..., j[valid]] = a[..., valid]
    return out


def data_count(n_steps: int, store_every: int) -> int:
    if store_every >= 1:
        return n_steps // store_every + 1 + (1 if n_steps % store_every else 0)
    return 2


def data_indices(n_steps: int, store_every: int) -> np.ndarray:
    """Time indices at which data is stored (sims.py:84-95)."""
    idx = np.arange(n_steps + 1)
    mask = (idx == 0) | (idx == n_steps)
    if store_every >= 1:
        mask |= idx % store_every == 0
    return idx[mask]


# ----------------------------------------------------------------- compiled simulations


@dataclass
class CompiledSH:
    """One SphericalHarmonicMesh simulation in device table form."""

    r_points: int
    l_bound: int
    P: int
    m: int
    n_steps: int
    store_every: int
    lb: int
    delta_r: float
    corr0: float
    centrifugal: float  # tau0 hbar^2 / (2 mu); U_j = centrifugal / r_j^2
    tables: np.ndarray  # (6, npad); (7, npad) in the velocity gauge
    fs: np.ndarray  # (N,)
    ts: np.ndarray  # (N,)
    g0: np.ndarray  # (L, npad) complex128, position order
    basis: np.ndarray  # (nb, npad) float64, position order
    basis_l: np.ndarray  # (nb,) int32
    norm_by_l: bool = False  # also reduce the norm in every l at data times
    gauge: str = "length"  # "length" or "velocity"
    method: str = "sio"  # "sio" (SplitInteractionOperator) or "adi" (AlternatingDirectionImplicit)

    @property
    def npad(self):
        return self.P * self.m

    @property
    def mesh_points(self):
        return self.r_points * self.l_bound

    @property
    def footprint_bytes(self):
        """Device bytes this simulation holds (two g buffers, tables, snapshots, basis, data)."""
        nd = data_count(self.n_steps, self.store_every)
        return int(
            2 * 16 * self.l_bound * self.npad
            + 8 * len(self.tables) * self.npad
            + 16 * self.n_steps
            + 16 * self.lb * self.npad
            + 8 * len(self.basis_l) * self.npad
            + nd * (16 + 16 * max(1, len(self.basis_l)))
        )


def sh_radial_tables(*, r_bound, r_points, potential, mask, time_step, test_charge=u.electron_charge,
                     hydrogen_correction=True, P, m):
    """
    Radial tables of SphericalHarmonicLengthGaugeOperators in tau0 units,
    tau0 = time_step / (2 hbar), position order:
      T_j = tau0 (beta_j (-2 P_r) + V(r_j))               (mesh_operators.py:898-928, 244-269)
      U_j = tau0 hbar^2 / (2 mu r_j^2)                     (centrifugal factor per l(l+1))
      A_j = tau0 P_r alpha_j (j < n-1), stored as A_{j-1} and A_j
      RQ_j = tau0 (-q) r_j                                 (mesh_operators.py:988-1006)
      MASK_j = mask(r_j)                                   (masks.py:76-89)
    with P_r = -hbar^2 / (2 mu dr^2), the reduced mass mu regardless of test_mass
    (mesh_operators.py:892-894). Returns (tables, corr0, delta_r, r, centrifugal) with
    centrifugal = tau0 hbar^2 / (2 mu), from which the R-phase recomputes U_j.
    """
    r = np.linspace(0, r_bound, r_points)
    delta_r = r[1] - r[0]
    r = r + delta_r / 2
    mu = u.electron_mass_reduced
    tau0 = time_step / (2 * u.hbar)
    pr = -(u.hbar ** 2) / (2 * mu * (delta_r ** 2))
    j = np.arange(r_points, dtype=np.float64)
    xb = 2 * (j ** 2) + (2 * j)
    beta = (xb + 1) / (xb + 0.5)
    xa = (j ** 2) + (2 * j)
    alpha = (xa + 1) / (xa + 0.75)
    v = np.asarray(potential(r), dtype=np.float64) * np.ones_like(r)
    T = tau0 * (beta * (-2 * pr) + v)
    centrifugal = tau0 * ((u.hbar ** 2) / (2 * mu))
    U = centrifugal / (r ** 2)
    A = tau0 * pr * alpha
    A[-1] = 0.0
    AP = np.concatenate([[0.0], A[:-1]])
    RQ = tau0 * (-test_charge) * r
    MK = np.asarray(mask(r), dtype=np.float64) * np.ones_like(r) if mask is not None else np.ones_like(r)
    corr0 = 0.0
    if hydrogen_correction:
        dr = delta_r / u.bohr_radius
        corr0 = tau0 * (-2 * pr) * (dr * (1 + dr) / 8)
    tables = to_positions(np.stack([T, U, AP, A, RQ, MK]), P, m)
    return tables, corr0, delta_r, r, centrifugal


def vector_potential_table(field: Callable, times) -> np.ndarray:
    """
    A_n for every step n: the velocity gauge's step n takes
    get_vector_potential_amplitude_numeric(sim.times_to_current) with
    time_index already advanced, i.e. -integral of E over times[0 .. n + 1]
    (mesh_operators.py:1182-1186, sims.py:195-196, 317-321), by the rule of
    ``potentials.simps_cumulative``.
    """
    t = np.asarray(times, dtype=np.float64)
    return -potentials.simps_cumulative(np.asarray(field(t), dtype=np.float64) * np.ones_like(t), t)[1:]


def sh_velocity_rows(*, r_bound, r_points, time_step, test_charge, test_mass, P, m):
    """
    The velocity gauge's two radial factor rows in tau0 units, position order
    (mesh_operators.py:1143-1180):
      h1 (l coupling at fixed r):     tau0 hbar (q / m) / r_j          (times c_l (l + 1) on the device)
      h2 (l and r coupling together): tau0 hbar q / (2 m dr) alpha_j  (times c_l), zero for j >= n - 1
    With these, the rotation angles of split_h1 / split_h2 are -i tau (i ...) A = tau A (...)
    (mesh_operators.py:1201-1215, 1252-1262).
    """
    r = np.linspace(0, r_bound, r_points)
    delta_r = r[1] - r[0]
    r = r + delta_r / 2
    tau0 = time_step / (2 * u.hbar)
    h1 = tau0 * u.hbar * (test_charge / test_mass) / r
    j = np.arange(r_points, dtype=np.float64)
    xa = (j ** 2) + (2 * j)
    alpha = (xa + 1) / (xa + 0.75)
    h2 = tau0 * u.hbar * test_charge / (test_mass * 2 * delta_r) * alpha
    h2[-1] = 0.0
    return to_positions(np.stack([h1, h2]), P, m)


def compile_sh(*, r_bound, r_points, l_bound, potential, field: Callable, mask, times, time_step,
               g0: np.ndarray, basis: Sequence[Tuple[np.ndarray, int]] = (), store_every=1,
               test_charge=u.electron_charge, hydrogen_correction=True, rows=None,
               norm_by_l=False, layout: Optional[Tuple[int, int]] = None, gauge: str = "length",
               test_mass=u.electron_mass_reduced, method: str = "sio") -> CompiledSH:
    """
    Compile a SphericalHarmonicMesh + SplitInteractionOperator simulation.
    ``times`` is the reference time grid (sims.py:198-205); step n uses
    tau_n = (t_{n+1} - t_n) / 2 hbar. Length gauge: the field at
    t_{n+1} + time_step / 2 (mesh_operators.py:1011-1013). Velocity gauge
    (``gauge="velocity"``, SphericalHarmonicVelocityGaugeOperators): the
    vector potential of ``vector_potential_table`` and ``test_mass`` in the
    interaction (mesh_operators.py:1143-1186). ``method="adi"``:
    AlternatingDirectionImplicit with the total Hamiltonian [H0 along r, E H_int
    along l] (evolution_methods.py:46-77, mesh_operators.py:1020-1035), length
    gauge, the field at t_{n+1} + time_step / 2. ``layout`` = (P, m) forces the
    device layout (the autotuner measures candidates this way).
    """
    if gauge not in ("length", "velocity"):
        raise DeviceError(f"unknown gauge {gauge!r}")
    if method not in ("sio", "adi") or (method == "adi" and gauge != "length"):
        raise DeviceError(f"unsupported SH method {method!r} with the {gauge} gauge")
    if gauge == "velocity" and (r_points < 3 or l_bound < 2):
        raise DeviceError("the velocity gauge needs r_points >= 3 and l_bound >= 2")
    if layout is not None:
        P, m = int(layout[0]), int(layout[1])
        check_layout(r_points, P, m, "sh")
    else:
        P, m = choose_layout(r_points, "sh", rows)
    tables, corr0, delta_r, _, centrifugal = sh_radial_tables(
        r_bound=r_bound, r_points=r_points, potential=potential, mask=mask, time_step=time_step,
        test_charge=test_charge, hydrogen_correction=hydrogen_correction, P=P, m=m,
    )
    times = np.asarray(times, dtype=np.float64)
    steps = np.diff(times)
    ts = steps / time_step
    if gauge == "velocity":
        vrows = sh_velocity_rows(r_bound=r_bound, r_points=r_points, time_step=time_step, test_charge=test_charge,
                                 test_mass=test_mass, P=P, m=m)
        tables = np.concatenate([tables, vrows[1:]])
        tables[SH_RQ] = vrows[0]
        fs = vector_potential_table(field, times) * ts
    else:
        fs = np.asarray(field(times[1:] + time_step / 2), dtype=np.float64) * np.ones_like(ts) * ts
    g0 = np.asarray(g0, dtype=np.complex128).reshape(l_bound, r_points)
    lb = 1 + max([l for _, l in basis], default=0)
    bmat = np.zeros((len(basis), r_points), dtype=np.float64)
    bl = np.zeros(len(basis), dtype=np.int32)
    for i, (vec, l) in enumerate(basis):
        if np.iscomplexobj(vec) and np.any(np.imag(vec) != 0):
            raise DeviceError("SH basis states must be real radial functions")
        bmat[i] = np.real(vec)
        bl[i] = l
    return CompiledSH(
        r_points=r_points, l_bound=l_bound, P=P, m=m, n_steps=len(steps), store_every=int(store_every),
        lb=min(lb, l_bound), delta_r=delta_r, corr0=corr0, centrifugal=float(centrifugal),
        tables=np.ascontiguousarray(tables),
        fs=np.ascontiguousarray(fs), ts=np.ascontiguousarray(ts),
        g0=np.ascontiguousarray(to_positions(g0, P, m)),
        basis=np.ascontiguousarray(to_positions(bmat, P, m)), basis_l=bl, norm_by_l=bool(norm_by_l),
        gauge=gauge, method=method,
    )


def relayout_sh(c: CompiledSH, layout: Optional[Tuple[int, int]] = None) -> CompiledSH:
    """The same simulation under another device layout (the tuned or rule one when None):
    position-ordered arrays are permuted, nothing is recomputed."""
    P, m = layout if layout is not None else choose_layout(c.r_points, "sh")
    check_layout(c.r_points, P, m, "sh")
    if (P, m) == (c.P, c.m):
        return c
    n = c.r_points
    move = lambda a: np.ascontiguousarray(to_positions(from_positions(a, n, c.P, c.m), P, m))
    return dataclasses.replace(c, P=P, m=m, tables=move(c.tables), g0=move(c.g0), basis=move(c.basis))


@dataclass
class CompiledLine:
    """One LineMesh simulation in device table form."""

    z_points: int
    P: int
    m: int
    n_steps: int
    store_every: int
    delta_z: float
    renormalize: bool
    offdiag: float  # tau0 P_z
    zq_first: float  # tau0 (-q) z_0
    zq_step: float  # tau0 (-q) dz
    tables: np.ndarray  # (5, npad)
    wt: np.ndarray  # (N,) complex: i tau_n / tau0
    fs: np.ndarray  # (N,)
    g0: np.ndarray  # (npad,) complex
    basis: np.ndarray  # (nb, npad) complex
    method: int = 0  # 0: ADI; 1: split operator, length gauge; 2: split operator, velocity gauge
    vel_factor: float = 0.0  # method 2: tau0 hbar (q / m) / (2 dz)

    @property
    def npad(self):
        return self.P * self.m

    @property
    def mesh_points(self):
        return self.z_points

    @property
    def footprint_bytes(self):
        nd = data_count(self.n_steps, self.store_every)
        return int(16 * self.npad * (1 + len(self.basis)) + 8 * LINE_NTAB * self.npad + 24 * self.n_steps
                   + nd * 16 * (1 + max(1, len(self.basis))))


def compile_line(*, z_bound, z_points, potential, field: Callable, mask, times, g0,
                 basis: Sequence[np.ndarray] = (), store_every=1, test_mass=u.electron_mass_reduced,
                 test_charge=u.electron_charge, time_steps=None, renormalize=False, rows=None,
                 method: str = "adi", gauge: str = "length") -> CompiledLine:
    """
    Compile a LineMesh simulation (meshes.py:263-325). ``method="adi"``:
    LineLengthGaugeOperators + AlternatingDirectionImplicit, H = K + V(z) -
    z q E(t_{n+1}) (mesh_operators.py:304-327). ``method="sio"``:
    SplitInteractionOperator with the CN step in K + V and the interaction
    split around it, ``gauge="length"`` (the phase exp(-i tau V_int),
    mesh_operators.py:329-342) or ``gauge="velocity"`` (z-pair rotations by
    tau A hbar (q / m) / (2 dz), A from ``vector_potential_table``,
    mesh_operators.py:352-427). tau0 = (t_1 - t_0) / 2 hbar. ``time_steps``
    (complex, ADI only) overrides the real steps of ``times`` (imaginary-time
    relaxation); the mask is evaluated at the signed z (meshes.py:318-320).
    """
    if method not in ("adi", "sio") or gauge not in ("length", "velocity"):
        raise DeviceError(f"unknown LineMesh method {method!r} / gauge {gauge!r}")
    if method == "adi" and gauge != "length":
        raise DeviceError("the LineMesh ADI step implements the length gauge")
    if method == "sio" and (time_steps is not None or renormalize):
        raise DeviceError("the LineMesh split operator takes real time steps without renormalisation")
    P, m = choose_layout(z_points, "line", rows)
    z = np.linspace(-z_bound, z_bound, z_points)
    delta_z = np.abs(z[1] - z[0])
    times = np.asarray(times, dtype=np.float64)
    steps = np.diff(times).astype(np.complex128) if time_steps is None else np.asarray(time_steps, np.complex128)
    tau0 = abs(steps[0]) / (2 * u.hbar)
    pz = -(u.hbar ** 2) / (2 * test_mass * (delta_z ** 2))
    v = np.asarray(potential(z), dtype=np.float64) * np.ones_like(z)
    H = tau0 * ((-2 * pz) + v)
    AN = np.full(z_points, tau0 * pz)
    AN[-1] = 0.0
    AP = np.full(z_points, tau0 * pz)
    AP[0] = 0.0
    ZQ = tau0 * (-z * test_charge)
    MK = np.asarray(mask(z), dtype=np.float64) * np.ones_like(z) if mask is not None else np.ones_like(z)
    tables = to_positions(np.stack([H, AP, AN, ZQ, MK]), P, m)
    wt = 1j * (steps / (2 * u.hbar)) / tau0
    code, vel_factor = 0, 0.0
    if method == "sio" and gauge == "velocity":
        code = 2
        vel_factor = float(tau0 * u.hbar * (test_charge / test_mass) / (2 * delta_z))
        fs = vector_potential_table(field, times)
    else:
        code = 1 if method == "sio" else 0
        fs = np.asarray(field(times[1:]), dtype=np.float64) * np.ones(len(steps))
    bmat = np.array([np.asarray(b, dtype=np.complex128) for b in basis]).reshape(len(basis), z_points)
    return CompiledLine(
        z_points=z_points, P=P, m=m, n_steps=len(steps), store_every=int(store_every), delta_z=delta_z,
        renormalize=bool(renormalize), offdiag=float(tau0 * pz), zq_first=float(ZQ[0]),
        zq_step=float(tau0 * (-test_charge) * (z[1] - z[0])),
        tables=np.ascontiguousarray(tables), wt=np.ascontiguousarray(wt),
        fs=np.ascontiguousarray(fs), g0=np.ascontiguousarray(to_positions(np.asarray(g0, np.complex128), P, m)),
        basis=np.ascontiguousarray(to_positions(bmat, P, m)), method=code, vel_factor=vel_factor,
    )


# rho-sweep shapes the library instantiates: (rows, lines per workgroup) -> most chunks
# per line (threads per workgroup / lines), in order of preference
CYL_RHO_SHAPES = {(8, 4): 128, (4, 4): 256, (4, 8): 128, (16, 4): 128, (8, 8): 128}


def choose_cyl_layout(z_points: int, rho_points: int, z_rows: Optional[int] = None,
                      rho_shape: Optional[Tuple[int, int]] = None) -> Tuple[int, int, int, int, int]:
    """
    (Pz, Mz, Cr, Mr, lanes) for a cylindrical mesh: the z-sweep layout as for an
    SH radial line, and a rho-sweep shape of Cr chunks of Mr rows with `lanes`
    z-lines per workgroup (Cr * lanes a multiple of 64, Cr within the shape's
    limit); least rho padding first, then the order of CYL_RHO_SHAPES ((8, 4)
    ahead of (4, 4): a shorter reduced system, 1.02 -> 0.99 ms per rho-sweep of
    the 32-simulation 2000 x 1000 workload).
    """
    Pz, Mz = choose_layout(z_points, "cylz", z_rows)
    best = None
    shapes = [tuple(rho_shape)] if rho_shape is not None else list(CYL_RHO_SHAPES)
    for pref, (mr, lanes) in enumerate(shapes):
        if (mr, lanes) not in CYL_RHO_SHAPES:
            raise DeviceError(f"unsupported rho-sweep shape {(mr, lanes)}")
        step = 64 // lanes
        cr = -(-rho_points // mr)
        cr = -(-cr // step) * step
        if cr > CYL_RHO_SHAPES[(mr, lanes)]:
            continue
        key = (cr * mr - rho_points, pref)
        if best is None or key < best[0]:
            best = (key, cr, mr, lanes)
    if best is None:
        raise DeviceError(f"no device rho layout for {rho_points} rho points")
    return Pz, Mz, best[1], best[2], best[3]


def to_cyl_positions(g: np.ndarray, Pz: int, Mz: int, nrpad: int) -> np.ndarray:
    """
    (..., nz, nr) natural order -> the device order, zero padding: rho rows r < nrpad
    (even) by z positions p < Pz Mz (position p holds z index (p % Pz) Mz + p // Pz), with
    the two rows of each pair (2 q, 2 q + 1) interleaved position by position: element
    (r, p) at flat index ((r // 2) npz + p) 2 + r % 2 of the trailing (nrpad, npz) block.
    """
    g = np.asarray(g)
    nz, nr = g.shape[-2:]
    if nrpad % 2:
        raise DeviceError("the padded rho rows of a cylindrical mesh must be even")
    npz = Pz * Mz
    jz = position_index(nz, Pz, Mz)
    live = jz < nz
    rows = np.zeros(g.shape[:-2] + (nrpad, npz), dtype=g.dtype)
    rows[..., :nr, live] = np.swapaxes(g[..., jz[live], :], -1, -2)
    pairs = rows.reshape(g.shape[:-2] + (nrpad // 2, 2, npz))
    return np.ascontiguousarray(np.swapaxes(pairs, -1, -2)).reshape(g.shape[:-2] + (nrpad, npz))


def from_cyl_positions(a: np.ndarray, nz: int, nr: int, Pz: int, Mz: int) -> np.ndarray:
    """Inverse of to_cyl_positions."""
    a = np.asarray(a)
    nrpad, npz = a.shape[-2:]
    rows = np.swapaxes(a.reshape(a.shape[:-2] + (nrpad // 2, npz, 2)), -1, -2).reshape(a.shape[:-2] + (nrpad, npz))
    jz = position_index(nz, Pz, Mz)
    live = jz < nz
    out = np.empty(a.shape[:-2] + (nz, nr), dtype=a.dtype)
    out[..., jz[live], :] = np.swapaxes(rows[..., :nr, live], -1, -2)
    return out


@dataclass
class CompiledCyl:
    """One CylindricalSliceMesh simulation in device table form."""

    z_points: int
    rho_points: int
    Pz: int
    Mz: int
    Cr: int
    Mr: int
    lanes: int
    n_steps: int
    store_every: int
    z_prefactor: float
    rho_prefactor: float
    coulomb: float
    test_charge: float
    delta: float
    mask: Optional[Tuple[float, float, float]]
    z_first: float
    z_step: float
    rho_first: float
    rho_step: float
    zt: np.ndarray  # (npz,)
    rt: np.ndarray  # (nrpad,)
    roff: np.ndarray  # (nrpad,)
    fs: np.ndarray  # (N,)
    ts: np.ndarray  # (N,)
    g0: np.ndarray  # (nrpad, npz) complex
    basis: np.ndarray  # (nb, nrpad, npz) complex

    @property
    def npz(self):
        return self.Pz * self.Mz

    @property
    def nrpad(self):
        return self.Cr * self.Mr

    @property
    def mesh_points(self):
        return self.z_points * self.rho_points

    @property
    def footprint_bytes(self):
        nd = data_count(self.n_steps, self.store_every)
        pts = self.npz * self.nrpad
        nb = len(self.basis)
        return int(16 * pts * (1 + nb) + 8 * (self.npz + 2 * self.nrpad) + 16 * self.n_steps
                   + 8 * (self.npz // self.lanes) * (1 + 2 * nb) + nd * 16 * (1 + max(1, nb)) + 256 * 12)


def compile_cyl(*, z_bound, rho_bound, z_points, rho_points, field: Callable, times, g0,
                basis: Sequence[np.ndarray] = (), store_every=1, charge=u.proton_charge,
                test_mass=u.electron_mass_reduced, test_charge=u.electron_charge, mask=None,
                z_rows=None, rho_shape=None) -> CompiledCyl:
    """
    Compile a CylindricalSliceMesh + CylindricalSliceLengthGaugeOperators + ADI
    simulation (meshes.py:509-607, mesh_operators.py:550-631,
    evolution_methods.py:45-77) with the Coulomb potential of a point charge
    ``charge`` (potentials/static.py:12-38). ``mask`` is None or the
    (inner_radius, outer_radius, smoothness) of a RadialCosineMask
    (potentials/masks.py:38-89). g0 and basis are (z_points, rho_points).
    """
    Pz, Mz, Cr, Mr, lanes = choose_cyl_layout(z_points, rho_points, z_rows, rho_shape)
    nrpad = Cr * Mr
    z = np.linspace(-z_bound, z_bound, z_points)
    rho = np.delete(np.linspace(0, rho_bound, rho_points + 1), 0)
    delta_z = z[1] - z[0]
    delta_rho = rho[1] - rho[0]
    rho = rho - delta_rho / 2
    zp = -(u.hbar ** 2) / (2 * test_mass * (delta_z ** 2))
    rp = -(u.hbar ** 2) / (2 * test_mass * (delta_rho ** 2))
    jz = position_index(z_points, Pz, Mz)
    zt = np.where(jz < z_points, z[np.minimum(jz, z_points - 1)], 0.0)
    rt = np.zeros(nrpad)
    rt[:rho_points] = rho
    j = np.arange(1, rho_points, dtype=np.float64)  # c(j) couples rho indices j - 1 and j
    roff = np.zeros(nrpad)
    roff[: rho_points - 1] = (j / np.sqrt((j ** 2) - 0.25)) * rp
    times = np.asarray(times, dtype=np.float64)
    steps = np.diff(times)
    fs = np.asarray(field(times[1:]), dtype=np.float64) * np.ones(len(steps))
    ts = steps / (2 * u.hbar)
    bmat = np.array([np.asarray(b, dtype=np.complex128) for b in basis]).reshape(len(basis), z_points, rho_points)
    return CompiledCyl(
        z_points=int(z_points), rho_points=int(rho_points), Pz=Pz, Mz=Mz, Cr=Cr, Mr=Mr, lanes=lanes,
        n_steps=len(steps), store_every=int(store_every), z_prefactor=float(zp), rho_prefactor=float(rp),
        coulomb=float(u.coulomb_constant * charge * test_charge), test_charge=float(test_charge),
        delta=float(delta_z * delta_rho), mask=None if mask is None else tuple(float(v) for v in mask),
        z_first=float(z[0]), z_step=float(delta_z), rho_first=float(rho[0]), rho_step=float(delta_rho),
        zt=np.ascontiguousarray(zt), rt=rt, roff=roff, fs=np.ascontiguousarray(fs), ts=np.ascontiguousarray(ts),
        g0=np.ascontiguousarray(to_cyl_positions(np.asarray(g0, np.complex128), Pz, Mz, nrpad)),
        basis=np.ascontiguousarray(to_cyl_positions(bmat, Pz, Mz, nrpad)),
    )


# ----------------------------------------------------------------- context


@dataclass
class ShResult:
    g: np.ndarray  # (L, n) natural order
    norms: np.ndarray
    lnorms: np.ndarray
    ips: np.ndarray  # (n_data, nb)
    norm_by_l: Optional[np.ndarray] = None  # (n_data, L) when the simulation reduces it


class DeviceContext:
    """One tdse_ctx: a HIP stream and a device memory pool holding many simulations."""

    def __init__(self, device: int = 0, pool_bytes: int = 4 << 30):
        self._lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(self._lib.tdse_ctx_create(int(device), ctypes.c_size_t(int(pool_bytes)), ctypes.byref(handle)))
        self._ctx = handle
        self.pool_bytes = int(pool_bytes)
        self._sh = {}
        self._line = {}
        self._cyl = {}

    # -- lifetime
    def close(self):
        if self._ctx:
            self._lib.tdse_ctx_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        _lib.check(rc, self._ctx)

    def sync(self):
        self._check(self._lib.tdse_sync(self._ctx))

    def pool_usage(self):
        used, total = ctypes.c_size_t(), ctypes.c_size_t()
        self._check(self._lib.tdse_pool_usage(self._ctx, ctypes.byref(used), ctypes.byref(total)))
        return used.value, total.value

    def copy_probe(self, n_bytes, launches=3):
        """Average ms of a streaming copy of n_bytes inside the pool (needs 2 n_bytes free)."""
        ms = ctypes.c_double()
        self._check(self._lib.tdse_copy_probe(self._ctx, int(n_bytes), int(launches), ctypes.byref(ms)))
        return ms.value

    # -- op level
    def tdma(self, sub, diag, sup, d):
        """x = T^-1 d (cy.tdma semantics); sub[i] couples x[i-1], sup[i] couples x[i+1]."""
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (sub, diag, sup, d)]
        n = len(arrs[3])
        x = np.empty(n, dtype=np.complex128)
        self._check(self._lib.tdse_tdma(self._ctx, *[_lib.dptr(a) for a in arrs], _lib.dptr(x), n))
        return x

    def tdma_batched(self, sub, diag, sup, d, offsets, lengths):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (sub, diag, sup, d)]
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        ln = np.ascontiguousarray(lengths, dtype=np.int64)
        x = np.zeros(len(arrs[3]), dtype=np.complex128)
        self._check(self._lib.tdse_tdma_batched(
            self._ctx, *[_lib.dptr(a) for a in arrs], _lib.dptr(x),
            off.ctypes.data_as(_lib.c_int64_p), ln.ctypes.data_as(_lib.c_int64_p), len(off)))
        return x

    # -- simulations
    def add_sh(self, c: CompiledSH) -> int:
        desc = _lib.ShDesc(c.r_points, c.l_bound, c.P, c.m, c.n_steps, c.store_every, len(c.basis_l), c.lb,
                           c.delta_r, c.corr0, c.centrifugal, 1 if c.norm_by_l else 0,
                           1 if c.gauge == "velocity" else 0, 1 if c.method == "adi" else 0, 0)
        slot = ctypes.c_int32()
        basis = np.ascontiguousarray(c.basis, dtype=np.float64) if len(c.basis_l) else None
        bl = np.ascontiguousarray(c.basis_l, dtype=np.int32)
        self._check(self._lib.tdse_sh_add(
            self._ctx, ctypes.byref(desc), _lib.dptr(c.tables), _lib.dptr(c.fs), _lib.dptr(c.ts), _lib.dptr(c.g0),
            _lib.dptr(basis), bl.ctypes.data_as(_lib.c_int32_p) if len(bl) else None, ctypes.byref(slot)))
        self._sh[slot.value] = c
        return slot.value

    def add_line(self, c: CompiledLine) -> int:
        desc = _lib.LineDesc(c.z_points, c.P, c.m, c.n_steps, c.store_every, len(c.basis), int(c.renormalize),
                             int(c.method), c.delta_z, c.offdiag, c.zq_first, c.zq_step, c.vel_factor)
        slot = ctypes.c_int32()
        basis = np.ascontiguousarray(c.basis, dtype=np.complex128) if len(c.basis) else None
        self._check(self._lib.tdse_line_add(
            self._ctx, ctypes.byref(desc), _lib.dptr(c.tables), _lib.dptr(c.wt), _lib.dptr(c.fs), _lib.dptr(c.g0),
            _lib.dptr(basis), ctypes.byref(slot)))
        self._line[slot.value] = c
        return slot.value

    def add_cyl(self, c: CompiledCyl) -> int:
        mk = c.mask if c.mask is not None else (0.0, 0.0, 1.0)
        desc = _lib.CylDesc(c.z_points, c.rho_points, c.Pz, c.Mz, c.Cr, c.Mr, c.lanes, c.n_steps, c.store_every,
                            len(c.basis), 0 if c.mask is None else 1, 0, c.z_prefactor, c.rho_prefactor, c.coulomb,
                            c.test_charge, c.delta, mk[0], mk[1], mk[2], c.z_first, c.z_step, c.rho_first,
                            c.rho_step)
        slot = ctypes.c_int32()
        basis = np.ascontiguousarray(c.basis, dtype=np.complex128) if len(c.basis) else None
        self._check(self._lib.tdse_cyl_add(
            self._ctx, ctypes.byref(desc), _lib.dptr(c.zt), _lib.dptr(c.rt), _lib.dptr(c.roff), _lib.dptr(c.fs),
            _lib.dptr(c.ts), _lib.dptr(c.g0), _lib.dptr(basis), ctypes.byref(slot)))
        self._cyl[slot.value] = c
        return slot.value

    def set_hold(self, kind: int, slot: int, hold: bool):
        """Cyl: end every later step at the exact reference g (see tdse_set_hold)."""
        self._check(self._lib.tdse_set_hold(self._ctx, kind, slot, 1 if hold else 0))

    def download_cyl(self, slot: int, with_g: bool = True):
        """(g (z_points, rho_points), norms, inner products) of a cylindrical simulation."""
        c = self._cyl[slot]
        _, _, nd = self.state(KIND_CYL, slot)
        g = np.empty((c.nrpad, c.npz), dtype=np.complex128) if with_g else None
        norms = np.empty(nd)
        ips = np.empty((nd, max(1, len(c.basis))), dtype=np.complex128)
        self._check(self._lib.tdse_download(self._ctx, KIND_CYL, slot, _lib.dptr(g), _lib.dptr(norms), None,
                                            _lib.dptr(ips)))
        gn = from_cyl_positions(g, c.z_points, c.rho_points, c.Pz, c.Mz) if with_g else None
        return gn, norms, ips[:, : len(c.basis)]

    def run(self, n_steps: int, use_graph: bool = True) -> float:
        """Advance every simulation by n_steps steps; returns the device time in ms."""
        self._check(self._lib.tdse_run(self._ctx, int(n_steps), 1 if use_graph else 0))
        return self.last_run_ms()

    def line_run(self, n_steps: int) -> float:
        self._check(self._lib.tdse_line_run(self._ctx, int(n_steps)))
        return self.last_run_ms()

    def last_run_ms(self) -> float:
        ms = ctypes.c_double()
        self._check(self._lib.tdse_last_run_ms(self._ctx, ctypes.byref(ms)))
        return ms.value

    def state(self, kind: int, slot: int):
        st, nn, nd = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._check(self._lib.tdse_sim_state(self._ctx, kind, slot, ctypes.byref(st), ctypes.byref(nn),
                                             ctypes.byref(nd)))
        return st.value, nn.value, nd.value

    def finished(self, kind: int, slot: int) -> bool:
        st, nn, _ = self.state(kind, slot)
        return st >= nn if kind == KIND_CYL else st > nn

    def download_sh(self, slot: int, with_g: bool = True) -> ShResult:
        c = self._sh[slot]
        _, _, nd = self.state(KIND_SH, slot)
        g = np.empty((c.l_bound, c.npad), dtype=np.complex128) if with_g else None
        norms = np.empty(nd)
        lnorms = np.empty(nd)
        ips = np.empty((nd, max(1, len(c.basis_l))), dtype=np.complex128)
        self._check(self._lib.tdse_download(self._ctx, KIND_SH, slot, _lib.dptr(g), _lib.dptr(norms),
                                            _lib.dptr(lnorms), _lib.dptr(ips)))
        gn = from_positions(g, c.r_points, c.P, c.m) if with_g else None
        nbl = None
        if c.norm_by_l:
            nbl = np.empty((nd, c.l_bound))
            self._check(self._lib.tdse_download_norm_by_l(self._ctx, slot, _lib.dptr(nbl)))
        return ShResult(gn, norms, lnorms, ips[:, : len(c.basis_l)], nbl)

    def download_line(self, slot: int):
        c = self._line[slot]
        _, _, nd = self.state(KIND_LINE, slot)
        g = np.empty(c.npad, dtype=np.complex128)
        norms = np.empty(nd)
        ips = np.empty((nd, max(1, len(c.basis))), dtype=np.complex128)
        self._check(self._lib.tdse_download(self._ctx, KIND_LINE, slot, _lib.dptr(g), _lib.dptr(norms), None,
                                            _lib.dptr(ips)))
        return from_positions(g, c.z_points, c.P, c.m), norms, ips[:, : len(c.basis)]

    def upload_state(self, kind: int, slot: int, g_natural: np.ndarray, step: int):
        if kind == KIND_CYL:
            c = self._cyl[slot]
            gp = np.ascontiguousarray(to_cyl_positions(np.asarray(g_natural, np.complex128), c.Pz, c.Mz, c.nrpad))
        else:
            c = self._sh[slot] if kind == KIND_SH else self._line[slot]
            gp = np.ascontiguousarray(to_positions(np.asarray(g_natural, np.complex128), c.P, c.m))
        self._check(self._lib.tdse_upload_state(self._ctx, kind, slot, _lib.dptr(gp), int(step)))

    def materialize(self, kind: int, slot: int):
        """Make the device g equal to the reference g after the current step count."""
        self._check(self._lib.tdse_materialize(self._ctx, kind, slot))

    def flags(self, kind: int, first_slot: int = 0, n_slots: Optional[int] = None) -> np.ndarray:
        """Health flags (FLAG_NONFINITE | FLAG_NORM_GROWTH) of a range of slots of one kind."""
        if n_slots is None:
            n_slots = 1
        out = np.zeros(int(n_slots), dtype=np.int32)
        self._check(self._lib.tdse_sim_flags(self._ctx, kind, int(first_slot), int(n_slots),
                                             out.ctypes.data_as(_lib.c_int32_p)))
        return out

    def remove(self, kind: int, slot: int):
        self._check(self._lib.tdse_remove(self._ctx, kind, slot))
        {KIND_SH: self._sh, KIND_LINE: self._line, KIND_CYL: self._cyl}[kind].pop(slot, None)

    def profile(self, enable: bool):
        self._check(self._lib.tdse_profile(self._ctx, 1 if enable else 0))

    def profile_read(self):
        lm, rm = ctypes.c_double(), ctypes.c_double()
        ln, rn = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._lib.tdse_profile_read(self._ctx, ctypes.byref(lm), ctypes.byref(ln), ctypes.byref(rm),
                                                ctypes.byref(rn)))
        return {"lphase_ms": lm.value, "lphase_launches": ln.value, "rphase_ms": rm.value,
                "rphase_launches": rn.value}

    def profile_kernel(self, kernel: int):
        """(ms, launches) accumulated for kernel 0 = SH L-phase, 1 = SH R-phase, 2 = Cyl z, 3 = Cyl rho,
        4 = SH velocity-gauge h2 factors, 5 = SH velocity-gauge odd l-pairs (R-phase with their h2 factors)."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        self._check(self._lib.tdse_profile_kernel(self._ctx, int(kernel), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
