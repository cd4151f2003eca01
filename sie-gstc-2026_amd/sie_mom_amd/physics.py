"""Solver options and the float-rounded kernel parameters derived from them.

Mirrors the reference option structs and their derived quantities:
GstcSieOptions / make_physical_parameters (include/sie_mom/solvers/gstc_sie.hpp:32-100),
PecOptions (include/sie_mom/solvers/pec_solver.hpp:17-28, derivation pec_solver.cu:291-300),
DielectricOptions (include/sie_mom/solvers/dielectric/dielectric_solver.hpp:29-42,
derivation dielectric_solver.cu:132-160) and prepare_source
(src/sie_mom/excitation/excitation.cu:110-137).
"""
from dataclasses import dataclass, field
import math

import numpy as np

from . import _lib

CODATA2022_EPS0 = 8.8541878188e-12
CODATA2022_MU0 = 1.25663706127e-6
CODATA2018_EPS0 = 8.8541878128e-12
CODATA2018_MU0 = 1.25663706212e-6


@dataclass
class Medium:
    """IsotropicMedium (include/sie_mom/fields.hpp:13-16)."""
    relative_permittivity: float = 1.0
    relative_permeability: float = 1.0


@dataclass
class GstcOptions:
    frequency_hz: float = 0.0
    interior: Medium = field(default_factory=Medium)
    exterior: Medium = field(default_factory=Medium)
    quadrature_rule: int = 17
    singularity_threshold_lambda0: float = 0.1
    tensors_are_uniform: bool = False
    compute_residual: bool = True
    basis: str = "face"  # or "spherical"
    vacuum_permittivity: float = CODATA2022_EPS0
    vacuum_permeability: float = CODATA2022_MU0


@dataclass
class PecOptions:
    frequency_hz: float = 0.0
    exterior: Medium = field(default_factory=Medium)
    quadrature_rule: int = 7
    singularity_threshold_lambda0: float = 0.1
    compute_residual: bool = True
    vacuum_permittivity: float = CODATA2018_EPS0
    vacuum_permeability: float = CODATA2018_MU0


@dataclass
class DielectricOptions:
    frequency_hz: float = 0.0
    interior: Medium = field(default_factory=Medium)
    exterior: Medium = field(default_factory=Medium)
    quadrature_rule: int = 17
    singularity_threshold_lambda0: float = 0.1
    compute_residual: bool = True
    vacuum_permittivity: float = CODATA2018_EPS0
    vacuum_permeability: float = CODATA2018_MU0


@dataclass
class Physical:
    omega: float
    k_interior: float
    k_exterior: float
    eta0: float
    singularity_threshold: float


def physical_parameters(o):
    """make_physical_parameters (gstc_sie.hpp:75-100); the PEC and dielectric
    drivers derive the same quantities with lambda0 = 2 pi / (omega sqrt(mu0 eps0))."""
    if not o.frequency_hz > 0:
        raise ValueError("SIE frequency must be positive")
    pi = math.acos(-1.0)
    omega = 2.0 * pi * o.frequency_hz
    interior = getattr(o, "interior", o.exterior)
    ei = interior.relative_permittivity * o.vacuum_permittivity
    mi = interior.relative_permeability * o.vacuum_permeability
    ee = o.exterior.relative_permittivity * o.vacuum_permittivity
    me = o.exterior.relative_permeability * o.vacuum_permeability
    eta0 = math.sqrt(o.vacuum_permeability / o.vacuum_permittivity)
    if isinstance(o, GstcOptions):
        c0 = 1.0 / math.sqrt(o.vacuum_permeability * o.vacuum_permittivity)
        thr = o.singularity_threshold_lambda0 * c0 / o.frequency_hz
    else:
        lam0 = 2.0 * pi / (omega * math.sqrt(o.vacuum_permeability * o.vacuum_permittivity))
        thr = o.singularity_threshold_lambda0 * lam0
    return Physical(omega, omega * math.sqrt(mi * ei), omega * math.sqrt(me * ee), eta0, thr)


def frequency_for_k0(k0, vacuum_permittivity=CODATA2022_EPS0, vacuum_permeability=CODATA2022_MU0):
    """Frequency at which the vacuum wavenumber is k0 (rad/m)."""
    c0 = 1.0 / math.sqrt(vacuum_permeability * vacuum_permittivity)
    return k0 * c0 / (2.0 * math.pi)


def calderon_params(o):
    """Kernel launch parameters cast to float as the reference casts them
    (gstc_sie.cu:281-290, pec_solver.cu:352-359, dielectric_solver.cu:335-347)."""
    ph = physical_parameters(o)
    interior = getattr(o, "interior", o.exterior)
    f32 = np.float32
    p = _lib.CalderonParams()
    p.rule = int(o.quadrature_rule)
    p.k1 = float(f32(ph.k_interior if not isinstance(o, PecOptions) else ph.k_exterior))
    p.k2 = float(f32(ph.k_exterior))
    p.mu_r1 = float(f32(interior.relative_permeability))
    p.mu_r2 = float(f32(o.exterior.relative_permeability))
    p.eps_r1 = float(f32(interior.relative_permittivity))
    p.eps_r2 = float(f32(o.exterior.relative_permittivity))
    p.mu0 = float(f32(o.vacuum_permeability))
    p.eps0 = float(f32(o.vacuum_permittivity))
    p.omega = float(f32(ph.omega))
    p.threshold = float(f32(ph.singularity_threshold))
    return p


def plane_wave(o, electric_amplitude=(1.0, 0.0, 0.0), propagation_direction=(0.0, 0.0, 1.0),
               region="exterior"):
    """prepare_source for a PlaneWaveExcitation: bind the source medium's k and
    eta, normalise the direction, reject a non-transverse amplitude."""
    reg = 2 if region in ("exterior", 2) else 1
    medium = o.exterior if reg == 2 else getattr(o, "interior", o.exterior)
    if not o.frequency_hz > 0:
        raise ValueError("source frequency must be positive and finite")
    eps = medium.relative_permittivity * o.vacuum_permittivity
    mu = medium.relative_permeability * o.vacuum_permeability
    omega = 2.0 * math.acos(-1.0) * o.frequency_hz
    k = omega * math.sqrt(mu * eps)
    eta = math.sqrt(mu / eps)
    d = np.asarray(propagation_direction, float)
    a = np.asarray(electric_amplitude, float)
    dn = math.hypot(*d)
    an = math.hypot(*a)
    if not (dn > 0 and np.isfinite(dn)):
        raise ValueError("plane-wave direction magnitude must be positive and finite")
    if not (an > 0 and np.isfinite(an)):
        raise ValueError("plane-wave amplitude magnitude must be positive and finite")
    u = d / dn
    if abs(float(np.dot(u, a / an))) > 64 * np.finfo(float).eps:
        raise ValueError("plane-wave electric amplitude must be transverse to propagation")
    w = _lib.PlaneWaveC()
    for i in range(3):
        w.E0[i] = float(a[i])
        w.khat[i] = float(u[i])
    w.k = k
    w.eta = eta
    w.region = reg
    return w


def fibonacci_directions(count):
    """count unit vectors on the golden-angle (Fibonacci) spiral of the sphere:
    z_i = 1 - (2 i + 1) / count in (-1, 1), azimuth i pi (3 - sqrt 5)."""
    i = np.arange(count)
    z = 1.0 - (2.0 * i + 1.0) / count
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    r = np.sqrt(1.0 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def fibonacci_plane_waves(o, count, region="exterior"):
    """count unit plane waves incident along the Fibonacci directions d_i
    (fibonacci_directions), polarised along theta_hat(d_i) for even i and
    phi_hat(d_i) for odd i: the many-right-hand-side excitation of a
    dielectric scatterer."""
    waves = []
    for i, d in enumerate(fibonacci_directions(count)):
        th = math.acos(max(-1.0, min(1.0, d[2])))
        ph = math.atan2(d[1], d[0])
        if i % 2 == 0:
            e = (math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph), -math.sin(th))
        else:
            e = (-math.sin(ph), math.cos(ph), 0.0)
        waves.append(plane_wave(o, e, tuple(d), region))
    return waves
