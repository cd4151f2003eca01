"""Mie series for a perfectly conducting or a dielectric sphere (closed form,
test-only).

Scattering amplitudes in the convention of Bohren & Huffman, "Absorption and
Scattering of Light by Small Particles" (1983), sections 4.4 and 4.8;
psi_n = x j_n(x), xi_n = x h_n^(1)(x), x = k a (k of the exterior medium).
Dielectric sphere of relative refractive index m = sqrt(eps_r mu_r) with
mu_r = 1 (B&H eq. 4.53 with mu_1 = mu):
  a_n = [m psi_n(mx) psi_n'(x) - psi_n(x) psi_n'(mx)] / [m psi_n(mx) xi_n'(x) - xi_n(x) psi_n'(mx)],
  b_n = [psi_n(mx) psi_n'(x) - m psi_n(x) psi_n'(mx)] / [psi_n(mx) xi_n'(x) - m xi_n(x) psi_n'(mx)];
perfect conductor (the limit of infinite m): a_n = psi_n'(x) / xi_n'(x),
b_n = psi_n(x) / xi_n(x).
For an x-polarised wave travelling along +z the bistatic radar cross section in
the E-plane (phi = 0) is 4 pi |S2|^2 / k^2 and in the H-plane (phi = 90 deg)
4 pi |S1|^2 / k^2. Magnitudes do not depend on the time convention.
"""
import numpy as np
from scipy.special import spherical_jn, spherical_yn


def coefficients(x, nmax=None, m=None):
    """(n, a_n, b_n); m = None: perfect conductor."""
    nmax = nmax or int(np.ceil(x + 4.05 * x ** (1 / 3) + 2)) + 4
    n = np.arange(1, nmax + 1)
    jn, djn = spherical_jn(n, x), spherical_jn(n, x, derivative=True)
    yn, dyn = spherical_yn(n, x), spherical_yn(n, x, derivative=True)
    hn, dhn = jn + 1j * yn, djn + 1j * dyn
    psi, dpsi = x * jn, jn + x * djn
    xi, dxi = x * hn, hn + x * dhn
    if m is None:
        return n, dpsi / dxi, psi / xi
    mx = m * x
    jm, djm = spherical_jn(n, mx), spherical_jn(n, mx, derivative=True)
    psim, dpsim = mx * jm, jm + mx * djm
    a = (m * psim * dpsi - psi * dpsim) / (m * psim * dxi - xi * dpsim)
    b = (psim * dpsi - m * psi * dpsim) / (psim * dxi - m * xi * dpsim)
    return n, a, b


def angular(n, theta):
    """pi_n = P_n^1(cos t) / sin t and tau_n = d P_n^1(cos t) / dt by the
    upward recurrences (Bohren & Huffman eq. 4.47), sign convention of B&H."""
    mu = np.cos(theta)
    N = int(n[-1])
    pi = np.zeros((N + 1,) + np.shape(theta))
    tau = np.zeros_like(pi)
    pi[1] = 1.0
    if N >= 2:
        pi[2] = 3.0 * mu
    for k in range(3, N + 1):
        pi[k] = (2 * k - 1) / (k - 1) * mu * pi[k - 1] - k / (k - 1) * pi[k - 2]
    for k in range(1, N + 1):
        tau[k] = k * mu * pi[k] - (k + 1) * pi[k - 1]
    return pi[1:], tau[1:]


def amplitudes(ka, theta, m=None):
    n, a, b = coefficients(ka, m=m)
    pi, tau = angular(n, np.asarray(theta, float))
    w = ((2 * n + 1) / (n * (n + 1)))[:, None] if np.ndim(theta) else (2 * n + 1) / (n * (n + 1))
    a_, b_ = (a[:, None], b[:, None]) if np.ndim(theta) else (a, b)
    S1 = np.sum(w * (a_ * pi + b_ * tau), axis=0)
    S2 = np.sum(w * (a_ * tau + b_ * pi), axis=0)
    return S1, S2


def bistatic_rcs(ka, theta, k=1.0, m=None):
    """(sigma_E, sigma_H): E-plane and H-plane bistatic RCS for wavenumber k."""
    S1, S2 = amplitudes(ka, theta, m=m)
    return 4 * np.pi * np.abs(S2) ** 2 / k ** 2, 4 * np.pi * np.abs(S1) ** 2 / k ** 2


def cross_sections(ka, k=1.0, m=None):
    """(sigma_sca, sigma_ext) from the coefficient sums (B&H eqs. 4.61, 4.62)."""
    n, a, b = coefficients(ka, m=m)
    s = 2 * np.pi / k ** 2
    return s * np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2)), s * np.sum((2 * n + 1) * np.real(a + b))
