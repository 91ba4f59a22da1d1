"""Test infrastructure, not product code: the Heston (1993) European price by
Fourier inversion, the closed-form checker of the Heston path stepping.

kiyosi prices Heston Europeans analytically (engines/analytic/heston.rs:515-549,
log characteristic function; Gauss-Kronrod Fourier integration) and the GPU
engine steps the same model by Monte Carlo. This module is written from the
published form, not from that file: the characteristic function of ln S_T in
the "little Heston trap" branch of Albrecher, Mayer, Schoutens and Tistaert
(2007), which stays on the principal branch of the complex logarithm, and the
Gil-Pelaez probabilities P1, P2 of Heston (1993):

    call = S e^{-qT} P1 - K e^{-rT} P2,
    P2 = 1/2 + 1/pi int_0^inf Re[e^{-iu ln K} phi(u) / (iu)] du,
    P1 = 1/2 + 1/pi int_0^inf Re[e^{-iu ln K} phi(u - i) / (iu phi(-i))] du,

with phi(-i) = E[S_T] = S e^{(r-q)T}. Puts follow from put-call parity. The
module is pinned by the reference's analytic Heston fixture
(benchmarks/quantlib/references/heston.json: 9.011278407541738) and by the
Black-Scholes limit xi -> 0 in tests/test_heston.py.
"""

import math
import warnings

import numpy as np
from scipy import integrate


def _log_cf(u, spot, rate, dividend, time, v0, theta, kappa, xi, rho):
    """ln E[exp(iu ln S_T)] for complex u."""
    iu = 1j * u
    b = kappa - rho * xi * iu
    d = np.sqrt(b * b + xi * xi * (iu + u * u))
    g = (b - d) / (b + d)
    e = np.exp(-d * time)
    c = kappa * theta / (xi * xi) * ((b - d) * time - 2.0 * np.log((1.0 - g * e) / (1.0 - g)))
    dd = (b - d) / (xi * xi) * (1.0 - e) / (1.0 - g * e)
    return iu * (math.log(spot) + (rate - dividend) * time) + c + dd * v0


def heston_call(spot, strike, rate, dividend, time, v0, theta, kappa, xi, rho, upper=400.0):
    """European call under Heston by Gil-Pelaez inversion (adaptive quadrature
    to 1e-13 absolute on [0, upper]; the integrands decay like exp(-c u))."""
    args = (spot, rate, dividend, time, v0, theta, kappa, xi, rho)
    lk = math.log(strike)
    forward = spot * math.exp((rate - dividend) * time)

    def p2_integrand(u):
        return (np.exp(-1j * u * lk + _log_cf(u, *args)) / (1j * u)).real

    def p1_integrand(u):
        return (np.exp(-1j * u * lk + _log_cf(u - 1j, *args)) / (1j * u * forward)).real

    opts = dict(limit=2000, epsabs=1e-13, epsrel=1e-13)
    with warnings.catch_warnings():
        # quad reports when roundoff stops it short of 1e-13; the result is
        # then still within about 1e-12 of the reference's value.
        warnings.simplefilter("ignore", integrate.IntegrationWarning)
        i2 = integrate.quad(p2_integrand, 1e-12, upper, **opts)[0]
        i1 = integrate.quad(p1_integrand, 1e-12, upper, **opts)[0]
    p1 = 0.5 + i1 / math.pi
    p2 = 0.5 + i2 / math.pi
    return spot * math.exp(-dividend * time) * p1 - strike * math.exp(-rate * time) * p2


def heston_put(spot, strike, rate, dividend, time, v0, theta, kappa, xi, rho, upper=400.0):
    call = heston_call(spot, strike, rate, dividend, time, v0, theta, kappa, xi, rho, upper)
    return call - spot * math.exp(-dividend * time) + strike * math.exp(-rate * time)
