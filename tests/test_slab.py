"""Slab walls (wallWidth > 0, an extension): the oracle's coefficients against closed forms, CPU.

The reference models every wall as an infinite half-space when useFresnelCoeff is on, so a
transmission carries no field (FieldCompute.cpp:250-262, 279-291); its slab branch is only reached
with width 0.  The slab here is the textbook dielectric slab in air (ITU-R P.2040): its pins are
energy conservation for lossless walls, the half-wave slab (no reflection, unit transmission), the
normal-incidence transmittance of a lossless slab, the thin-slab limit (R = 0, T = 1), and the
thick lossy limit (the half-space coefficient, no transmission).  A wrong phase thickness (the
reference branch's sqrt(n - sin^2)) or a two-pass phase on the transmitted wave fails the first
three."""
import numpy as np
import pytest

from oracle import oracle as O

C0 = 2.99792458e8
FREQ = 28e9
LAM = C0 / FREQ


def slab(theta, eps, width, sigma=0.0):
    return O.fresnel(theta, eps, sigma, FREQ, width=width, slab=True)


@pytest.mark.parametrize("eps", [2.0, 5.31, 9.5])
def test_lossless_slab_conserves_energy(eps):
    rng = np.random.default_rng(3)
    for _ in range(200):
        th = rng.uniform(0, np.pi / 2 - 1e-3)
        d = rng.uniform(1e-4, 0.5)
        c = slab(th, eps, d)
        for pol in ("TM", "TE"):
            r, t = c["r" + pol], c["t" + pol]
            assert abs(abs(r) ** 2 + abs(t) ** 2 - 1.0) < 1e-12, (th, d, pol, r, t)


@pytest.mark.parametrize("eps,theta", [(4.0, 0.0), (5.31, 0.6), (2.5, 1.2)])
def test_half_wave_slab_is_transparent(eps, theta):
    d = LAM / (2 * np.sqrt(eps - np.sin(theta) ** 2))   # phase thickness pi
    c = slab(theta, eps, d)
    for pol in ("TM", "TE"):
        assert abs(c["r" + pol]) < 1e-12
        assert abs(c["t" + pol] + 1.0) < 1e-12   # e^{-j pi}


def test_normal_incidence_transmittance():
    for eps in (2.0, 5.31, 9.5):
        n = np.sqrt(eps)
        for d in (0.001, 0.0123, 0.05, 0.2):
            kd = 2 * np.pi / LAM * n * d
            want = 1.0 / (1.0 + ((n * n - 1) / (2 * n)) ** 2 * np.sin(kd) ** 2)
            c = slab(0.0, eps, d)
            assert abs(abs(c["tTE"]) ** 2 - want) < 1e-12 and abs(abs(c["tTM"]) ** 2 - want) < 1e-12


def test_thin_and_thick_limits():
    th, eps, sigma = 0.7, 5.31, 0.4
    thin = slab(th, eps, 1e-12, sigma)
    half = O.fresnel(th, eps, sigma, FREQ, inf_wall=True)
    thick = slab(th, eps, 5.0, sigma)
    for pol in ("TM", "TE"):
        # first order in the phase thickness q = k d sqrt(eps - sin^2) (about 1.3e-9 here)
        assert abs(thin["r" + pol]) < 1e-8 and abs(thin["t" + pol] - 1.0) < 1e-8
        assert abs(thick["r" + pol] - half["r" + pol]) < 1e-9 and abs(thick["t" + pol]) < 1e-9


def test_reference_branch_differs():
    """The reference's own slab branch (never reached by it with width > 0) is not this slab."""
    c_ref = O.fresnel(0.3, 5.31, 0.0, FREQ, width=0.01, inf_wall=False)
    c_new = slab(0.3, 5.31, 0.01)
    assert abs(c_ref["tTE"] - c_new["tTE"]) > 1e-3
