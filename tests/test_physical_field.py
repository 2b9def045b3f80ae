"""Physical field transport (physicalField = 1, an extension), in the oracle, CPU, against image
theory and the textbook Fresnel coefficients.  Over a flat ground, the field of the ground-reflected
path equals the direct field of the transmitter's image (mirrored below the ground, same antenna
frame) times the reflection coefficient of the polarisation: for a vertical dipole the field lies in
the plane of incidence (p), for a horizontal dipole seen broadside it is normal to it (s).  With a
perfect conductor (Gamma_p = +1, Gamma_s = -1) that is image theory for any receiver.  The
reference's transport (through the surface frames) does not satisfy these."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_long_miss import ground_soup

FREQ = 1.8e9
E0 = 8.854187817e-12
H = 12.0
VERT = np.eye(3)
HORZ = np.array([[0.0, 0.0, -1.0], [0.0, 1.0, 0.0], [1.0, 0.0, 0.0]])   # dipole (local z) along global x


def far_scene():
    c = np.array([[5000, 5000, 5000, 5001, 5000, 5000, 5000, 5001, 5000]], np.float32)
    return O.Scene.from_soup(c.reshape(-1), np.array([0, 0, 1], np.float32))


def fields(scene, mats, tx, cs, rx, **kw):
    p = O.make_params(tx=tx, tess=8, maxR=1, maxT=0, frequency=FREQ, receptionModel=1, longMiss=1, txCoordSys=cs,
                      physicalField=1, **kw)
    r = O.trace(scene, p, rx, mats, nthreads=8)
    out = {}
    for k in range(len(r.prx)):
        out[(int(r.prx[k]), int(r.pnseg[k]))] = r.path_fields[k]
    return out


def gammas(theta, eps, sigma):
    ec = eps - 1j * sigma / (E0 * 2 * np.pi * FREQ)
    ct = np.sqrt(1 - np.sin(theta) ** 2 / ec)   # n_t cos(theta_t) / n_t
    n = np.sqrt(ec)
    gs = (np.cos(theta) - n * ct) / (np.cos(theta) + n * ct)
    gp = (n * np.cos(theta) - ct) / (n * np.cos(theta) + ct)
    return gs, gp


def check(eps, sigma, cs, rx, pol):
    mats = np.array([[FREQ, eps, 0.0, eps, 0.0, sigma]])
    g = O.Scene.from_soup(*ground_soup())
    refl = fields(g, mats, (0.0, 0.0, H), cs, rx)
    img = fields(far_scene(), mats, (0.0, 0.0, -H), cs, rx)
    for r, p in enumerate(rx):
        assert (r, 2) in refl and (r, 1) in img, r
        rho = np.hypot(p[0], p[1])
        theta = np.arctan2(rho, H + p[2])           # incidence angle at the reflection point
        gs, gp = gammas(theta, eps, sigma)
        want = (gp if pol == "p" else gs) * img[(r, 1)]
        got = refl[(r, 2)]
        assert np.linalg.norm(got - want) <= 1e-6 * np.linalg.norm(want), (r, got, want)


def test_vertical_dipole_lossy_ground_is_p():
    rx = np.array([[60.0, 0.0, 2.0], [0.0, 150.0, 8.0], [-90.0, 70.0, 1.5], [300.0, -200.0, 20.0]])
    check(5.0, 0.02, VERT, rx, "p")


def test_horizontal_dipole_broadside_lossy_ground_is_s():
    rx = np.array([[0.0, 60.0, 2.0], [0.0, -150.0, 8.0], [0.0, 400.0, 30.0]])
    check(15.0, 0.5, HORZ, rx, "s")


@pytest.mark.parametrize("cs", [VERT, HORZ], ids=["vertical", "horizontal"])
def test_perfect_conductor_is_image_theory(cs):
    rx = np.array([[60.0, 10.0, 2.0], [-40.0, 150.0, 8.0], [90.0, -70.0, 1.5], [300.0, 200.0, 20.0]])
    mats = np.array([[FREQ, 1.0, 0.0, 1.0, 0.0, 1e13]])
    refl = fields(O.Scene.from_soup(*ground_soup()), mats, (0.0, 0.0, H), cs, rx)
    img = fields(far_scene(), mats, (0.0, 0.0, -H), cs, rx)
    sign = 1.0 if cs is VERT else -1.0   # a horizontal dipole's image is reversed
    for r in range(len(rx)):
        want = sign * img[(r, 1)]
        assert np.linalg.norm(refl[(r, 2)] - want) <= 1e-5 * np.linalg.norm(want), r


def test_reference_transport_differs():
    rx = np.array([[0.0, 60.0, 2.0]])
    mats = np.array([[FREQ, 15.0, 0.0, 15.0, 0.0, 0.5]])
    g = O.Scene.from_soup(*ground_soup())
    a = fields(g, mats, (0.0, 0.0, H), HORZ, rx)
    p = O.make_params(tx=(0.0, 0.0, H), tess=7, maxR=1, maxT=0, frequency=FREQ, receptionModel=1, longMiss=1,
                      txCoordSys=HORZ)
    r = O.trace(g, p, rx, mats, nthreads=8)
    ref = [r.path_fields[k] for k in range(len(r.prx)) if r.pnseg[k] == 2][0]
    assert np.linalg.norm(ref - a[(0, 2)]) > 1e-3 * np.linalg.norm(a[(0, 2)])
