"""Tabulated transmitter pattern (antennaType 3) in the oracle, CPU.  The reference reads
TXPattern.dat (Preprocessor.cpp:3-35) into GainInfo rows (DataStructures.h:32-37) but its field
computation always uses the analytic pattern (FieldCompute.cpp:117), so no reference output covers
a table: the oracle's lookup is pinned here by closed forms.

* Bilinear interpolation reproduces exactly a function that is bilinear on every grid cell: a
  product of a linear function of theta and a triangle wave of phi with its corners on grid angles
  (periodic, so the cell across phi = 360 is covered too).  An index or weight mistake, a swapped
  axis, a missing wrap or a degree/radian slip breaks it by orders of magnitude more than the
  float32 rounding of the dB / degree rows.
* A table of the Hertzian dipole (gain 1.5 sin^2 theta, phase 0) gives the analytic Hertzian field
  at grid angles, and whole traces with that table keep the analytic traces' paths with per-path
  fields within the interpolation error."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from sbrsolver_amd import scenes

PATTERN_FILE = "/root/reference/SBRSolver_rev2/data/TXPattern.dat"


def table_from(f_theta, f_phi, n_theta, n_phi):
    """GainInfo rows (gain dB, phase deg for theta and phi) whose coefficients j sqrt(60 G) e^{j psi}
    equal f(theta_deg, phi_deg) at the grid points."""
    th = np.arange(n_theta) * 180.0 / (n_theta - 1)
    ph = np.arange(n_phi) * 360.0 / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    out = np.zeros((n_theta, n_phi, 4), np.float32)
    for h, f in enumerate((f_theta, f_phi)):
        c = np.broadcast_to(f(T, P), T.shape).astype(complex)
        out[..., 2 * h] = 10 * np.log10(np.maximum(np.abs(c) ** 2 / 60.0, 1e-300))
        out[..., 2 * h + 1] = np.degrees(np.angle(c)) - 90.0
    return out


def tri(phi_deg):
    return np.abs(np.mod(phi_deg + 180.0, 360.0) - 180.0)


F_TH = lambda t, p: (2.0 + 0.5j + (0.03 - 0.01j) * t) * (1.5 - 0.2j + (0.004 + 0.002j) * tri(p))
F_PH = lambda t, p: (0.3 - 1.1j + (-0.002 + 0.004j) * t) * (0.7 + 0.9j + (-0.003 + 0.001j) * tri(p))


def _angles(rng, n):
    th = np.concatenate([[0.0, np.pi, np.pi / 2, 1e-9, np.pi - 1e-9], rng.uniform(0, np.pi, n)])
    ph = np.concatenate([[0.0, 2 * np.pi - 1e-12, np.pi, 1e-9, 3.0], rng.uniform(0, 2 * np.pi, n)])
    return th, ph


@pytest.mark.parametrize("n_theta,n_phi", [(181, 360), (7, 8), (2, 1)])
def test_table_reproduces_cellwise_bilinear_pattern(n_theta, n_phi):
    if n_phi == 1:   # one phi column: the pattern does not depend on phi
        fth, fph = (lambda t, p: (2.0 + 0.5j + (0.03 - 0.01j) * t) * (1 + 0 * p)), (lambda t, p: (0.3 - 1.1j) + 0 * t + 0 * p)
    else:
        fth, fph = F_TH, F_PH
    O.set_pattern(table_from(fth, fph, n_theta, n_phi))
    rng = np.random.default_rng(7)
    th, ph = _angles(rng, 400)
    pt = 2.5
    scale = max(np.abs(fth(np.array([0.0, 180.0]), np.array([0.0, 180.0]))).max(),
                np.abs(fph(np.array([0.0, 180.0]), np.array([0.0, 180.0]))).max())
    for t, p in zip(th, ph):
        et, ep = O.pattern_eval(t, p, pt)
        td, pd = np.degrees(t), np.degrees(p)
        want_t, want_p = np.sqrt(pt) * fth(td, pd), np.sqrt(pt) * fph(td, pd)
        assert abs(et - want_t) < 2e-6 * np.sqrt(pt) * scale, (t, p, et, want_t)
        assert abs(ep - want_p) < 2e-6 * np.sqrt(pt) * scale, (t, p, ep, want_p)


def test_table_detects_a_half_cell_shift():
    """The check above has teeth: evaluating half a cell off in phi misses by far more than its
    tolerance."""
    O.set_pattern(table_from(F_TH, F_PH, 7, 8))
    et, _ = O.pattern_eval(np.radians(50.0), np.radians(100.0 + 22.5), 1.0)
    assert abs(et - F_TH(50.0, 100.0)) > 1e-3


def hertz_table(n_theta=181, n_phi=360):
    return table_from(lambda t, p: 1j * np.sqrt(60 * 1.5) * np.abs(np.sin(np.radians(t))) + 0 * p,
                      lambda t, p: 0 * t, n_theta, n_phi)


def test_hertzian_table_at_grid_angles():
    O.set_pattern(hertz_table())
    for td in (0.0, 1.0, 37.0, 90.0, 143.0, 179.0, 180.0):
        for pd in (0.0, 17.3, 359.9):
            et, ep = O.pattern_eval(np.radians(td), np.radians(pd), 3.0)
            want = 1j * np.sqrt(60 * 3.0 * 1.5) * np.sin(np.radians(td))
            assert abs(et - want) < 1e-6 * np.sqrt(60 * 3.0 * 1.5) and abs(ep) < 1e-100


def test_oracle_trace_with_hertzian_table_keeps_the_analytic_paths():
    cfg = scenes.config2(tess=3)
    sc = O.Scene.from_soup(*cfg["soup"])
    O.set_pattern(hertz_table())
    tr = dict(cfg["trace"], receptionModel=1, antennaType=1)
    a = O.trace(sc, O.make_params(**tr), cfg["rx"], cfg["materials"], nthreads=8)
    tr["antennaType"] = 3
    b = O.trace(sc, O.make_params(**tr), cfg["rx"], cfg["materials"], nthreads=8)
    assert len(a.prx) > 100
    assert np.array_equal(a.prx, b.prx) and np.array_equal(a.stri, b.stri) and np.array_equal(a.stype, b.stype)
    mag = np.linalg.norm(a.path_fields, axis=1)
    err = np.linalg.norm(b.path_fields - a.path_fields, axis=1)
    # bilinear interpolation of sin(theta) over 1-degree cells: at most (pi/180)^2 / 8 relative
    assert np.all(err <= 5e-5 * mag + 1e-12 * mag.max())
    assert err.max() > 0   # the table path is taken


@pytest.mark.skipif(not os.path.exists(PATTERN_FILE), reason="reference data not present")
def test_reference_pattern_file_is_a_full_grid():
    """TXPattern.dat (the reference's own data): 181 x 360 rows at 1 degree after its 4-token
    header, which a GainMap reader skips.  The file runs theta fastest; a GainMap orders by
    (theta, phi), which is the table layout."""
    rows = []
    with open(PATTERN_FILE) as f:
        for line in f:
            w = line.split()
            if len(w) == 6:
                rows.append([float(x) for x in w])
    a = np.array(rows)
    assert a.shape == (181 * 360, 6)
    a = a[np.lexsort((a[:, 1], a[:, 0]))]
    th, ph = a[:, 0].reshape(181, 360), a[:, 1].reshape(181, 360)
    assert np.array_equal(th, np.repeat(np.arange(181.0)[:, None], 360, axis=1))
    assert np.array_equal(ph, np.repeat(np.arange(360.0)[None, :], 181, axis=0))
    O.set_pattern(a[:, 2:].reshape(181, 360, 4))
    for k in (0, 181 * 360 - 1, 90 * 360 + 45):
        et, ep = O.pattern_eval(np.radians(a[k, 0]), np.radians(a[k, 1]), 1.0)
        assert abs(abs(et) - np.sqrt(60 * 10 ** (np.float32(a[k, 2]) / 10))) < 1e-6 * abs(et)
        assert abs(abs(ep) - np.sqrt(60 * 10 ** (np.float32(a[k, 4]) / 10))) < 1e-6 * abs(ep)
