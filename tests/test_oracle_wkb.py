"""Pins of the oracle's WKB ground-to-ground cut-off (ESSModeSolver.cpp:561-621).

The reference's own WKB output (Nby2D_tloss_1d.nm of the Modess examples) is not held in
the repository, so these checks come from what the arithmetic fixes: a 20-row grid worked
by hand, the closed-form tunnelling integral of a linear k_eff^2 profile, and the two
branches that bypass the stepping (c_eff smallest at the ground; the cut-off below k_min).
"""
import math

import numpy as np

from oracle import ess
from oracle.atmos import PI


class _Grid:
    """The three attributes of a grid modal_trace reads: Nz, dz and c_eff(azimuth)."""

    def __init__(self, ceff, dz):
        self.c = np.asarray(ceff, dtype=np.float64)
        self.Nz, self.dz = self.c.size, float(dz)

    def ceff(self, azimuth):
        return self.c.copy()


def _hand_grid():
    """k_eff^2 = K_i on 20 rows, omega = 340 rad/s (c_eff = 340/sqrt(K_i) m/s), dz = 46 m:
    K_0 = 1 (ground), K_i = 1.013 + 0.016 (i - 1) for i = 1..11, K_12 = 2 (the c_eff minimum),
    K_13..19 = 0.5 (the top rows: k_min = omega/c_eff[19] = sqrt(0.5))."""
    K = [1.0] + [1.013 + 0.016 * (i - 1) for i in range(1, 12)] + [2.0] + [0.5] * 7
    freq = 340.0 / (2 * PI)
    omega = 2 * PI * freq
    return _Grid(omega / np.sqrt(np.array(K)), 46.0), freq


def test_wkb_hand_worked_grid():
    """dkk = (2 - 1)/100 = 0.01; kk = 1 + 0.01 n.
    n = 0: row 0 has |kk - K_0| = 0 <= dkk: integral 46 sqrt(0) = 0.
    n = 1: |kk - K_0| = dkk, a tie that rounding decides: max k_eff^2 rounds to
           2.0000000000000004, dkk to 0.010000000000000004 and fl(1 + dkk) - 1 to
           0.010000000000000009 > dkk, so row 0 does not stop the sum; row 1
           (|1.01 - 1.013| = 0.003) does: 46 (0.1 + 0.0548) = 7.12 (4.6 with the tie the
           other way), below 10 either way.
    n = 2: kk = 1.02; row 0: 0.02 > dkk, row 1: 0.007 <= dkk, so rows 0..1 (inclusive):
           46 (sqrt(0.02) + sqrt(0.007)) = 46 (0.141421 + 0.083666) = 10.354 >= 10: stop.
    k_max = sqrt(1.02) = 1.00995049383620780.  Leaving the stopping row out would give
    46 sqrt(0.02) = 6.505 at n = 2 and stop at n = 3 (k_max = sqrt(1.03)); dz in km, a
    dropped sqrt or |.| would stop elsewhere or never."""
    grid, freq = _hand_grid()
    diag, ceffz, k_min, k_max, used = ess.modal_trace(grid, freq, 0.0, 0.0, 0.0, False)
    assert used
    assert abs(k_max - 1.00995049383620780) <= 1e-12
    assert abs(k_min - math.sqrt(0.5)) <= 1e-12
    kk, steps = ess.wkb_cutoff(ceffz, 2 * PI * freq, grid.dz)
    assert len(steps) == 3
    assert steps[2][1] == 1   # stopping row of n = 2
    assert abs(steps[2][2] - 46.0 * (math.sqrt(0.02) + math.sqrt(0.007))) <= 1e-9
    assert steps[1][1] == 1 and abs(steps[1][2] - 7.119523764523721) <= 1e-9
    # source or receiver above 1 mm, or WKB turned off: k_max = omega / min c_eff = sqrt(2)
    for zs, zr, off in ((0.0, 0.0, True), (2e-3, 0.0, False), (0.0, 2e-3, False)):
        _, _, _, km, u = ess.modal_trace(grid, freq, 0.0, zs, zr, off)
        assert not u and abs(km - math.sqrt(2.0)) <= 1e-12


def test_wkb_slow_layer_below_turning_point():
    """The hand grid with a slow row under the turning point: K_1 = 1.3 (c_eff 298 m/s), rows
    2..11 = 1.013 + 0.016 (i - 2).  n = 1 (kk = 1.01): row 0 does not stop (as above), row 1
    has |1.01 - 1.3| = 0.29 > dkk, row 2 has 0.003 <= dkk: rows 0..2,
    46 (0.1 + sqrt(0.29) + sqrt(0.003)) = 46 (0.1 + 0.538516 + 0.054772) = 31.891 >= 10, so
    k_max = sqrt(1.01) = 1.004987562112089.  Without the |.| row 1's term would be negative,
    stop the sum there with sqrt(-0.29) and never reach 10."""
    K = [1.0, 1.3] + [1.013 + 0.016 * (i - 2) for i in range(2, 12)] + [2.0] + [0.5] * 7
    freq = 340.0 / (2 * PI)
    grid = _Grid(2 * PI * freq / np.sqrt(np.array(K)), 46.0)
    _, ceffz, _, k_max, used = ess.modal_trace(grid, freq, 0.0, 0.0, 0.0, False)
    assert used and abs(k_max - 1.004987562112089) <= 1e-12
    kk, steps = ess.wkb_cutoff(ceffz, 2 * PI * freq, grid.dz)
    assert len(steps) == 2 and steps[1][1] == 2
    assert abs(steps[1][2] - 31.891281877342447) <= 1e-9


def test_wkb_linear_profile_closed_form():
    """k_eff^2(z) = K0 + b z rising linearly to z_p = 60 km, then falling to K_top < K0.
    For kk = K0 + n dkk the sum runs from the ground to the first row with kk - k_eff^2 <= dkk,
    z_s = (n - 1) dkk / b, one step of dkk short of the turning point z* = n dkk / b, so it is a
    rectangle rule of int_0^z_s sqrt(b (z* - z)) dz = (2/3) (dkk^(3/2) / b) (n^(3/2) - 1), and
    the stepping stops at the first n >= n_c = (15 b / dkk^(3/2) + 1)^(2/3).  The rule's
    relative error (~dz/z_s) is far below the distance of n_c from an integer in the cases
    below."""
    omega = 2 * PI * 0.4
    Nz, H = 20000, 150e3
    dz = H / (Nz - 1)
    z = np.arange(Nz) * dz
    zp = 60e3
    for K0, Kp in ((6.0e-6, 8.5e-6), (5.0e-6, 8.0e-6), (6.0e-6, 8.2e-6)):
        b = (Kp - K0) / zp
        K = np.where(z <= zp, K0 + b * z, K0 - 1e-6 + (Kp - K0 + 1e-6) * np.exp(-(z - zp) / 2e3))
        K[z > zp] = np.minimum(K[z > zp], Kp)
        ceff = omega / np.sqrt(K)
        dkk = (K.max() - K0) / 100.0   # the grid's largest k_eff^2 sits within b dz of Kp
        n_c = (15.0 * b / dkk ** 1.5 + 1.0) ** (2.0 / 3.0)
        frac = n_c - math.floor(n_c)
        assert 0.05 < frac < 0.95, n_c   # not a knife edge
        n_expect = math.ceil(n_c)
        kk, steps = ess.wkb_cutoff(ceff, omega, dz)
        assert steps[-1][2] >= 10.0 and all(s[2] < 10.0 for s in steps[:-1])
        assert len(steps) - 1 == n_expect, (len(steps) - 1, n_c)
        # the stopping row is z_s (to one grid row)
        assert abs(steps[-1][1] * dz - (n_expect - 1) * dkk / b) <= 2 * dz
        # through modal_trace (k_min from the falling top rows, below the cut-off)
        _, _, k_min, k_max, used = ess.modal_trace(_Grid(ceff, dz), 0.4, 0.0, 0.0, 0.0, False)
        assert used and abs(k_max ** 2 - (K0 + n_expect * dkk)) <= 1e-12 * Kp
        assert k_min < k_max


def test_wkb_ceff_minimum_at_ground():
    """c_eff smallest at the ground: k_max_full = k_gnd, dkk = 0 <= 1e-10, the loop is skipped
    and k_max = k_gnd (ESSModeSolver.cpp:584-606)."""
    omega = 2 * PI * 0.1
    Nz = 200
    ceff = 320.0 + 0.5 * np.arange(Nz)
    kk, steps = ess.wkb_cutoff(ceff, omega, 750.0)
    assert steps == [] and kk == (omega / 320.0) ** 2
    _, _, k_min, k_max, used = ess.modal_trace(_Grid(ceff, 750.0), 0.1, 0.0, 0.0, 0.0, False)
    assert used and abs(k_max - omega / 320.0) <= 1e-15
    assert abs(k_min - omega / ceff[Nz - Nz // 10 + 1]) <= 1e-18


def test_wkb_below_kmin_falls_back():
    """A cut-off below k_min turns WKB off: k_max = omega / min c_eff (ESSModeSolver.cpp:616-621).
    A fast ground layer (c_eff 500 m/s falling to 300 m/s over 6 km) below slower rows
    (c_eff rising back to 400 m/s at the top): the integral reaches 10 while the turning point
    is still inside the fast layer, at a phase speed above 400 m/s."""
    freq = 1.0
    omega = 2 * PI * freq
    Nz, dz = 2000, 50.0
    z = np.arange(Nz) * dz
    ceff = np.where(z < 6000.0, 500.0 - 200.0 * z / 6000.0,
                    np.minimum(300.0 + 100.0 * (z - 6000.0) / 50e3, 400.0))
    ceff[Nz - Nz // 10:] = 400.0
    kk, steps = ess.wkb_cutoff(ceff, omega, dz)
    assert omega / math.sqrt(kk) > 400.0   # the cut-off itself is below k_min
    _, _, k_min, k_max, used = ess.modal_trace(_Grid(ceff, dz), freq, 0.0, 0.0, 0.0, False)
    assert abs(k_min - omega / 400.0) <= 1e-15
    assert abs(k_max - omega / 300.0) <= 1e-15
