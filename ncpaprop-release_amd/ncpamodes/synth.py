"""Synthetic atmosphere profiles for tests and benchmarks.

The profile family P(seed) is a z,u,v,w,t,d,p table (Atmosphere1D text format,
src/atmosphere/file-format-description.txt) at 0.1 km spacing from 0 to 150 km:

* T(z) piecewise linear: 288 K at the ground, -6.5 K/km to 11 km, isothermal to 20 km,
  +1 K/km to 32 km, +2.8 K/km to 47 km, isothermal to 51 km, -2.8 K/km to 71 km,
  -2 K/km to 86 km, +4 K/km to 150 km (T(150 km) = 440.5 K).
* P hydrostatic from 101325 Pa with g = 9.80665 m/s^2 and R = 287 J/(kg K), integrated
  exactly per linear segment; rho = P/(R T); W = 0.
* U = A1 exp(-(z-50)^2/(2*8^2)) + A2 exp(-(z-11)^2/(2*2^2)),
  V = A3 exp(-(z-110)^2/(2*10^2)), nominal (A1, A2, A3) = (60, 25, 20) m/s, each scaled by
  an independent U(1-amp, 1+amp) factor drawn from numpy.random.default_rng(seed) in the
  order A1, A2, A3 (amp = 0.2).  The serving variant also shifts the jet heights by
  U(-dz, +dz) km (drawn after the amplitudes).

Units written: Z km, U/V/W m/s, T degK, RHO g/cm3, P mbar, with a scalar Z0 = 0 m.
"""
import numpy as np

G0 = 9.80665
R_AIR = 287.0

# (top of segment [km], lapse rate [K/km])
_T_SEGMENTS = [(11.0, -6.5), (20.0, 0.0), (32.0, 1.0), (47.0, 2.8), (51.0, 0.0),
               (71.0, -2.8), (86.0, -2.0), (150.0, 4.0)]


def temperature_pressure(z_km):
    """T [K] and hydrostatic P [Pa] at altitudes z_km (array, km)."""
    z_km = np.asarray(z_km, dtype=np.float64)
    T = np.empty_like(z_km)
    P = np.empty_like(z_km)
    zb, Tb, Pb = 0.0, 288.0, 101325.0
    prev_top = 0.0
    for top, lapse in _T_SEGMENTS:
        sel = (z_km >= prev_top) & ((z_km <= top) if top == _T_SEGMENTS[-1][0] else (z_km < top))
        dz = z_km[sel] - zb
        T[sel] = Tb + lapse * dz
        L = lapse / 1000.0  # K/m
        if lapse == 0.0:
            P[sel] = Pb * np.exp(-G0 * dz * 1000.0 / (R_AIR * Tb))
        else:
            P[sel] = Pb * (T[sel] / Tb) ** (-G0 / (R_AIR * L))
        # advance base to the top of this segment
        Tt = Tb + lapse * (top - zb)
        if lapse == 0.0:
            Pt = Pb * np.exp(-G0 * (top - zb) * 1000.0 / (R_AIR * Tb))
        else:
            Pt = Pb * (Tt / Tb) ** (-G0 / (R_AIR * L))
        zb, Tb, Pb = top, Tt, Pt
        prev_top = top
    return T, P


def jet_parameters(seed, amp=0.2, height_shift_km=0.0):
    """(A1, A2, A3, z1, z2, z3) for seed; see module docstring for the draw order."""
    rng = np.random.default_rng(seed)
    f = rng.uniform(1.0 - amp, 1.0 + amp, size=3)
    A = np.array([60.0, 25.0, 20.0]) * f
    heights = np.array([50.0, 11.0, 110.0])
    if height_shift_km > 0.0:
        heights = heights + rng.uniform(-height_shift_km, height_shift_km, size=3)
    return A[0], A[1], A[2], heights[0], heights[1], heights[2]


def profile_columns(seed=0, amp=0.2, height_shift_km=0.0, dz_km=0.1, ztop_km=150.0):
    """Columns z [km], u, v, w [m/s], T [K], rho [g/cm3], P [mbar] of profile P(seed)."""
    n = int(round(ztop_km / dz_km)) + 1
    z = np.arange(n, dtype=np.float64) * dz_km
    T, P = temperature_pressure(z)
    rho = P / (R_AIR * T)  # kg/m3
    A1, A2, A3, z1, z2, z3 = jet_parameters(seed, amp, height_shift_km)
    u = A1 * np.exp(-(z - z1) ** 2 / (2 * 8.0 ** 2)) + A2 * np.exp(-(z - z2) ** 2 / (2 * 2.0 ** 2))
    v = A3 * np.exp(-(z - z3) ** 2 / (2 * 10.0 ** 2))
    w = np.zeros_like(z)
    return z, u, v, w, T, rho * 1e-3, P * 0.01


HEADER = ("#% 0, Z0, m, 0.0\n#% 1, Z, km\n#% 2, U, m/s\n#% 3, V, m/s\n#% 4, W, m/s\n"
          "#% 5, T, degK\n#% 6, RHO, g/cm3\n#% 7, P, mbar\n")


def profile_text(seed=0, amp=0.2, height_shift_km=0.0):
    cols = profile_columns(seed, amp, height_shift_km)
    lines = [HEADER]
    data = np.column_stack(cols)
    for row in data:
        lines.append(" ".join("%.10e" % x for x in row) + "\n")
    return "".join(lines)


def write_profile(path, seed=0, amp=0.2, height_shift_km=0.0):
    with open(path, "w") as fh:
        fh.write(profile_text(seed, amp, height_shift_km))
    return path


def hydrostatic(z_km, T, P0=101325.0):
    """P [Pa] on the nodes z_km for temperatures T [K]: d ln P/dz = -g/(R T), trapezoidal
    in 1/T between nodes (for profiles other than the family's piecewise-linear T)."""
    z_m = np.asarray(z_km, dtype=np.float64) * 1000.0
    inv = 1.0 / np.asarray(T, dtype=np.float64)
    lnp = np.concatenate([[0.0], np.cumsum(0.5 * (inv[1:] + inv[:-1]) * np.diff(z_m))])
    return P0 * np.exp(-G0 / R_AIR * lnp)


def write_columns(path, z_km, u, v, T, P=None):
    """Write a z,u,v,w,t,d,p profile (header and units as the family's) from node arrays:
    z [km], u, v [m/s], T [K], P [Pa] (hydrostatic from T when None); w = 0, rho = P/(R T)."""
    z_km = np.asarray(z_km, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    P = hydrostatic(z_km, T) if P is None else np.asarray(P, dtype=np.float64)
    rho = P / (R_AIR * T)
    data = np.column_stack([z_km, np.asarray(u, dtype=np.float64), np.asarray(v, dtype=np.float64),
                            np.zeros_like(z_km), T, rho * 1e-3, P * 0.01])
    with open(path, "w") as fh:
        fh.write(HEADER)
        for row in data:
            fh.write(" ".join("%.10e" % x for x in row) + "\n")
    return path
