"""Oracle: atmosphere profile reader, units, Steffen splines and derived quantities.

Test infrastructure only (see oracle/__init__.py).
"""
import math
import numpy as np

PI = 3.141592653589793          # src/common/units.h:63-65, src/common/util.h:13-15
GAMMA_FOR_C = 1.4               # src/atmosphere/AtmosphericModel.cpp:5-7
R_FOR_C = 287.0                 # src/atmosphere/AtmosphericModel.cpp:9-11

# --------------------------------------------------------------------------------------
# Units (src/common/units.cpp:131-178, string table :225-283)
# --------------------------------------------------------------------------------------
_UNIT_ALIASES = {
    "": "none", "N/A": "none",
    "K": "K", "DEGK": "K", "DEG K": "K", "DEGREES K": "K", "KELVIN": "K",
    "C": "C", "DEGC": "C", "DEG C": "C", "DEGREES C": "C", "CELSIUS": "C", "CENTIGRADE": "C",
    "F": "F", "DEGF": "F", "DEG F": "F", "DEGREES F": "F", "FAHRENHEIT": "F",
    "M": "m", "METERS": "m", "KM": "km", "KILOMETERS": "km",
    "M/S": "m/s", "MPS": "m/s", "MPERS": "m/s", "M PER S": "m/s", "METERS PER SECOND": "m/s",
    "KM/S": "km/s", "KMPS": "km/s", "KMPERS": "km/s", "KM PER S": "km/s",
    "KILOMETERS PER SECOND": "km/s",
    "PA": "Pa", "PASCAL": "Pa", "PASCALS": "Pa",
    "MBAR": "mbar", "MILLIBAR": "mbar", "MILLIBARS": "mbar",
    "HECTOPASCALS": "hPa", "HPA": "hPa", "ATMOSPHERES": "atm", "ATM": "atm",
    "KG/M3": "kg/m3", "KGPM3": "kg/m3", "KILOGRAMS PER CUBIC METER": "kg/m3",
    "G/CM3": "g/cm3", "GPCM3": "g/cm3", "GRAMS PER CUBIC CENTIMETER": "g/cm3",
    "DEGREES CLOCKWISE FROM NORTH": "degcwn", "DEG CW FROM N": "degcwn", "AZIMUTH": "degcwn",
    "DEGREES COUNTERCLOCKWISE FROM EAST": "degccwe", "DEG CCW FROM E": "degccwe",
    "DEG": "deg", "DEGREES": "deg", "RAD": "rad", "RADIANS": "rad",
}

_CONV = {
    ("C", "F"): lambda x: x * 1.8 + 32.0,
    ("C", "K"): lambda x: x + 273.15,
    ("K", "F"): lambda x: (x - 273.15) * 1.8 + 32.0,
    ("K", "C"): lambda x: x - 273.15,
    ("F", "K"): lambda x: (x - 32.0) * 5.0 / 9.0 + 273.15,
    ("F", "C"): lambda x: (x - 32.0) * 5.0 / 9.0,
    ("m", "km"): lambda x: x * 0.001,
    ("km", "m"): lambda x: x * 1000.0,
    ("m/s", "km/s"): lambda x: x * 0.001,
    ("km/s", "m/s"): lambda x: x * 1000.0,
    ("Pa", "mbar"): lambda x: x * 0.01,
    ("Pa", "hPa"): lambda x: x * 0.01,
    ("mbar", "Pa"): lambda x: x * 100.0,
    ("mbar", "hPa"): lambda x: x,
    ("hPa", "Pa"): lambda x: x * 100,
    ("hPa", "mbar"): lambda x: x,
    ("atm", "hPa"): lambda x: x * 1013.25,
    ("atm", "Pa"): lambda x: x * 101325.0,
    ("atm", "mbar"): lambda x: x * 1013.25,
    ("hPa", "atm"): lambda x: x * 0.000986923,
    ("Pa", "atm"): lambda x: x * 0.00000986923,
    ("mbar", "atm"): lambda x: x * 0.000986923,
    ("kg/m3", "g/cm3"): lambda x: x * 0.001,
    ("g/cm3", "kg/m3"): lambda x: x * 1000.0,
    ("rad", "deg"): lambda x: x * 180.0 / PI,
    ("deg", "rad"): lambda x: x * PI / 180.0,
}


def units_from_string(s):
    """Units::fromString (units.cpp:47-66): case-insensitive lookup, None if unknown."""
    return _UNIT_ALIASES.get(s.upper())


def convert(values, u_from, u_to):
    """Units::convert (units.cpp:310-350); identity when units agree."""
    if u_from == u_to:
        return values
    try:
        f = _CONV[(u_from, u_to)]
    except KeyError:
        raise ValueError("Undefined conversion requested from %s to %s" % (u_from, u_to))
    return f(np.asarray(values, dtype=np.float64)) if not np.isscalar(values) else f(values)


# --------------------------------------------------------------------------------------
# Profile file reader (src/atmosphere/Atmosphere1D.cpp:85-285)
# --------------------------------------------------------------------------------------
def _split(s, delims):
    """NCPA::split (src/common/util.cpp:101-128): tokens separated by any delimiter char."""
    out, cur = [], []
    for ch in s:
        if ch in delims:
            if cur:
                out.append("".join(cur))
                cur = []
        else:
            cur.append(ch)
    if cur:
        out.append("".join(cur))
    return out


class Profile:
    """A parsed 1-D profile: altitude vector plus named vectors and scalars with units."""

    def __init__(self, z, z_units, vectors, scalars):
        self.z = np.asarray(z, dtype=np.float64)
        self.z_units = z_units
        self.vectors = {k: (np.asarray(v, dtype=np.float64), u) for k, (v, u) in vectors.items()}
        self.scalars = dict(scalars)

    def vec(self, key, units=None):
        v, u = self.vectors[key]
        return v if units is None else convert(v, u, units)


def read_profile(path, header_path=""):
    """Atmosphere1D(file, headerfile): header lines '#% col, KEY, units[, value]'.

    Column 0 entries are scalars, column 1 is the altitude, data lines are separated by
    whitespace or commas and '#' lines are comments (Atmosphere1D.cpp:85-105,107-285).
    """
    hfile = header_path if header_path else path
    headerlines = []
    with open(hfile) as fh:
        for line in fh:
            line = line.strip(" \t\n\r")
            if line[:1] == "#" and line[1:2] == "%":
                headerlines.append(line[2:])
    if not headerlines:
        raise ValueError("Atmosphere1D - No descriptive fields found.")
    cols, keys, units, values = [], [], [], []
    for h in headerlines:
        f = _split(h.strip(" \t\n\r"), " ,")
        if len(f) not in (3, 4):
            raise ValueError("Atmosphere1D - Error parsing descriptive line: " + h)
        cols.append(int(f[0]))
        u = units_from_string(f[2].strip())
        if u is None or u == "none":
            raise ValueError("Unrecognized units string: " + f[2])
        keys.append(f[1].strip())
        units.append(u)
        values.append(float(f[3]) if len(f) == 4 else 0.0)
    if len(set(keys)) != len(keys):
        raise ValueError("Atmosphere1D: key values are not unique")
    ncols = max(cols)
    zunits = None
    for c, u in zip(cols, units):
        if c == 1:
            zunits = u
    rows = []
    with open(path) as fh:
        for line in fh:
            line = line.strip(" \t\n\r")
            if not line or line[0] == "#":
                continue
            f = _split(line, " \t,")
            if len(f) != ncols:
                raise ValueError("Error parsing data line: %d columns expected, %d found"
                                 % (ncols, len(f)))
            rows.append([float(x) for x in f])
    data = np.array(rows, dtype=np.float64).reshape(-1, ncols)
    vectors, scalars = {}, {}
    for c, k, u, v in zip(cols, keys, units, values):
        if c == 0:
            scalars[k] = (v, u)
        elif c > 1:
            vectors[k] = (data[:, c - 1].copy(), u)
    return Profile(data[:, 0].copy(), zunits, vectors, scalars)


# --------------------------------------------------------------------------------------
# Steffen monotone cubic (GSL gsl_interp_steffen, selected by AtmosphericProperty1D.cpp:76-80)
# Restated from M. Steffen, A&A 239, 443 (1990): interior slopes eq. 11, one-sided
# "simplest possibility" end slopes (sec. 2.2), Hermite coefficients eqs. 2-5.
# --------------------------------------------------------------------------------------
class Steffen:
    def __init__(self, x, y):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        n = x.size
        if n < 3:
            raise ValueError("Steffen spline needs at least 3 points")
        h = np.diff(x)
        s = np.diff(y) / h
        yp = np.empty(n)
        yp[0] = s[0]
        yp[-1] = s[-1]
        him1, hi = h[:-1], h[1:]
        sim1, si = s[:-1], s[1:]
        p = (sim1 * hi + si * him1) / (him1 + hi)
        sg = lambda v: np.where(v < 0.0, -1.0, 1.0)
        yp[1:-1] = (sg(sim1) + sg(si)) * np.minimum(np.abs(sim1),
                                                      np.minimum(np.abs(si), 0.5 * np.abs(p)))
        self.x = x
        self.a = (yp[:-1] + yp[1:] - 2.0 * s) / h / h
        self.b = (3.0 * s - 2.0 * yp[:-1] - yp[1:]) / h
        self.c = yp[:-1].copy()
        self.d = y[:-1].copy()
        self.y = y

    def index(self, xq):
        """Interval index i with x[i] <= xq < x[i+1], capped at n-2 (gsl_interp_bsearch)."""
        i = np.searchsorted(self.x, xq, side="right") - 1
        return np.clip(i, 0, self.x.size - 2)

    def eval_raw(self, xq):
        xq = np.asarray(xq, dtype=np.float64)
        i = self.index(xq)
        dx = xq - self.x[i]
        return self.d[i] + dx * (self.c[i] + dx * (self.b[i] + dx * self.a[i]))

    def get(self, xq):
        """AtmosphericProperty1D::get (AtmosphericProperty1D.cpp:116-126): clamp outside."""
        xq = np.asarray(xq, dtype=np.float64)
        out = self.eval_raw(xq)
        out = np.where(xq > self.x[-1], self.y[-1], out)
        out = np.where(xq < self.x[0], self.y[0], out)
        return out

    def deriv(self, xq):
        """AtmosphericProperty1D::get_first_derivative (:128-136): 0 outside the range."""
        xq = np.asarray(xq, dtype=np.float64)
        i = self.index(xq)
        dx = xq - self.x[i]
        d1 = self.c[i] + dx * (2.0 * self.b[i] + dx * 3.0 * self.a[i])
        inside = (xq >= self.x[0]) & (xq <= self.x[-1])
        return np.where(inside, d1, 0.0)


def read_attenuation_file(path):
    """Atmosphere1D::read_attenuation_from_file (Atmosphere1D.cpp:307-367): lines that do not
    start with '#' hold 'altitude[km] attenuation[Np/m]', split at blanks or commas.
    Returns (z_km, alpha)."""
    zs, aa = [], []
    with open(path) as fh:
        for line in fh:
            t = line.strip()
            if not t or t[0] == "#":
                continue
            f = [x for x in t.replace(",", " ").split() if x]
            if len(f) != 2:
                raise ValueError("Atmosphere1D - Error parsing attenuation line:\n" + line +
                                 "Must be formatted as:\naltitude  attenuation")
            zs.append(float(f[0]))
            aa.append(float(f[1]))
    return np.array(zs), np.array(aa)


def attenuation_file_nodes(path, z_nodes_km):
    """_ALPHA_ at the profile nodes from an attenuation file (Atmosphere1D.cpp:368-414): the
    table is extended by a node at the profile's lowest (highest) altitude carrying its first
    (last) value when it starts above (ends below) the profile, then Steffen-interpolated at
    the nodes.  (The reference's extension copies bytes instead of doubles; the build
    implements the evident intent, which holds the end values.)"""
    z, a = read_attenuation_file(path)
    z_nodes_km = np.asarray(z_nodes_km, dtype=np.float64)
    if z[0] > z_nodes_km[0]:
        z, a = np.concatenate([[z_nodes_km[0]], z]), np.concatenate([[a[0]], a])
    if z[-1] < z_nodes_km[-1]:
        z, a = np.concatenate([z, [z_nodes_km[-1]]]), np.concatenate([a, [a[-1]]])
    return Steffen(z, a).get(z_nodes_km)


# --------------------------------------------------------------------------------------
# Sutherland-Bass attenuation (src/atmosphere/AtmosphericModel.cpp:86-219)
# --------------------------------------------------------------------------------------
def attenuation_sutherland_bass(z, T, P, D, freq):
    """alpha [Np/m] at z [km], T [K], P [Pa], D [kg/m3], freq [Hz]; vectorised."""
    with np.errstate(over="ignore"):   # np.where evaluates the untaken branch too
        return _attenuation_sutherland_bass(z, T, P, D, freq)


def _attenuation_sutherland_bass(z, T, P, D, freq):
    z = np.asarray(z, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    P = np.asarray(P, dtype=np.float64)
    D = np.asarray(D, dtype=np.float64)
    mu_o, T_o, P_o, S = 18.192e-6, 293.15, 101325.0, 117.0
    Cv_R = [5.0 / 2.0, 5.0 / 2.0, 3.0, 3.0]
    Cp_R = [7.0 / 2.0, 7.0 / 2.0, 4.0, 4.0]
    theta = [2239.1, 3352.0, 915.0, 1037.0]
    c_snd = np.sqrt(GAMMA_FOR_C * P / D)
    mu = mu_o * np.sqrt(T / T_o) * ((1.0 + S / T_o) / (1.0 + S / T))
    nu = (8.0 * PI * freq * mu) / (3.0 * P)
    X = [None] * 7
    X[0] = np.where(z > 90.0, 10.0 ** (49.296 - 1.5524 * z + 1.8714e-2 * z ** 2
                                       - 1.1069e-4 * z ** 3 + 3.199e-7 * z ** 4
                                       - 3.6211e-10 * z ** 5), 10.0 ** -0.67887)
    X[1] = np.where(z > 76.0, 10.0 ** (1.3972e-1 - 5.6269e-3 * z + 3.9407e-5 * z ** 2
                                       - 1.0737e-7 * z ** 3), 10.0 ** -0.10744)
    X[2] = np.full_like(z, 10.0 ** -3.3979)
    X[3] = np.where(z > 80.0, 10.0 ** (-4.234 - 3.0975e-2 * z),
                    10.0 ** (-19.027 + 1.3093 * z - 4.6496e-2 * z ** 2 + 7.8543e-4 * z ** 3
                             - 6.5169e-6 * z ** 4 + 2.1343e-8 * z ** 5))
    X[4] = np.where(z > 95.0, 10.0 ** (-3.2456 + 4.6642e-2 * z - 2.6894e-4 * z ** 2
                                       + 5.264e-7 * z ** 3),
                    10.0 ** (-11.195 + 1.5408e-1 * z - 1.4348e-3 * z ** 2 + 1.0166e-5 * z ** 3))
    X[5] = 10.0 ** (-53.746 + 1.5439 * z - 1.8824e-2 * z ** 2 + 1.1587e-4 * z ** 3
                    - 3.5399e-7 * z ** 4 + 4.2609e-10 * z ** 5)
    # H2O: the 'z > 100' branch sits inside 'else' of 'z > 30' and is unreachable
    # (AtmosphericModel.cpp:151-162); restated as written.
    X[6] = np.where(z > 30.0,
                    10.0 ** (-4.2563 + 7.6245e-2 * z - 2.1824e-3 * z ** 2 - 2.3010e-6 * z ** 3
                             + 2.4265e-7 * z ** 4 - 1.2500e-09 * z ** 5),
                    10.0 ** (-1.7491 + 4.4986e-2 * z - 6.8549e-2 * z ** 2 + 5.4639e-3 * z ** 3
                             - 1.5539e-4 * z ** 4 + 1.5063e-06 * z ** 5))
    X_ON = (X[0] + X[1]) / 0.9903
    Z_rot_0 = 54.1 * np.exp(-17.3 * T ** (-1.0 / 3.0))
    Z_rot_1 = 63.3 * np.exp(-16.7 * T ** (-1.0 / 3.0))
    Z_rot = 1.0 / ((X[1] / Z_rot_1) + (X[0] / Z_rot_0))
    sigma = 5.0 / math.sqrt(21.0)
    nn = (4.0 / 5.0) * math.sqrt(3.0 / 7.0) * Z_rot
    chi = 3.0 * nn * nu / 4.0
    cchi = 2.36 * chi
    a_cl = (2 * PI * freq / c_snd) * np.sqrt(0.5 * (np.sqrt(1 + nu ** 2) - 1) * (1 + cchi ** 2)
                                              / ((1 + nu ** 2) * (1 + (sigma * cchi) ** 2)))
    a_rot = (2 * PI * freq / c_snd) * X_ON * ((sigma ** 2 - 1) * chi / (2 * sigma)) \
        * np.sqrt(0.5 * (np.sqrt(1 + nu ** 2) + 1) / ((1 + nu ** 2) * (1 + cchi ** 2)))
    a_diff = 0.003 * a_cl
    Tr = (T / T_o) ** (-1.0 / 3.0) - 1.0
    A1 = (X[0] + X[1]) * 24.0 * np.exp(-9.16 * Tr)
    A2 = (X[4] + X[5]) * 2400.0
    B = 40400.0 * np.exp(10.0 * Tr)
    C = 0.02 * np.exp(-11.2 * Tr)
    Dd = 0.391 * np.exp(8.41 * Tr)
    E = 9.0 * np.exp(-19.9 * Tr)
    F = 60000.0
    G = 28000.0 * np.exp(-4.17 * Tr)
    H = 22000.0 * np.exp(-7.68 * Tr)
    I = 15100.0 * np.exp(-10.4 * Tr)
    J = 11500.0 * np.exp(-9.17 * Tr)
    K = 8.48e08 * np.exp(9.17 * Tr)
    L = np.exp(-7.72 * Tr)
    ZZ = H * X[2] + I * (X[0] + 0.5 * X[4]) + J * (X[1] + 0.5 * X[5]) + K * (X[6] + X[3])
    hu = 100.0 * (X[3] + X[6])
    pf = (P / P_o) * (mu_o / mu)
    f_vib = [pf * (A1 + A2 + B * hu * (C + hu) * (Dd + hu)),
             pf * (E + F * X[3] + G * X[6]),
             pf * ZZ,
             pf * 1.2e5 * L]
    a_vib = np.zeros_like(z)
    for m in range(4):
        C_R = ((theta[m] / T) ** 2) * np.exp(-theta[m] / T) / (1 - np.exp(-theta[m] / T)) ** 2
        A_max = (X[m] * (PI / 2) * C_R) / (Cp_R[m] * (Cv_R[m] + C_R))
        a_vib = a_vib + (A_max / c_snd) * ((2 * freq ** 2 / f_vib[m]) / (1 + (freq / f_vib[m]) ** 2))
    return a_cl + a_rot + a_diff + a_vib


# --------------------------------------------------------------------------------------
# Node-level derived quantities used by ESSModeSolver::setParams
# --------------------------------------------------------------------------------------
class AtmosphereNodes:
    """Profile converted to SI at the nodes, as ESSModeSolver::setParams leaves it.

    z [m], u, v [m/s], T [K], P [Pa], rho [kg/m3], c0 [m/s] (C0 column, else
    sqrt(1.4 P/rho), else sqrt(1.4*287*T); ESSModeSolver.cpp:159-183), wind speed and
    direction (Atmosphere1D.cpp:484-544), ground height Z0 [m].
    """

    def __init__(self, prof):
        self.z = convert(prof.z, prof.z_units, "m")
        if "U" not in prof.vectors:
            raise ValueError("U wind vector not found in atmosphere file")
        if "V" not in prof.vectors:
            raise ValueError("V wind vector not found in atmosphere file")
        self.u = prof.vec("U", "m/s")
        self.v = prof.vec("V", "m/s")
        if "C0" in prof.vectors:
            self.c0 = prof.vec("C0", "m/s")
            T = prof.vec("T", "K") if "T" in prof.vectors else self.c0 * self.c0 / GAMMA_FOR_C / R_FOR_C
        else:
            T = prof.vec("T", "K") if "T" in prof.vectors else None
            if "P" in prof.vectors and "RHO" in prof.vectors:
                self.c0 = np.sqrt(GAMMA_FOR_C * prof.vec("P", "Pa") / prof.vec("RHO", "kg/m3"))
            elif T is not None:
                self.c0 = np.sqrt(GAMMA_FOR_C * R_FOR_C * T)
            else:
                raise ValueError("Can't calculate static sound speed, either temperature T or "
                                 "pressure P and density RHO missing")
        if T is None:
            raise ValueError("T temperature vector not found in atmosphere file")
        if "P" not in prof.vectors:
            raise ValueError("P pressure vector not found in atmosphere file")
        if "RHO" not in prof.vectors:
            raise ValueError("RHO density vector not found in atmosphere file")
        self.T = T
        self.P = prof.vec("P", "Pa")
        self.rho = prof.vec("RHO", "kg/m3")
        self.Z0 = None
        if "Z0" in prof.scalars:
            v, u = prof.scalars["Z0"]
            self.Z0 = convert(v, u, "m")
        self.ws = np.sqrt(self.u * self.u + self.v * self.v)
        self.wd = (PI / 2.0 - np.arctan2(self.v, self.u)) * 180.0 / PI
        self._splines = {}

    def spline(self, key):
        if key not in self._splines:
            self._splines[key] = Steffen(self.z, getattr(self, key))
        return self._splines[key]

    def ceff_nodes(self, azimuth_deg):
        """_WC_ + _C0_ at the nodes (Atmosphere1D.cpp:546-610)."""
        az = azimuth_deg * PI / 180.0
        wc = self.ws * np.cos(self.wd * PI / 180.0 - az)
        return wc + self.c0

    def alpha_nodes(self, freq):
        """_ALPHA_ at the nodes (Atmosphere1D.cpp:1074-1112): z converted to km."""
        return attenuation_sutherland_bass(self.z * 0.001, self.T, self.P, self.rho, freq)
