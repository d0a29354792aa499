"""Synthetic inputs and closed-form solutions for the benchmark configurations.

All functions take unit-sphere coordinates (numpy arrays of any shape).
Sign conventions of the reference kernels: psi = int G f with
G(x, y) = -1/(4 pi) log(1 - x.y) solves  Lap psi = -(f - mean f), so a degree-l
spherical harmonic maps to Y/(l(l+1)); the Biot-Savart velocity is
u = -x cross grad psi (bve_vel_impl.hpp:61-66 is -x cross grad_x of G).
"""
import numpy as np

FOUR_PI = 4.0 * np.pi


def latlon(x, y, z):
    """xyz_to_latlon (general_utils.cpp:34-38)."""
    lat = np.pi / 2.0 - np.arctan2(np.sqrt(x * x + y * y), z)
    lon = np.arctan2(y, x)
    return lat, lon


# ---------------------------------------------------------------- Rossby-Haurwitz wave 4

def rh4_vorticity(x, y, z, omega0=7.848e-6, K=7.848e-6):
    """2(w0+K) sin(lat) - 30 K cos^4(lat) sin(lat) cos(4 lon): an l=1 plus an l=5 harmonic.
    cos^4(lat) cos(4 lon) = Re((x + i y)^4) = x^4 - 6 x^2 y^2 + y^4."""
    c4 = x ** 4 - 6 * x ** 2 * y ** 2 + y ** 4
    return 2 * (omega0 + K) * z - 30 * K * z * c4


def rh4_stream(x, y, z, omega0=7.848e-6, K=7.848e-6):
    """psi = int G omega: degree-1 part / 2, degree-5 part / 30."""
    c4 = x ** 4 - 6 * x ** 2 * y ** 2 + y ** 4
    return (omega0 + K) * z - K * z * c4


def rh4_velocity(x, y, z, omega0=7.848e-6, K=7.848e-6):
    """u = -x cross grad psi, grad of the polynomial extension of rh4_stream."""
    c4 = x ** 4 - 6 * x ** 2 * y ** 2 + y ** 4
    gx = -K * z * (4 * x ** 3 - 12 * x * y ** 2)
    gy = -K * z * (4 * y ** 3 - 12 * x ** 2 * y)
    gz = (omega0 + K) - K * c4
    ux = -(y * gz - z * gy)
    uy = -(z * gx - x * gz)
    uz = -(x * gy - y * gx)
    return ux, uy, uz


# ---------------------------------------------------------------- real spherical harmonics

def real_sph_harm_all(lmax, x, y, z):
    """Orthonormal real spherical harmonics Y_lm, l = 0..lmax, m = -l..l, at unit
    vectors; returns a dict {(l, m): array}.  Normalised associated Legendre
    recurrence (Holmes & Featherstone 2002)."""
    ct = np.clip(z, -1.0, 1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    lon = np.arctan2(y, x)
    out = {}
    pmm = np.full_like(ct, np.sqrt(1.0 / FOUR_PI))  # P_00
    for m in range(lmax + 1):
        if m > 0:
            pmm = -np.sqrt((2 * m + 1) / (2.0 * m)) * st * pmm
        plm2, plm1 = None, pmm
        for l in range(m, lmax + 1):
            if l == m:
                p = pmm
            elif l == m + 1:
                p = np.sqrt(2 * m + 3.0) * ct * pmm
            else:
                a = np.sqrt((4.0 * l * l - 1) / (l * l - m * m))
                b = np.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1))
                p = a * (ct * plm1 - b * plm2)
            if l > m:
                plm2, plm1 = plm1, p
            if m == 0:
                out[(l, 0)] = p
            else:
                out[(l, m)] = np.sqrt(2.0) * p * np.cos(m * lon)
                out[(l, -m)] = np.sqrt(2.0) * p * np.sin(m * lon)
    return out


def random_sh_forcing(x, y, z, lmax=32, seed=0):
    """Sum a_lm Y_lm, 1 <= l <= lmax, a_lm ~ N(0, (1+l)^-2) (zero mean: l = 0 excluded).
    Returns (f, psi_exact) with psi_exact = sum a_lm Y_lm / (l (l+1))."""
    rng = np.random.default_rng(seed)
    Y = real_sph_harm_all(lmax, x, y, z)
    f = np.zeros_like(x)
    psi = np.zeros_like(x)
    for l in range(1, lmax + 1):
        for m in range(-l, l + 1):
            a = rng.normal(0.0, 1.0 / (1.0 + l))
            f += a * Y[(l, m)]
            psi += a * Y[(l, m)] / (l * (l + 1.0))
    return f, psi


# ---------------------------------------------------------------- Gaussian vortices

def gaussian_vortices(x, y, z, area, count=12, seed=0):
    """sum_k a_k exp(-(gcdist(x, c_k)/w_k)^2): centres uniform on the sphere,
    amplitudes +-U(1e-5, 3e-5) 1/s, widths U(0.05, 0.2) rad; then the
    area-weighted mean is removed (balance_ic, initial_conditions.cpp:377-384)."""
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(count, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    amp = rng.uniform(1e-5, 3e-5, count) * rng.choice([-1.0, 1.0], count)
    width = rng.uniform(0.05, 0.2, count)
    w = np.zeros_like(x)
    for k in range(count):
        dp = np.clip(x * c[k, 0] + y * c[k, 1] + z * c[k, 2], -1.0, 1.0)
        w += amp[k] * np.exp(-(np.arccos(dp) / width[k]) ** 2)
    w -= np.sum(w * area) / FOUR_PI
    return w


# ---------------------------------------------------------------- shallow water on a rotating sphere

EARTH_RADIUS = 6.37122e6   # m
GRAVITY = 9.80616          # m s^-2


def gaussian_mountain(seed=0):
    """Seeded Gaussian mountain: height U(1000, 2500) m, width U(0.1, 0.3) rad, centre
    uniform on the sphere.  Returns (height, width, lat, lon)."""
    rng = np.random.default_rng(seed)
    h0 = rng.uniform(1000.0, 2500.0)
    w = rng.uniform(0.1, 0.3)
    c = rng.normal(size=3)
    c /= np.linalg.norm(c)
    return float(h0), float(w), float(np.arcsin(c[2])), float(np.arctan2(c[1], c[0]))


def mountain_height(x, y, z, mountain):
    """h_s = H exp(-(r/w)^2), r the great-circle distance to the centre."""
    h0, w, lat, lon = mountain
    c = np.array([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])
    cx, cy, cz = y * c[2] - z * c[1], z * c[0] - x * c[2], x * c[1] - y * c[0]
    r = np.arctan2(np.sqrt(cx * cx + cy * cy + cz * cz), x * c[0] + y * c[1] + z * c[2])
    return h0 * np.exp(-(r / w) ** 2)


def swe_zonal_flow(x, y, z, mountain=None, u0=20.0, depth=5960.0, omega=7.292e-5, a=EARTH_RADIUS, g=GRAVITY):
    """Balanced zonal flow u = u0 cos(lat) on the unit sphere (velocities in rad/s, time in s,
    heights in m): relative vorticity 2 (u0/a) sin(lat), zero divergence, and the fluid depth
    under the free surface depth - (a Omega u0 + u0^2/2) sin^2(lat) / g (Williamson et al. 1992,
    test cases 2 and 5), minus the mountain.  Returns (vor, div, h)."""
    vor = 2.0 * (u0 / a) * z
    surface = depth - (a * omega * u0 + 0.5 * u0 * u0) * z * z / g
    hs = 0.0 if mountain is None else mountain_height(x, y, z, mountain)
    return vor, np.zeros_like(z), surface - hs


def random_unit_vectors(n, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def vec_rel_l2(u, v):
    """rel-L2 of a 3-component field (lists of arrays)."""
    num = sum(np.sum((np.asarray(a) - np.asarray(b)) ** 2) for a, b in zip(u, v))
    den = sum(np.sum(np.asarray(b) ** 2) for b in v)
    return float(np.sqrt(num / max(den, 1e-300)))


def vec_rel_linf(u, v):
    num = max(float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) for a, b in zip(u, v))
    den = max(float(np.max(np.abs(np.asarray(b)))) for b in v)
    return num / max(den, 1e-300)
