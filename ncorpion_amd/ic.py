"""Synthetic initial conditions: thin Keplerian disks around a central mass.

Body records use the layout of NcorpiON's struct moonlet (structure.h:74-84):
columns x, y, z, vx, vy, vz, mass, radius (float64, 64 bytes per body).

The disk recipe: semi-major axis uniform in [a_min, a_max]; eccentricity and
inclination Rayleigh-distributed (sigma_e, sigma_i in radians); mean anomaly,
argument of periapsis and longitude of the ascending node uniform in
[0, 2 pi).  Kepler's equation is solved by Newton iterations to machine
precision.  The central body (mass `mu`/G) sits at the origin.
"""
import numpy as np

__all__ = ["disk", "disk_c3_masses", "write_init_txt", "seeded_ensemble_sizes"]


def _kepler_E(M, e):
    E = M + e * np.sin(M)
    for _ in range(50):
        f = E - e * np.sin(E) - M
        d = f / (1.0 - e * np.cos(E))
        E = E - d
        if np.max(np.abs(d)) < 1e-16:
            break
    return E


def disk(n, seed=0, mass=1e-8, radius=2e-4, a_range=(0.9, 1.1), sigma_e=0.01, sigma_i=0.005,
         mu=1.0, masses=None, radii=None):
    """Return an (n, 8) float64 array of bodies on a thin disk.

    `mass`/`radius` give equal bodies; `masses`/`radii` (arrays of length n)
    override them.  `mu` = G * M_central.
    """
    rng = np.random.default_rng(seed)
    a = rng.uniform(a_range[0], a_range[1], n)
    e = np.minimum(rng.rayleigh(sigma_e, n), 0.9)
    inc = rng.rayleigh(sigma_i, n)
    M = rng.uniform(0.0, 2.0 * np.pi, n)
    om = rng.uniform(0.0, 2.0 * np.pi, n)
    Om = rng.uniform(0.0, 2.0 * np.pi, n)
    E = _kepler_E(M, e)
    cE, sE = np.cos(E), np.sin(E)
    b = a * np.sqrt(1.0 - e * e)
    # Position and velocity in the orbital plane (periapsis along +x).
    px = a * (cE - e)
    py = b * sE
    r = a * (1.0 - e * cE)
    nmot = np.sqrt(mu / a ** 3)
    vx = -a * nmot * sE * a / r
    vy = b * nmot * cE * a / r
    co, so = np.cos(om), np.sin(om)
    cO, sO = np.cos(Om), np.sin(Om)
    ci, si = np.cos(inc), np.sin(inc)
    # Rotation R = Rz(Om) Rx(inc) Rz(om)
    r11 = cO * co - sO * so * ci
    r12 = -cO * so - sO * co * ci
    r21 = sO * co + cO * so * ci
    r22 = -sO * so + cO * co * ci
    r31 = so * si
    r32 = co * si
    out = np.empty((n, 8), dtype=np.float64)
    out[:, 0] = r11 * px + r12 * py
    out[:, 1] = r21 * px + r22 * py
    out[:, 2] = r31 * px + r32 * py
    out[:, 3] = r11 * vx + r12 * vy
    out[:, 4] = r21 * vx + r22 * vy
    out[:, 5] = r31 * vx + r32 * vy
    out[:, 6] = mass if masses is None else masses
    out[:, 7] = radius if radii is None else radii
    return out


def disk_c3_masses(n, seed=0):
    """Masses log-uniform in [1e-11, 1e-9] and radii 1e-4 (m / 1e-10)^(1/3)."""
    rng = np.random.default_rng(seed + 0x5EED)
    m = np.exp(rng.uniform(np.log(1e-11), np.log(1e-9), n))
    R = 1e-4 * np.cbrt(m / 1e-10)
    return m, R


def seeded_ensemble_sizes(count=512, lo=256, hi=8192, seed=12345):
    """System sizes log-uniform in [lo, hi] (rounded), one per system."""
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)


def write_init_txt(path, bodies):
    """Write cartesian initial conditions in the format NcorpiON's populate()
    reads with fscanf("%lf") (structure.c:408-427): 8 numbers per body, %.17g
    so every double round-trips exactly."""
    np.savetxt(path, np.asarray(bodies, dtype=np.float64), fmt="%.17g")
