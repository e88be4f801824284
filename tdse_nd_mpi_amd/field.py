"""Laser field tables for the Python entry points (benchmark, tests).

Same definitions as the drop-in executables (tdse_nd_mpi_amd/host/qnd_params.c):
Field_vectors of src/parameters.c:899-949 (sum of modes, A along x with the
ellipticity on y, E = -dA/dt analytically, envelopes constant / trapz / cos2 /
cos4) evaluated at the integrator stage times of src/integrators.c:94-101.
"""
import numpy as np

AU_TIME_NM = 45.5637639947958      # omega [a.u.] = 45.56.../lambda [nm]  (src/parameters.c:809)
AU_FIELD = 5.33803e-9              # F0 [a.u.] = 5.33803e-9 sqrt(I [W/cm^2]) (src/parameters.c:808)
STAGES = {"Verlet": [.5],
          "BM4": [.0792036964311957, .1303114101821663, .2228614958676077,
                  -.3667132690474257, .3246481886897062, .1096884778767498],
          "BM6": [.050262764400392, .098553683500650, .314960616927694, -.447346482695478,
                  .492426372489876, -.425118767797691, .237063913978122, .195602488600053,
                  .346358189850727, -.362762779254345],
          "RKN4": [.082984406417405, .162314550766866, .233995250731502, .370877414979578,
                   -.409933719901926, .059762097006575]}


def coefficients(name, dt):
    first = STAGES[name]
    a = np.array(first + first[::-1])
    return len(first), a, a * dt


class Mode:
    """One laser mode.  `angle` (radians, extension: the reference fixes the
    major axis on x, src/parameters.c:942-947) rotates the mode's (x, y)
    field vector in the xy plane."""

    def __init__(self, intensity=1e14, wavelength=800.0, ellipticity=0.0, cep=0.0, delay=0.0,
                 envelope="cos2", duration=None, cycles=None, angle=0.0):
        self.intensity, self.wavelength = intensity, wavelength
        self.w = AU_TIME_NM / wavelength
        self.F0 = AU_FIELD * np.sqrt(intensity)
        self.xi = ellipticity
        self.phi = cep
        self.delay = delay
        self.env = envelope
        self.angle = angle
        if duration is None:
            duration = (cycles or 4) * 2 * np.pi / self.w
        self.tau = duration

    def envelope(self, ts):
        tau, w = self.tau, self.w
        if self.env == "constant":
            return np.ones_like(ts), np.zeros_like(ts)
        if self.env == "cos2":
            inside = (ts >= 0) & (ts <= tau)
            return np.sin(np.pi * ts / tau) ** 2 * inside, np.pi * np.sin(2 * np.pi * ts / tau) / tau * inside
        if self.env == "cos4":
            inside = (ts >= 0) & (ts <= tau)
            s = np.sin(np.pi * ts / tau)
            return s ** 4 * inside, 4 * np.pi * np.cos(np.pi * ts / tau) * s ** 3 / tau * inside
        if self.env == "trapz":
            tr = 2 * 2 * np.pi / w
            up = (ts >= 0) & (ts <= tr)
            flat = (ts > tr) & (ts < tau - tr)
            down = (ts >= tau - tr) & (ts <= tau)
            f = np.sin(.5 * np.pi * ts / tr) ** 4 * up + flat + np.sin(.5 * np.pi * (ts - tau) / tr) ** 4 * down
            df = np.pi / tr * (np.sin(.5 * np.pi * ts / tr) ** 2 * np.sin(np.pi * ts / tr) * up
                               + np.sin(.5 * np.pi * (ts - tau) / tr) ** 2 * np.sin(np.pi * (ts - tau) / tr) * down)
            return f, df
        raise ValueError(f"Unknown envelope: {self.env}")


def field_vectors(modes, t, dim):
    """A(t), E(t) arrays of shape t.shape + (dim,)."""
    t = np.asarray(t, dtype=np.float64)
    A = np.zeros(t.shape + (dim,))
    E = np.zeros(t.shape + (dim,))
    for m in modes:
        ts = t - m.delay
        f, df = m.envelope(ts)
        arg = m.w * ts + m.phi
        ex = -m.F0 * (df * np.cos(arg) - f * m.w * np.sin(arg)) / m.w
        ax = m.F0 * f * np.cos(arg) / m.w
        ey = -m.F0 * m.xi * (df * np.sin(arg) + f * m.w * np.cos(arg)) / m.w
        ay = m.F0 * m.xi * f * np.sin(arg) / m.w
        ca, sa = np.cos(m.angle), np.sin(m.angle)
        E[..., 0] += ca * ex - sa * ey
        A[..., 0] += ca * ax - sa * ay
        if dim >= 2:
            E[..., 1] += sa * ex + ca * ey
            A[..., 1] += sa * ax + ca * ay
    return A, E


def stage_times(integrator, dt, ntime):
    s, a, asc = coefficients(integrator, dt)
    time = np.arange(ntime) * dt
    ts = np.zeros((ntime, s))
    t = time + asc[0]                 # same sequential accumulation as the reference
    for j in range(s):
        ts[:, j] = t
        if j < s - 1:
            t = t + (asc[2 * j + 1] + asc[2 * j + 2])
    return time, ts


def tables(modes, integrator, dt, ntime, dim):
    """Et, At [ntime][s][dim] at stage times; Eobs, Aobs [ntime][dim] at t_n."""
    time, ts = stage_times(integrator, dt, ntime)
    At, Et = field_vectors(modes, ts, dim)
    Ao, Eo = field_vectors(modes, time, dim)
    return Et, At, Eo, Ao
