"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

OBS = {"none": 0, "Lz": 1, "energy": 2, "dipole_r": 3, "dipole_v": 4,
       "norm": 5, "dipole_a": 6}
ENV = {"constant": 0, "trapz": 1, "cos2": 2, "cos4": 3}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        i64p = C.POINTER(C.c_int64)
        ip = C.POINTER(C.c_int)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.c_int, i64p, dp, C.c_char_p, C.c_double, dp,
                                C.c_char_p, C.c_double, C.c_char_p, C.c_int]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_pointwise_parallel.argtypes = [C.c_int]
        L.or_npoints.restype = C.c_int64
        L.or_npoints.argtypes = [C.c_void_p]
        L.or_stages.argtypes = [C.c_void_p]
        L.or_coefficients.argtypes = [C.c_void_p, dp, dp]
        L.or_tables.argtypes = [C.c_void_p, dp, dp, dp, dp, dp]
        L.or_scalings.argtypes = [C.c_void_p, dp, dp, dp]
        L.or_integrator_coefficients.argtypes = [C.c_char_p, C.c_double, dp, dp]
        L.or_fftnd.argtypes = [dp, C.c_int, i64p, C.c_int]
        L.or_step_imag.argtypes = [C.c_void_p, dp]
        L.or_step_length.argtypes = [C.c_void_p, dp, dp]
        L.or_step_velocity.argtypes = [C.c_void_p, dp, dp]
        L.or_gram_schmidt.argtypes = [C.c_void_p, C.c_int, dp]
        L.or_energies.argtypes = [C.c_void_p, dp, C.c_int, dp]
        L.or_obs_width.argtypes = [C.c_int, C.c_int]
        L.or_observables.argtypes = [C.c_void_p, dp, ip, C.c_int, dp, dp, dp]
        L.or_set_mask.argtypes = [C.c_void_p, C.c_double]
        L.or_apply_mask.argtypes = [C.c_void_p, dp]
        L.or_imag_run.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int,
                                  C.c_uint64, dp, dp, dp]
        L.or_initial_guess.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_uint64, dp]
        L.or_imag_step_count.restype = C.c_int64
        L.or_imag_step_count.argtypes = [C.c_double, C.c_double]
        L.or_real_run.argtypes = [C.c_void_p, C.c_int, dp, C.c_int64, dp, dp, dp,
                                  dp, ip, C.c_int, dp]
        L.or_field_vectors.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp, dp,
                                       ip, C.c_double, dp, dp]
        L.or_field_tables.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, ip,
                                      dp, C.c_int, dp, C.c_int64, dp, dp]
        L.or_field_vectors_angle.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp, dp,
                                             ip, dp, C.c_double, dp, dp]
        L.or_field_tables_angle.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, ip, dp,
                                            dp, C.c_int, dp, C.c_int64, dp, dp]
        L.or_final_observables.restype = C.c_int64
        L.or_final_observables.argtypes = [C.c_void_p, dp, dp, dp, C.c_int,
                                           C.c_double, C.c_double, dp, dp, i64p]
        L.or_gaussian.argtypes = [C.c_uint64, C.c_int64, dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def fftnd(a, sign):
    """Unnormalised N-D DFT of a complex128 array (copy), sign -1 or +1."""
    x = np.ascontiguousarray(a, dtype=np.complex128).copy()
    nr = np.array(x.shape, dtype=np.int64)
    lib().or_fftnd(_d(x.view(np.float64)), x.ndim, _i64(nr), int(sign))
    return x


def integrator_coefficients(name, dt):
    a = np.zeros(20)
    b = np.zeros(20)
    s = lib().or_integrator_coefficients(name.encode(), float(dt), _d(a), _d(b))
    return s, a, b


def gaussian(seed, n):
    out = np.zeros(n)
    lib().or_gaussian(C.c_uint64(seed), n, _d(out))
    return out


def imag_step_count(T, dt):
    return lib().or_imag_step_count(float(T), float(dt))


def field_vectors(dim, modes, t):
    """modes: dict of arrays freq, amp, ellip, duration, cep, delay, env and,
    optionally, angle (the polarisation-angle extension, radians)."""
    m = _modes(modes)
    A = np.zeros(dim)
    E = np.zeros(dim)
    args = [_d(m[k]) for k in ("freq", "amp", "ellip", "duration", "cep", "delay")]
    if m["angle"] is None:
        rc = lib().or_field_vectors(dim, m["n"], *args, _i(m["env"]), float(t), _d(A), _d(E))
    else:
        rc = lib().or_field_vectors_angle(dim, m["n"], *args, _i(m["env"]), _d(m["angle"]), float(t),
                                          _d(A), _d(E))
    if rc:
        raise ValueError("bad envelope")
    return A, E


def field_tables(dim, modes, as_scaled, s, time):
    m = _modes(modes)
    time = np.ascontiguousarray(time, dtype=np.float64)
    nt = len(time)
    A = np.zeros(nt * s * dim)
    E = np.zeros(nt * s * dim)
    args = [_d(m[k]) for k in ("freq", "amp", "ellip", "duration", "cep", "delay")]
    asc = _d(np.ascontiguousarray(as_scaled, dtype=np.float64))
    if m["angle"] is None:
        rc = lib().or_field_tables(dim, m["n"], *args, _i(m["env"]), asc, s, _d(time), nt, _d(A), _d(E))
    else:
        rc = lib().or_field_tables_angle(dim, m["n"], *args, _i(m["env"]), _d(m["angle"]), asc, s, _d(time),
                                         nt, _d(A), _d(E))
    if rc:
        raise ValueError("bad envelope")
    return A.reshape(nt, s, dim), E.reshape(nt, s, dim)


def _modes(modes):
    out = {}
    for k in ("freq", "amp", "ellip", "duration", "cep", "delay"):
        out[k] = np.ascontiguousarray(modes[k], dtype=np.float64)
    out["env"] = np.ascontiguousarray(modes["env"], dtype=np.int32)
    out["angle"] = (np.ascontiguousarray(modes["angle"], dtype=np.float64)
                    if modes.get("angle") is not None else None)
    out["n"] = len(out["freq"])
    return out


class Oracle:
    """FP64 CPU restatement of one QuantumND grid + integrator."""

    def __init__(self, nr, dr, potential="H_1D", integrator="BM4", dt=0.1,
                 cap=(0.0, 0.0), softcore=-1.0):
        self.nr = tuple(int(n) for n in nr)
        self.dim = len(self.nr)
        self.shape = self.nr
        nr_a = np.array(self.nr, dtype=np.int64)
        dr_a = np.array(dr if hasattr(dr, "__len__") else [dr] * self.dim, dtype=np.float64)
        self.dr = dr_a
        cap_a = np.array(cap, dtype=np.float64)
        err = C.create_string_buffer(256)
        h = lib().or_create(self.dim, _i64(nr_a), _d(dr_a), potential.encode(),
                            float(softcore), _d(cap_a), integrator.encode(), float(dt),
                            err, 256)
        if not h:
            raise ValueError(err.value.decode())
        self.h = h
        self.N = int(np.prod(self.nr))
        self.s = lib().or_stages(h)
        self.dt = dt

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def coefficients(self):
        a = np.zeros(20)
        b = np.zeros(20)
        lib().or_coefficients(self.h, _d(a), _d(b))
        return a, b

    def tables(self):
        tot = sum(self.nr)
        r = np.zeros(tot)
        p = np.zeros(tot)
        K = np.zeros(self.N)
        Vc = np.zeros(self.N)
        V = np.zeros(self.N)
        lib().or_tables(self.h, _d(r), _d(p), _d(K), _d(Vc), _d(V))
        return r, p, K.reshape(self.shape), Vc.reshape(self.shape), V.reshape(self.shape)

    def scalings(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        lib().or_scalings(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def _c(self, psi):
        x = np.ascontiguousarray(psi, dtype=np.complex128).reshape(self.shape).copy()
        return x

    def step_imag(self, psi, nsteps=1):
        x = self._c(psi)
        for _ in range(nsteps):
            lib().or_step_imag(self.h, _d(x.view(np.float64)))
        return x

    def step_length(self, psi_p, Et):
        x = self._c(psi_p)
        Et = np.ascontiguousarray(Et, dtype=np.float64)
        lib().or_step_length(self.h, _d(x.view(np.float64)), _d(Et))
        return x

    def step_velocity(self, psi, At):
        x = self._c(psi)
        At = np.ascontiguousarray(At, dtype=np.float64)
        lib().or_step_velocity(self.h, _d(x.view(np.float64)), _d(At))
        return x

    def gram_schmidt(self, states, n):
        st = np.ascontiguousarray(states, dtype=np.complex128).copy()
        rc = lib().or_gram_schmidt(self.h, n, _d(st.view(np.float64)))
        return rc, st

    def energies(self, psi, M=2):
        x = self._c(psi)
        E = np.zeros(M)
        lib().or_energies(self.h, _d(x.view(np.float64)), M, _d(E))
        return E

    def observables(self, psi, names, alpha=None, Efield=None):
        x = self._c(psi)
        codes = np.array([OBS[n] for n in names], dtype=np.int32)
        width = sum(lib().or_obs_width(int(c), self.dim) for c in codes)
        out = np.zeros(max(width, 1))
        al = np.zeros(self.dim) if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        Ef = None if Efield is None else np.ascontiguousarray(Efield, dtype=np.float64)
        lib().or_observables(self.h, _d(x.view(np.float64)), _i(codes), len(codes), _d(al),
                             _d(Ef) if Ef is not None else None, _d(out))
        return out[:width]

    def set_mask(self, frac):
        lib().or_set_mask(self.h, float(frac))

    def apply_mask(self, psi):
        x = self._c(psi)
        lib().or_apply_mask(self.h, _d(x.view(np.float64)))
        return x

    def initial_guess(self, n_states, guess_mode=0, seed=0):
        out = np.zeros((n_states,) + self.shape, dtype=np.complex128)
        lib().or_initial_guess(self.h, n_states, guess_mode, C.c_uint64(seed),
                               _d(out.view(np.float64)))
        return out

    def imag_run(self, n_states, T, guess_mode=0, seed=0):
        psi = np.zeros((n_states,) + self.shape, dtype=np.complex128)
        E = np.zeros((n_states, 2))
        Lz = np.zeros(n_states)
        rc = lib().or_imag_run(self.h, n_states, float(T), guess_mode, C.c_uint64(seed),
                               _d(psi.view(np.float64)), _d(E), _d(Lz))
        if rc:
            raise RuntimeError("Gram-Schmidt norm collapse")
        return psi, E, Lz

    def real_run(self, psi, gauge, Et, At, Eobs=None, Aobs=None, names=()):
        x = self._c(psi)
        Et = np.ascontiguousarray(Et, dtype=np.float64)
        At = np.ascontiguousarray(At, dtype=np.float64)
        nt = Et.shape[0]
        Eo = np.zeros((nt, self.dim)) if Eobs is None else np.ascontiguousarray(Eobs, dtype=np.float64)
        Ao = np.zeros((nt, self.dim)) if Aobs is None else np.ascontiguousarray(Aobs, dtype=np.float64)
        codes = np.array([OBS[n] for n in names], dtype=np.int32)
        width = sum(lib().or_obs_width(int(c), self.dim) for c in codes)
        out = np.zeros((nt, max(width, 1)))
        g = {"length": 0, "velocity": 1}[gauge]
        lib().or_real_run(self.h, g, _d(x.view(np.float64)), nt, _d(Et), _d(At), _d(Eo),
                          _d(Ao), _i(codes), len(codes), _d(out))
        return x, out[:, :width]

    def final_observables(self, psi, R_mask, pr_max, Ne, de):
        x = self._c(psi)
        Rm = np.ascontiguousarray(R_mask, dtype=np.float64)
        pm = np.ascontiguousarray(pr_max, dtype=np.float64)
        npr = np.zeros(self.dim, dtype=np.int64)
        ped = np.zeros(Ne)
        cnt = lib().or_final_observables(self.h, _d(x.view(np.float64)), _d(Rm), _d(pm), Ne,
                                         float(de), float(de), _d(ped), None, _i64(npr))
        x2 = self._c(psi)
        pmd = np.zeros(2 * max(cnt, 1))
        lib().or_final_observables(self.h, _d(x2.view(np.float64)), _d(Rm), _d(pm), Ne,
                                   float(de), float(de), None, _d(pmd), _i64(npr))
        return x2, ped, pmd[:2 * cnt].reshape(cnt, 2), npr
