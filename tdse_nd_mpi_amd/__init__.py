"""tdse_nd_mpi_amd -- MI355X split-operator TDSE hot path of QuantumND.

Python view of the C ABI in include/qnd.h (libqnd.so, built in-tree from
tdse_nd_mpi_amd/csrc).  There is no CPU fallback: importing works anywhere,
but creating a Grid needs the HIP library and a GPU, and fails loudly
otherwise.

The methods mirror the reference's hot-path functions (src/integrators.c,
src/calculate_observables.c, src/QuantumND_imag.c, src/QuantumND_real.c);
each docstring names the one it replaces.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# QND_LIBRARY selects another build of the library (e.g. a variant built
# with different compile-time options); default: the in-tree libqnd.so
LIB_PATH = os.environ.get("QND_LIBRARY") or os.path.join(HERE, "libqnd.so")
BIN_DIR = os.path.join(HERE, "bin")

C128, C64, C64F64 = 0, 1, 2
GAUGES = {"length": 0, "velocity": 1}
OBS = {"none": 0, "Lz": 1, "energy": 2, "dipole_r": 3, "dipole_v": 4, "norm": 5, "dipole_a": 6}
PRECISIONS = {"c128": C128, "complex128": C128, "fp64": C128, "c64": C64, "complex64": C64, "mixed": C64,
              "c64f64": C64F64}

EXPORTS = ("qnd_create", "qnd_destroy", "qnd_last_error", "qnd_set_potential", "qnd_set_integrator",
           "qnd_set_mask", "qnd_upload_state", "qnd_download_state", "qnd_imag_steps",
           "qnd_gram_schmidt", "qnd_transform", "qnd_energy_moments", "qnd_observables", "qnd_set_field",
           "qnd_real_steps", "qnd_final_observables", "qnd_event_record", "qnd_event_elapsed",
           "qnd_synchronize", "qnd_profile_enable", "qnd_profile_read", "qnd_version",
           "qnd_set_field_batch", "qnd_real_steps_batch", "qnd_set_tuning", "qnd_forward_checkpointed",
           "qnd_adjoint", "qnd_observables_noalloc", "qnd_group_real_steps", "qnd_set_field_modes",
           "qnd_get_field_batch")
ENVELOPES = {"constant": 0, "trapz": 1, "cos2": 2, "cos4": 3}


class ModeStruct(C.Structure):
    """struct qnd_mode of include/qnd.h (one laser mode of Field_vectors)."""
    _fields_ = [("intensity", C.c_double), ("wavelength", C.c_double), ("ellipticity", C.c_double),
                ("cep", C.c_double), ("delay", C.c_double), ("duration", C.c_double), ("angle", C.c_double),
                ("envelope", C.c_int), ("pad_", C.c_int)]

_lib = None


class QNDError(RuntimeError):
    pass


def library():
    """Load libqnd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QNDError(f"{LIB_PATH} is not built: run `make -C tdse_nd_mpi_amd` "
                       "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    i64p = C.POINTER(C.c_int64)
    vp = C.c_void_p
    L.qnd_create.argtypes = [C.POINTER(vp), C.c_int, i64p, dp, C.c_int, C.c_int]
    L.qnd_destroy.argtypes = [vp]
    L.qnd_destroy.restype = None
    L.qnd_last_error.argtypes = [vp]
    L.qnd_last_error.restype = C.c_char_p
    L.qnd_set_potential.argtypes = [vp, C.c_char_p, C.c_double, dp]
    L.qnd_set_integrator.argtypes = [vp, C.c_char_p, C.c_double]
    L.qnd_set_mask.argtypes = [vp, C.c_double]
    L.qnd_upload_state.argtypes = [vp, C.c_int, dp]
    L.qnd_download_state.argtypes = [vp, C.c_int, dp]
    L.qnd_imag_steps.argtypes = [vp, C.c_int, C.c_int64, C.c_int]
    L.qnd_gram_schmidt.argtypes = [vp, C.c_int]
    L.qnd_transform.argtypes = [vp, C.c_int, C.c_int]
    L.qnd_energy_moments.argtypes = [vp, C.c_int, C.c_int, dp]
    L.qnd_observables.argtypes = [vp, C.c_int, ip, C.c_int, dp, dp, dp]
    L.qnd_observables_noalloc.argtypes = [vp, vp, ip, C.c_int, dp, vp]
    L.qnd_set_field.argtypes = [vp, C.c_int64, dp, dp, dp, dp]
    L.qnd_real_steps.argtypes = [vp, C.c_int, C.c_int, C.c_int64, C.c_int64, ip, C.c_int, dp]
    L.qnd_set_field_batch.argtypes = [vp, C.c_int, C.c_int64, dp, dp, dp]
    L.qnd_set_field_modes.argtypes = [vp, C.c_int, C.c_int64, C.c_int, C.POINTER(ModeStruct)]
    L.qnd_get_field_batch.argtypes = [vp, dp, dp, dp]
    L.qnd_set_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.qnd_forward_checkpointed.argtypes = [vp, C.c_int, C.c_int64, C.c_int64, dp, dp]
    L.qnd_adjoint.argtypes = [vp, C.c_int64, dp, C.c_double, dp, dp, dp]
    L.qnd_real_steps_batch.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64, ip, C.c_int, dp]
    L.qnd_group_real_steps.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int64, C.c_int64, ip, C.c_int,
                                       C.POINTER(dp), C.c_int]
    L.qnd_final_observables.argtypes = [vp, C.c_int, dp, dp, C.c_int, C.c_double, dp, dp]
    L.qnd_event_record.argtypes = [vp, C.c_int]
    L.qnd_event_elapsed.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_float)]
    L.qnd_synchronize.argtypes = [vp]
    L.qnd_profile_enable.argtypes = [vp, C.c_int]
    L.qnd_profile_read.argtypes = [vp, i64p, dp, C.c_int]
    L.qnd_version.restype = C.c_char_p
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _opt_d(a):
    return None if a is None else _d(a)


class Grid:
    """One N-D grid on one GPU with `n_slots` state vectors.

    Replaces Initialize_meshgrid (src/meshgrid.c:17-179) and the FFTW plans.
    """

    def __init__(self, nr, dr, precision="c128", n_slots=1):
        L = library()
        self.nr = tuple(int(n) for n in nr)
        self.dim = len(self.nr)
        self.shape = self.nr
        drs = dr if hasattr(dr, "__len__") else [dr] * self.dim
        self.dr = np.array(drs, dtype=np.float64)
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self.n_slots = n_slots
        self.N = int(np.prod(self.nr))
        h = C.c_void_p()
        nr_a = np.array(self.nr, dtype=np.int64)
        rc = L.qnd_create(C.byref(h), self.dim, nr_a.ctypes.data_as(C.POINTER(C.c_int64)),
                          _d(self.dr), self.precision, n_slots)
        if rc:
            raise QNDError(L.qnd_last_error(None).decode())
        self.h = h
        self.s = 0

    def close(self):
        if getattr(self, "h", None):
            library().qnd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc:
            raise QNDError(library().qnd_last_error(self.h).decode())

    # -- set-up -------------------------------------------------------
    def set_potential(self, name, softcore=-1.0, cap=None):
        """Vpot/Vcap of src/meshgrid.c:123-154 (analytic on the device)."""
        c = None if cap is None else np.array(cap, dtype=np.float64)
        self._ck(library().qnd_set_potential(self.h, name.encode(), float(softcore), _opt_d(c)))

    def set_integrator(self, name, dt):
        """Initialize_integrator (src/integrators.c:13-71)."""
        self._ck(library().qnd_set_integrator(self.h, name.encode(), float(dt)))
        self.s = {"Verlet": 1, "BM4": 6, "BM6": 10, "RKN4": 6}[name]
        self.dt = dt

    def set_mask(self, frac):
        self._ck(library().qnd_set_mask(self.h, float(frac)))

    # -- states -------------------------------------------------------
    def upload(self, slot, psi):
        x = np.ascontiguousarray(psi, dtype=np.complex128).reshape(self.shape)
        self._ck(library().qnd_upload_state(self.h, slot, _d(x.view(np.float64))))

    def download(self, slot):
        x = np.zeros(self.shape, dtype=np.complex128)
        self._ck(library().qnd_download_state(self.h, slot, _d(x.view(np.float64))))
        return x

    # -- imaginary time -----------------------------------------------
    def imag_steps(self, slot, nsteps, gram_schmidt=True):
        """Integration_step_imag + Gram_Schmidt_orthonormalization, nsteps times
        (src/integrators.c:135-163, src/QuantumND_imag.c:182-185,265-311)."""
        self._ck(library().qnd_imag_steps(self.h, slot, int(nsteps), 1 if gram_schmidt else 0))

    def transform(self, slot, sign):
        """fftw_execute of the forward (-1) / backward (+1) plan on a slot."""
        self._ck(library().qnd_transform(self.h, slot, int(sign)))

    def gram_schmidt(self, slot):
        self._ck(library().qnd_gram_schmidt(self.h, slot))

    def energy_moments(self, slot, M=2):
        """Calculation_global_energies (src/calculate_observables.c:89-116)."""
        E = np.zeros(M)
        self._ck(library().qnd_energy_moments(self.h, slot, M, _d(E)))
        return E

    def observables(self, slot, names, alpha=None, Efield=None):
        """Calculation_global_observables (src/calculate_observables.c:119-140)."""
        codes = np.array([OBS[n] for n in names], dtype=np.int32)
        width = sum({1: 1, 2: 1, 5: 1, 3: self.dim, 4: self.dim, 6: self.dim}.get(int(c), 0)
                    for c in codes)
        out = np.zeros(max(width, 1))
        al = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        Ef = None if Efield is None else np.ascontiguousarray(Efield, dtype=np.float64)
        self._ck(library().qnd_observables(self.h, slot, codes.ctypes.data_as(C.POINTER(C.c_int)),
                                           len(codes), _opt_d(al), _opt_d(Ef), _d(out)))
        return out[:width]

    def observables_noalloc(self, psi_dev, names, out_dev, Efield=None):
        """qnd_observables_noalloc: norm / dipole_r / dipole_a of the device
        state at address psi_dev into the device array at out_dev, without
        allocating; ordered on the context's stream (synchronize() to wait)."""
        codes = np.array([OBS[n] for n in names], dtype=np.int32)
        Ef = None if Efield is None else np.ascontiguousarray(Efield, dtype=np.float64)
        self._ck(library().qnd_observables_noalloc(self.h, C.c_void_p(int(psi_dev)),
                                                   codes.ctypes.data_as(C.POINTER(C.c_int)), len(codes),
                                                   _opt_d(Ef), C.c_void_p(int(out_dev))))

    # -- real time ----------------------------------------------------
    def _check_shape(self, name, a, shape):
        if tuple(a.shape) != tuple(shape):
            raise QNDError(f"{name} has shape {tuple(a.shape)}, expected {tuple(shape)} "
                           f"(the integrator has {self.s} stages)")

    def set_field(self, Et, At, Eobs=None, Aobs=None):
        """Field tables [ntime][s][dim] (src/integrators.c:74-129)."""
        if not self.s:
            raise QNDError("integrator not set (set_integrator) before the field tables")
        Et = np.ascontiguousarray(Et, dtype=np.float64)
        At = np.ascontiguousarray(At, dtype=np.float64)
        nt = Et.shape[0]
        Eo = np.zeros((nt, self.dim)) if Eobs is None else np.ascontiguousarray(Eobs, dtype=np.float64)
        Ao = np.zeros((nt, self.dim)) if Aobs is None else np.ascontiguousarray(Aobs, dtype=np.float64)
        self._check_shape("Et", Et, (nt, self.s, self.dim))
        self._check_shape("At", At, (nt, self.s, self.dim))
        self._check_shape("Eobs", Eo, (nt, self.dim))
        self._check_shape("Aobs", Ao, (nt, self.dim))
        self._ck(library().qnd_set_field(self.h, nt, _d(Et), _d(At), _d(Eo), _d(Ao)))
        self.ntime = nt

    def real_steps(self, slot, gauge, step0, nsteps, names=()):
        """Observables + Integration_step_length/velocity, nsteps times
        (src/QuantumND_real.c:169-186,195-223)."""
        codes = np.array([OBS[n] for n in names], dtype=np.int32)
        width = sum({1: 1, 2: 1, 5: 1, 3: self.dim, 4: self.dim, 6: self.dim}.get(int(c), 0)
                    for c in codes)
        out = np.zeros((max(nsteps, 1), max(width, 1)))
        self._ck(library().qnd_real_steps(
            self.h, slot, GAUGES[gauge], int(step0), int(nsteps),
            codes.ctypes.data_as(C.POINTER(C.c_int)) if len(codes) else None, len(codes),
            _d(out) if width else None))
        return out[:nsteps, :width]

    def set_tuning(self, vpt=0, tw_contig=0, tw_strided=0):
        """Launch geometry overrides (0 = automatic); see tdse_nd_mpi_amd.tuning."""
        self._ck(library().qnd_set_tuning(self.h, vpt, tw_contig, tw_strided))
        self.tuning = (vpt, tw_contig, tw_strided)

    def forward_checkpointed(self, slot, nsteps, every):
        """Length-gauge steps 0..nsteps-1 with checkpoints every `every` steps;
        returns (dipole acceleration [nsteps][dim], norms [nsteps], final norm)."""
        out = np.zeros((nsteps, self.dim + 1))
        nf = C.c_double(0.0)
        self._ck(library().qnd_forward_checkpointed(self.h, slot, int(nsteps), int(every), _d(out),
                                                    C.cast(C.pointer(nf), C.POINTER(C.c_double))))
        return out[:, :self.dim], out[:, self.dim], nf.value

    def adjoint(self, dL_da, dL_dnorm):
        """Back-propagate dL/da [nsteps][dim] and dL/dnorm_final through the
        last forward_checkpointed run: (dL/dEt [nsteps][s][dim],
        dL/dEobs [nsteps][dim], dL/d soft-core length)."""
        dL_da = np.ascontiguousarray(dL_da, dtype=np.float64)
        if dL_da.ndim != 2 or dL_da.shape[1] != self.dim:
            raise QNDError(f"dL_da has shape {dL_da.shape}, expected (nsteps, {self.dim})")
        nsteps = dL_da.shape[0]
        dEt = np.zeros((nsteps, self.s, self.dim))
        dEo = np.zeros((nsteps, self.dim))
        ds = C.c_double(0.0)
        self._ck(library().qnd_adjoint(self.h, nsteps, _d(dL_da), float(dL_dnorm), _d(dEt), _d(dEo),
                                       C.cast(C.pointer(ds), C.POINTER(C.c_double))))
        return dEt, dEo, ds.value

    def set_field_batch(self, Et, At, Eobs=None):
        """Per-instance field tables: Et, At [nb][ntime][s][dim], Eobs [nb][ntime][dim]."""
        if not self.s:
            raise QNDError("integrator not set (set_integrator) before the field tables")
        Et = np.ascontiguousarray(Et, dtype=np.float64)
        At = np.ascontiguousarray(At, dtype=np.float64)
        nb, nt = Et.shape[0], Et.shape[1]
        Eo = np.zeros((nb, nt, self.dim)) if Eobs is None else np.ascontiguousarray(Eobs, dtype=np.float64)
        self._check_shape("Et", Et, (nb, nt, self.s, self.dim))
        self._check_shape("At", At, (nb, nt, self.s, self.dim))
        self._check_shape("Eobs", Eo, (nb, nt, self.dim))
        self._ck(library().qnd_set_field_batch(self.h, nb, nt, _d(Et), _d(At), _d(Eo)))
        self.batch = nb
        self.field_ntime = nt

    def set_field_modes(self, modes, ntime):
        """Field tables formed on the device (qnd_set_field_modes): `modes` is
        one list of field.Mode per instance (nb = len(modes), the same number
        of modes each); nb == 1 also sets the single-instance tables."""
        if not self.s:
            raise QNDError("integrator not set (set_integrator) before the field")
        nb = len(modes)
        nm = len(modes[0]) if nb else 0
        if any(len(m) != nm for m in modes):
            raise QNDError("every instance needs the same number of modes")
        arr = (ModeStruct * max(nb * nm, 1))()
        for b, ms in enumerate(modes):
            for k, m in enumerate(ms):
                if m.env not in ENVELOPES:
                    raise QNDError(f"Unknown envelope: {m.env}")
                arr[b * nm + k] = ModeStruct(m.intensity, m.wavelength, m.xi, m.phi, m.delay, m.tau, m.angle,
                                             ENVELOPES[m.env], 0)
        self._ck(library().qnd_set_field_modes(self.h, nb, int(ntime), nm, arr))
        self.batch = nb
        self.field_ntime = int(ntime)
        if nb == 1:
            self.ntime = int(ntime)

    def field_batch(self):
        """The batch field tables on the device: Et, At [nb][ntime][s][dim],
        Eobs [nb][ntime][dim]."""
        nb, nt = self.batch, self.field_ntime
        Et = np.zeros((nb, nt, self.s, self.dim))
        At = np.zeros((nb, nt, self.s, self.dim))
        Eo = np.zeros((nb, nt, self.dim))
        self._ck(library().qnd_get_field_batch(self.h, _d(Et), _d(At), _d(Eo)))
        return Et, At, Eo

    def real_steps_batch(self, slot0, nb, gauge, step0, nsteps, names=()):
        """real_steps for slots slot0..slot0+nb-1 in one launch per pass, each
        under its own field; returns [nb][nsteps][width]."""
        codes = np.array([OBS[n] for n in names], dtype=np.int32)
        width = sum({5: 1, 3: self.dim, 6: self.dim}.get(int(c), 0) for c in codes)
        out = np.zeros((nb, max(nsteps, 1), max(width, 1)))
        self._ck(library().qnd_real_steps_batch(
            self.h, slot0, nb, GAUGES[gauge], int(step0), int(nsteps),
            codes.ctypes.data_as(C.POINTER(C.c_int)) if len(codes) else None, len(codes),
            _d(out) if width else None))
        return out[:, :nsteps, :width]

    def final_observables(self, slot, R_mask, pr_max, Ne, de):
        """Write_final_observables numerics (src/io_functions.c:136-287)."""
        Rm = np.ascontiguousarray(R_mask, dtype=np.float64)
        pm = np.ascontiguousarray(pr_max, dtype=np.float64)
        ped = np.zeros(max(Ne, 1))
        psi_p = np.zeros(self.shape, dtype=np.complex128)
        self._ck(library().qnd_final_observables(self.h, slot, _d(Rm), _d(pm), int(Ne), float(de),
                                                 _d(ped), _d(psi_p.view(np.float64))))
        return ped[:Ne], psi_p

    # -- timing ---------------------------------------------------------
    def event_record(self, idx):
        self._ck(library().qnd_event_record(self.h, idx))

    def event_elapsed(self, i0, i1):
        ms = C.c_float()
        self._ck(library().qnd_event_elapsed(self.h, i0, i1, C.byref(ms)))
        return ms.value

    def synchronize(self):
        self._ck(library().qnd_synchronize(self.h))

    def profile(self, enable):
        self._ck(library().qnd_profile_enable(self.h, 1 if enable else 0))

    def profile_read(self):
        cnt = np.zeros(5, dtype=np.int64)
        ms = np.zeros(5)
        self._ck(library().qnd_profile_read(self.h, cnt.ctypes.data_as(C.POINTER(C.c_int64)), _d(ms), 5))
        return cnt, ms


def group_real_steps(grids, gauge, step0, nsteps, names=(), use_graph=True, branches=False):
    """qnd_group_real_steps: every grid's batch (its slots 0 .. nb-1 under its
    set_field_batch tables) stepped together: one kernel launch per pass over
    all grids, or (branches) each grid's own kernels on its own stream, forked
    and joined every step; use_graph replays captured HIP graphs.  Returns one
    [nb][nsteps][width] record array per grid."""
    L = library()
    codes = np.array([OBS[n] for n in names], dtype=np.int32)
    outs, ptrs = [], []
    for g in grids:
        width = sum({5: 1, 3: g.dim, 6: g.dim}.get(int(c), 0) for c in codes)
        o = np.zeros((g.batch, max(nsteps, 1), max(width, 1)))
        outs.append((o, width))
        ptrs.append(_d(o))
    hs = (C.c_void_p * len(grids))(*[g.h.value if isinstance(g.h, C.c_void_p) else g.h for g in grids])
    pp = (C.POINTER(C.c_double) * len(grids))(*ptrs)
    rc = L.qnd_group_real_steps(hs, len(grids), GAUGES[gauge], int(step0), int(nsteps),
                                codes.ctypes.data_as(C.POINTER(C.c_int)) if len(codes) else None, len(codes),
                                pp, (1 if use_graph else 0) | (2 if branches else 0))
    if rc:
        raise QNDError(L.qnd_last_error(grids[0].h).decode())
    return [o[:, :nsteps, :w] for o, w in outs]


def imag_step_count(T, dt):
    """Steps of the reference loop for(time=0;time<T;time+=dt) (src/QuantumND_imag.c:182)."""
    n, t = 0, 0.0
    while t < T:
        n += 1
        t += dt
    return n
