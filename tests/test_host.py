"""CPU tests of the drop-in boundary: libqnd exports every symbol of
include/qnd.h, refuses to run without a GPU, and the QuantumND_imag /
QuantumND_real executables reproduce the reference's parameter handling
(src/parameters.c), error exits and the field files written before any GPU
work (src/integrators.c:74-129), which are compared with the oracle.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import tdse_nd_mpi_amd as q

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IMAG = os.path.join(q.BIN_DIR, "QuantumND_imag")
REAL = os.path.join(q.BIN_DIR, "QuantumND_real")


def have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "qnd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qnd_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = q.library()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(q.EXPORTS)
    assert L.qnd_version().startswith(b"libqnd")


def test_mode_struct_matches_header(tmp_path):
    """The ctypes mirror of struct qnd_mode has the C layout (size and
    offsets, from a C program compiled against include/qnd.h)."""
    import ctypes
    src = tmp_path / "m.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "qnd.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu\\n", sizeof(qnd_mode), offsetof(qnd_mode, duration),'
                   ' offsetof(qnd_mode, angle), offsetof(qnd_mode, envelope)); return 0;}\n')
    exe = tmp_path / "m"
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    M = q.ModeStruct
    assert got == [ctypes.sizeof(M), M.duration.offset, M.angle.offset, M.envelope.offset]
    assert q.ENVELOPES == {"constant": 0, "trapz": 1, "cos2": 2, "cos4": 3}


def test_constants_match_header(tmp_path):
    """The Python binding's precision, gauge and observable codes are the
    enum values of include/qnd.h (a C program compiled against it prints
    them)."""
    src = tmp_path / "c.c"
    src.write_text('#include <stdio.h>\n#include "qnd.h"\n'
                   'int main(void){printf("%d %d %d %d %d %d %d %d %d %d\\n", QND_C128, QND_C64, QND_C64F64,'
                   ' QND_GAUGE_LENGTH, QND_GAUGE_VELOCITY, QND_OBS_DIPOLE_R, QND_OBS_DIPOLE_V, QND_OBS_NORM,'
                   ' QND_OBS_DIPOLE_A, QND_OBS_LZ); return 0;}\n')
    exe = tmp_path / "c"
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert got == [q.PRECISIONS["c128"], q.PRECISIONS["c64"], q.PRECISIONS["c64f64"],
                   q.GAUGES["length"], q.GAUGES["velocity"], q.OBS["dipole_r"], q.OBS["dipole_v"], q.OBS["norm"],
                   q.OBS["dipole_a"], q.OBS["Lz"]]


@pytest.mark.skipif(have_gpu(), reason="checks the no-GPU failure path")
def test_create_fails_loudly_without_gpu():
    with pytest.raises(q.QNDError, match="no HIP device"):
        q.Grid((64,), 0.1)


def test_create_rejects_bad_grids():
    with pytest.raises(q.QNDError, match="power of two"):
        q.Grid((40,), 0.1)
    with pytest.raises(q.QNDError, match="power of two"):
        q.Grid((6,), 0.1)          # 3 * 2^k needs k >= 2
    with pytest.raises(q.QNDError, match="dimension"):
        q.Grid((8,) * 5, 0.1)


def run(exe, args, cwd):
    p = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=60)
    return p.returncode, p.stdout + p.stderr


IMAG_OK = """OUTPUT_IMAG \toutput.imag

DIMENSION\t1
N_STATES \t2

POTENTIAL\tH_1D

DR\t\t.2
NR\t\t256

INTEGRATOR \tBM4
DT \t\t.1
TINTEGRATION \t2.
"""


@pytest.mark.parametrize("text,args,msg", [
    ("FOO 1\n" + IMAG_OK, [], "Unknown parameter: FOO"),
    # an unknown word after a key's values is an extra argument of that key
    (IMAG_OK + "FOO 1\n", [], "Parameter 'TINTEGRATION' expected exactly 1 argument(s), but found extra"),
    (IMAG_OK.replace("NR\t\t256", "NR\t\t250"), [], "must be a power of 2"),
    (IMAG_OK.replace("NR\t\t256", "NR\t\t256 256"), [], "expected exactly 1 argument(s), but found extra"),
    (IMAG_OK.replace("N_STATES \t2", "N_STATES"), [], "expected 1 argument(s), but found only 0 (hit new parameter"),
    (IMAG_OK.replace("DIMENSION\t1", ""), [], "The dimension is not specified"),
    (IMAG_OK, ["CEP", "1"], "Unrecognized parameter key 'CEP'"),
    (IMAG_OK, ["DT"], "found without a corresponding value"),
    (IMAG_OK.replace("BM4", "BM5"), [], "Unknown integrator: BM5"),
    (IMAG_OK.replace("H_1D", "He_1D"), [], "Unknown potential: He_1D"),
    (IMAG_OK.replace("N_STATES \t2", "N_STATES \t0"), [], "Could not read or interpret"),
    (IMAG_OK + "PRECISION\tc32\n", [], "Unknown precision: c32 (c128, c64 or c64f64)"),
])
def test_imag_error_paths(tmp_path, text, args, msg):
    (tmp_path / "params.imag").write_text(text)
    rc, out = run(IMAG, ["params.imag"] + args, tmp_path)
    assert rc == 1, out
    assert msg in out, out


def test_imag_usage(tmp_path):
    rc, out = run(IMAG, [], tmp_path)
    assert rc == 1 and "Usage:" in out


@pytest.mark.skipif(have_gpu(), reason="checks the no-GPU failure path")
@pytest.mark.parametrize("prec", ["", "PRECISION\tc64\n", "PRECISION\tc64f64\n"])
def test_imag_parses_then_needs_gpu(tmp_path, prec):
    (tmp_path / "params.imag").write_text(IMAG_OK + prec)
    rc, out = run(IMAG, ["params.imag", "DT", "0.05", "OUTPUT_IMAG", "x.imag"], tmp_path)
    assert rc == 1
    assert "DT\t\t\t0.050000000000000" in out and "OUTPUT_IMAG\t\tx.imag" in out
    assert "no HIP device available" in out


def write_output_imag(path, dim, nr, dr, pot, states, E, DE):
    with open(path, "w") as f:
        f.write(f"DIMENSION\t{dim}\nN_STATES\t{len(states)}\nNR")
        for n in nr:
            f.write(f"\t{n}")
        f.write("\nDR")
        for d in dr:
            f.write(f"\t{d:16.15f}")
        f.write(f"\nPOTENTIAL\t{pot}\nE\t")
        for e in E:
            f.write(f"\t{e:16.15f}")
        f.write("\nDE")
        for e in DE:
            f.write(f"\t{e:16.15f}")
        f.write("\nPSI\n")
    with open(path, "ab") as f:
        for s in states:
            f.write(np.ascontiguousarray(s, dtype=np.complex128).tobytes())


REAL_OK = """OUTPUT_IMAG\toutput.imag

NR \t\t128\t64

N_STATES \t2
J_STATES \t1 \t2
COEFFS_STATES\t1.\t.5+.5j

CAP \t\t.1\t5.

GAUGE \t\tlength
N_MODES \t2
INTENSITY \t1.e14 \t1.e13
LAMBDA \t\t800 400
ELLIPTICITY \t0. \t.3
CEP \t\t0. \t1.
DELAY \t\t0. \t5.
ENVELOPE \tcos2 \ttrapz
DURATION \t200.\t150.

INTEGRATOR \tBM4
DT \t\t.05
TINTEGRATION \t3.

DT_SAVE \t1.
N_OBSERVABLES \t4
OBSERVABLES \tdipole_r \tPMD\tenergy \tLz

P_MAX \t\t2.\t2.
R_MASK \t\t5. \t10.
DE \t\t.01
NE \t\t50
"""


@pytest.fixture
def real_dir(tmp_path):
    nr, dr = (64, 32), (0.4, 0.5)
    st = np.zeros((2,) + nr, dtype=complex)
    st[0, 32, 16] = 1.0
    st[1, 31, 16] = 1.0
    write_output_imag(tmp_path / "output.imag", 2, nr, dr, "Ne_2D", st, [-1.0, -0.5], [0.0, 0.0])
    return tmp_path


@pytest.mark.parametrize("edit,args,msg", [
    (("OUTPUT_IMAG\toutput.imag", ""), [], "Critical parameters missing"),
    (("J_STATES \t1 \t2", "J_STATES \t1 \t3"), [], "failed to read"),
    (("DELAY \t\t0. \t5.", "DELAY \t\t0. \t-5."), [], "failed to read"),
    (("OUTPUT_IMAG\toutput.imag", "DIMENSION 2\nOUTPUT_IMAG\toutput.imag"), [], "Unknown parameter: DIMENSION"),
    (("NR \t\t128\t64", "NR \t\t128\t60"), [], "must be a power of 2"),
    (("ENVELOPE \tcos2 \ttrapz", "ENVELOPE \tcos2 \tgauss"), [], "Unknown envelope: gauss"),
    (("INTENSITY \t1.e14 \t1.e13", "INTENSITY \t1.e14"), [], "expected 2 argument(s), but found only 1"),
    (("", ""), ["NR", "64"], "Unrecognized parameter key 'NR'"),
    (("", ""), ["CEP"], "found without a corresponding value"),
    (("BM4", "BM7"), [], "Unknown integrator: BM7"),
])
def test_real_error_paths(real_dir, edit, args, msg):
    text = REAL_OK.replace(*edit) if edit[0] else REAL_OK
    (real_dir / "params.real").write_text(text)
    rc, out = run(REAL, ["params.real"] + args, real_dir)
    assert rc == 1, out
    assert msg in out, out


def test_real_field_files_match_oracle(real_dir):
    (real_dir / "params.real").write_text(REAL_OK)
    rc, out = run(REAL, ["params.real", "CEP", "0.25", "INTENSITY", "2e14"], real_dir)
    if not have_gpu():
        assert rc == 1 and "no HIP device available" in out
    assert "COEFFS_STATES\t\t0.816496580927726+I*0.000000000000000\t0.408248290463863+I*0.408248290463863" in out
    Ef = np.loadtxt(real_dir / "electric_field.real")
    Af = np.loadtxt(real_dir / "vector_potential.real")
    s, a, asc = O.integrator_coefficients("BM4", 0.05)
    nt = int(3.0 / 0.05)
    time = np.arange(nt) * 0.05
    lam = np.array([800., 400.])
    modes = dict(freq=45.5637639947958 / lam, amp=5.33803e-9 * np.sqrt([2e14, 1e13]),
                 ellip=[0., .3], duration=[200., 150.], cep=[0.25, 1.], delay=[0., 5.],
                 env=[O.ENV["cos2"], O.ENV["trapz"]])
    A, E = O.field_tables(2, modes, asc, s, time)
    assert Ef.shape == (nt * s, 3)
    assert np.allclose(Ef[:, 1:], E.reshape(-1, 2), rtol=0, atol=1e-15)
    assert np.allclose(Af[:, 1:], A.reshape(-1, 2), rtol=0, atol=1e-15)
    # header of observables.real, with the reference's dimensions[k] indexing
    hdr = (real_dir / "observables.real").read_text().splitlines()[0]
    assert hdr == "time\tdipole_r\tenergy\tLz"


def test_parse_complex_cases(real_dir):
    for text, expect in [("1.", "1.000000000000000+I*0.000000000000000"),
                         ("1j", "0.000000000000000+I*1.000000000000000"),
                         ("2-1j", "0.894427190999916+I*-0.447213595499958")]:
        (real_dir / "params.real").write_text(
            REAL_OK.replace("N_STATES \t2", "N_STATES \t1").replace("J_STATES \t1 \t2", "J_STATES \t1")
            .replace("COEFFS_STATES\t1.\t.5+.5j", "COEFFS_STATES\t" + text))
        rc, out = run(REAL, ["params.real"], real_dir)
        assert "COEFFS_STATES\t\t" + expect in out, out
