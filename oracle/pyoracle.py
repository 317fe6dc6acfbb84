"""ctypes bindings for the CPU oracle (oracle/oracle.c).

Test infrastructure: imported only by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker.  The product library
never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint32)
_lib = None


def build():
    """Compile oracle/liboracle.so with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        i, d, f = ctypes.c_int, ctypes.c_double, ctypes.c_float
        L.orc_debye_f64.argtypes = [_dp, i, _dp, i, _dp, _dp, i]
        L.orc_debye_rows.argtypes = [_dp, i, _dp, i, _dp, _ip, i, _dp, i]
        L.orc_debye_ref_rows.argtypes = [_fp, i, _fp, i, _fp, i, i, i, _dp, i]
        L.orc_body_factors.argtypes = [_fp, i, _fp, i, i, i, f, _fp, _dp]
        L.orc_fibonacci_dots.argtypes = [i, f, f, _fp]
        L.orc_dodeca_icosa_dots.argtypes = [f, f, _fp]
        L.orc_sasa_counts.argtypes = [_fp, i, _fp, _fp, i, _up, i]
        L.orc_sasa_counts_cells.argtypes = [_fp, i, _fp, i, _fp, i, _up, i]
        L.orc_atomic_radii.argtypes = [_fp]
        L.orc_atomic_ff.argtypes = [_fp, i, f, f, _fp]
        L.orc_scale_chi2.argtypes = [_dp, _dp, _dp, i, _dp, _dp]
        L.orc_chi2_grad_f64.argtypes = [_dp, i, _dp, i, _dp, _dp, _dp, _dp, _dp, _dp, _dp, i]
        L.orc_chi2_fd.argtypes = [_dp, i, _dp, i, _dp, _dp, _dp, _ip, i, d, _dp, i]
        _lib = L
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def bodies(pos, types):
    """float32 [N,4] array: x, y, z and the type id in the bit pattern of w."""
    pos = np.asarray(pos, dtype=np.float32).reshape(-1, 3)
    b = np.zeros((len(pos), 4), dtype=np.float32)
    b[:, :3] = pos
    b[:, 3] = np.asarray(types, dtype=np.uint32).view(np.float32)
    return b


def body_factors(pos, types, q, table, water_index=-1, w=0.0, s=None):
    """FP64 per-atom factors [N, n_q] = table[q][type] + s * w * table[q][water]."""
    b = bodies(pos, types)
    table = _c(table, np.float32)
    nq, nt = table.shape
    F = np.zeros((len(b), nq), dtype=np.float64)
    sp = None if s is None else _ptr(_c(s, np.float32), ctypes.c_float)
    lib().orc_body_factors(_ptr(b, ctypes.c_float), len(b), _ptr(table, ctypes.c_float), nq, nt,
                           int(water_index), float(w), sp, _ptr(F, ctypes.c_double))
    return F


def debye(q, pos, F, nthreads=0):
    """FP64 I(q) of FP32 positions (widened) and FP64 factors F[N, n_q]."""
    q = _c(np.asarray(q, dtype=np.float32), np.float64)
    xyz = _c(np.asarray(pos, dtype=np.float32).reshape(-1, 3), np.float64)
    F = _c(F, np.float64)
    out = np.zeros(len(q), dtype=np.float64)
    lib().orc_debye_f64(_ptr(q, ctypes.c_double), len(q), _ptr(xyz, ctypes.c_double), len(xyz),
                        _ptr(F, ctypes.c_double), _ptr(out, ctypes.c_double), int(nthreads))
    return out


def debye_xyz64(q, xyz, F, nthreads=0):
    """As debye() but with FP64 positions used as given."""
    q = _c(np.asarray(q, dtype=np.float32), np.float64)
    xyz = _c(xyz, np.float64).reshape(-1, 3)
    F = _c(F, np.float64)
    out = np.zeros(len(q), dtype=np.float64)
    lib().orc_debye_f64(_ptr(q, ctypes.c_double), len(q), _ptr(xyz, ctypes.c_double), len(xyz),
                        _ptr(F, ctypes.c_double), _ptr(out, ctypes.c_double), int(nthreads))
    return out


def debye_rows(q, pos, F, rows, nthreads=0):
    """sum over i in rows, all j, of F_i F_j sinc(q r_ij) (FP64)."""
    q = _c(np.asarray(q, dtype=np.float32), np.float64)
    xyz = _c(np.asarray(pos, dtype=np.float32).reshape(-1, 3), np.float64)
    F = _c(F, np.float64)
    rows = _c(rows, np.int32)
    out = np.zeros(len(q), dtype=np.float64)
    lib().orc_debye_rows(_ptr(q, ctypes.c_double), len(q), _ptr(xyz, ctypes.c_double), len(xyz),
                         _ptr(F, ctypes.c_double), _ptr(rows, ctypes.c_int), len(rows),
                         _ptr(out, ctypes.c_double), int(nthreads))
    return out


def debye_ref_rows(q, pos, types, table, y0=0, y1=None, nthreads=1):
    """The reference CPU loop (host_debye.cpp:54-109) over rows [y0, y1)."""
    b = bodies(pos, types)
    q = _c(q, np.float32)
    table = _c(table, np.float32)
    if y1 is None:
        y1 = len(b)
    out = np.zeros(len(q), dtype=np.float64)
    lib().orc_debye_ref_rows(_ptr(q, ctypes.c_float), len(q), _ptr(b, ctypes.c_float), len(b),
                             _ptr(table, ctypes.c_float), table.shape[1], int(y0), int(y1),
                             _ptr(out, ctypes.c_double), int(nthreads))
    return out


def fibonacci_dots(nhalf, radius, probe):
    out = np.zeros((2 * nhalf, 4), dtype=np.float32)
    lib().orc_fibonacci_dots(int(nhalf), float(radius), float(probe), _ptr(out, ctypes.c_float))
    return out


def dodeca_icosa_dots(radius, probe):
    out = np.zeros((32, 4), dtype=np.float32)
    lib().orc_dodeca_icosa_dots(float(radius), float(probe), _ptr(out, ctypes.c_float))
    return out


def sasa_counts(pos, types, R2, dots, cells=True, nthreads=0):
    """Exposed dot counts (uint32) per atom; dots [n_types, D, 4]; R2[type] = (p + r)^2."""
    b = bodies(pos, types)
    R2 = _c(R2, np.float32)
    dots = _c(dots, np.float32)
    D = dots.shape[1]
    out = np.zeros(len(b), dtype=np.uint32)
    if cells:
        lib().orc_sasa_counts_cells(_ptr(b, ctypes.c_float), len(b), _ptr(R2, ctypes.c_float), len(R2),
                                    _ptr(dots, ctypes.c_float), D, _ptr(out, ctypes.c_uint32), int(nthreads))
    else:
        lib().orc_sasa_counts(_ptr(b, ctypes.c_float), len(b), _ptr(R2, ctypes.c_float),
                              _ptr(dots, ctypes.c_float), D, _ptr(out, ctypes.c_uint32), int(nthreads))
    return out


def atomic_radii():
    out = np.zeros(16, dtype=np.float32)
    lib().orc_atomic_radii(_ptr(out, ctypes.c_float))
    return out


def atomic_ff(q, avg_radius=1.62, c=1.0):
    q = _c(q, np.float32)
    out = np.zeros((len(q), 16), dtype=np.float32)
    lib().orc_atomic_ff(_ptr(q, ctypes.c_float), len(q), float(avg_radius), float(c), _ptr(out, ctypes.c_float))
    return out


def scale_chi2(Iexp, sigma, I):
    Iexp, sigma, I = (_c(a, np.float64) for a in (Iexp, sigma, I))
    c, x2 = ctypes.c_double(), ctypes.c_double()
    lib().orc_scale_chi2(_ptr(Iexp, ctypes.c_double), _ptr(sigma, ctypes.c_double), _ptr(I, ctypes.c_double),
                         len(I), ctypes.byref(c), ctypes.byref(x2))
    return c.value, x2.value


def chi2_grad(q, xyz, F, Iexp, sigma, nthreads=0):
    """Analytic FP64 gradient [N,3], plus I, scale, chi2."""
    q = _c(np.asarray(q, dtype=np.float32), np.float64)
    xyz = _c(xyz, np.float64).reshape(-1, 3)
    F, Iexp, sigma = (_c(a, np.float64) for a in (F, Iexp, sigma))
    g = np.zeros_like(xyz)
    I = np.zeros(len(q))
    c, x2 = ctypes.c_double(), ctypes.c_double()
    lib().orc_chi2_grad_f64(_ptr(q, ctypes.c_double), len(q), _ptr(xyz, ctypes.c_double), len(xyz),
                            _ptr(F, ctypes.c_double), _ptr(Iexp, ctypes.c_double), _ptr(sigma, ctypes.c_double),
                            _ptr(g, ctypes.c_double), _ptr(I, ctypes.c_double), ctypes.byref(c), ctypes.byref(x2),
                            int(nthreads))
    return g, I, c.value, x2.value


def chi2_fd(q, xyz, F, Iexp, sigma, comps, h=1e-3, nthreads=0):
    """Central finite differences of chi2 (scale re-optimised) for flat coordinate ids."""
    q = _c(np.asarray(q, dtype=np.float32), np.float64)
    xyz = _c(xyz, np.float64).reshape(-1, 3)
    F, Iexp, sigma = (_c(a, np.float64) for a in (F, Iexp, sigma))
    comps = _c(comps, np.int32)
    fd = np.zeros(len(comps))
    lib().orc_chi2_fd(_ptr(q, ctypes.c_double), len(q), _ptr(xyz, ctypes.c_double), len(xyz),
                      _ptr(F, ctypes.c_double), _ptr(Iexp, ctypes.c_double), _ptr(sigma, ctypes.c_double),
                      _ptr(comps, ctypes.c_int), len(comps), float(h), _ptr(fd, ctypes.c_double), int(nthreads))
    return fd


# ---------------------------------------------------------------------------
# The reference's own CPU code, compiled from its sources by `make -C oracle ref`
# into oracle/_ref/libresaxs_ref.so (see oracle/ref_bridge.cpp).  Optional:
# every caller checks ref_lib() for None (the reference tree is only present
# where the library was built).
# ---------------------------------------------------------------------------
REF_TREE = "/root/reference"
REF_LIB_PATH = os.path.join(HERE, "_ref", "libresaxs_ref.so")
_ref = None


def build_ref():
    """Compile oracle/_ref/libresaxs_ref.so when the reference tree is present."""
    if os.path.isdir(os.path.join(REF_TREE, "resaxslib")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={REF_TREE}"], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def ref_lib():
    """The compiled reference subset, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        L = ctypes.CDLL(REF_LIB_PATH)
        i, f = ctypes.c_int, ctypes.c_float
        L.ref_host_debye.argtypes = [_fp, i, _fp, i, _fp, i, _fp]
        L.ref_atomic_ff.argtypes = [_fp, i, f, f, _fp]
        L.ref_atomic_radii.argtypes = [_fp]
        L.ref_avg_radius.argtypes = [_fp, i]
        L.ref_avg_radius.restype = f
        for name in ("ref_fibonacci_dots", "ref_dodeca_icosa_dots"):
            getattr(L, name).argtypes = [f, f, _fp, i]
        L.ref_density_dots.argtypes = [f, f, f, _fp, i]
        L.ref_rsasa_cpu.argtypes = [_fp, i, f, f, _dp]
        L.ref_pdb_read.argtypes = [ctypes.c_char_p, i, _fp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong),
                                   ctypes.POINTER(ctypes.c_int), i, ctypes.POINTER(ctypes.c_int)]
        L.ref_pdb_atom_type.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
        _ref = L
    return _ref


def ref_host_debye(q, pos, types, table):
    """host_debye<double, float, cl_float4>::calc_curve: FP32 I(q)."""
    b = bodies(pos, types)
    q = _c(q, np.float32)
    table = _c(table, np.float32)
    out = np.zeros(len(q), dtype=np.float32)
    ref_lib().ref_host_debye(_ptr(q, ctypes.c_float), len(q), _ptr(b, ctypes.c_float), len(b),
                             _ptr(table, ctypes.c_float), table.shape[1], _ptr(out, ctypes.c_float))
    return out


def ref_atomic_ff(q, avg_radius=1.62, c=1.0):
    q = _c(q, np.float32)
    out = np.zeros((len(q), 16), dtype=np.float32)
    ref_lib().ref_atomic_ff(_ptr(q, ctypes.c_float), len(q), float(avg_radius), float(c), _ptr(out, ctypes.c_float))
    return out


def ref_atomic_radii():
    out = np.zeros(16, dtype=np.float32)
    ref_lib().ref_atomic_radii(_ptr(out, ctypes.c_float))
    return out


def ref_avg_radius(b):
    b = _c(b, np.float32).reshape(-1, 4)
    return np.float32(ref_lib().ref_avg_radius(_ptr(b, ctypes.c_float), len(b)))


def _ref_dots(fn, *args):
    cap = 4096
    out = np.zeros((cap, 4), dtype=np.float32)
    n = fn(*[float(a) for a in args], _ptr(out, ctypes.c_float), cap)
    assert 0 < n <= cap
    return out[:n].copy()


def ref_fibonacci_dots(radius, probe):
    return _ref_dots(ref_lib().ref_fibonacci_dots, radius, probe)


def ref_dodeca_icosa_dots(radius, probe):
    return _ref_dots(ref_lib().ref_dodeca_icosa_dots, radius, probe)


def ref_density_dots(density, radius, probe):
    return _ref_dots(ref_lib().ref_density_dots, density, radius, probe)


def ref_rsasa_cpu(b, density, probe):
    """rsasa_cpu_alg::calc_rsasa exposed fractions (FP64) for bodies b [N, 4]."""
    b = _c(b, np.float32).reshape(-1, 4)
    out = np.zeros(len(b), dtype=np.float64)
    rc = ref_lib().ref_rsasa_cpu(_ptr(b, ctypes.c_float), len(b), float(density), float(probe),
                                 _ptr(out, ctypes.c_double))
    assert rc == 0
    return out


def ref_pdb_read(path, all_models=True):
    """parse_pdb<float> of the reference: (bodies [N, 4], offsets [M + 1]), or the error code."""
    L = ref_lib()
    n, m = ctypes.c_longlong(), ctypes.c_int()
    cap, ocap = 1 << 20, 4096
    out = np.zeros((cap, 4), dtype=np.float32)
    off = np.zeros(ocap, dtype=np.int32)
    rc = L.ref_pdb_read(str(path).encode(), int(all_models), _ptr(out, ctypes.c_float), cap, ctypes.byref(n),
                        off.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ocap, ctypes.byref(m))
    if rc != 0:
        return rc
    assert n.value <= cap and m.value < ocap
    return out[:n.value].copy(), off[:m.value + 1].copy()


def ref_pdb_atom_type(label, res, element=None):
    """map_pdb_atom_to_ff_type of the reference (element symbol, or None for the label overload)."""
    return ref_lib().ref_pdb_atom_type(label.encode(), res.encode(), None if element is None else element.encode())
