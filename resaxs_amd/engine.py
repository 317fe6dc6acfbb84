"""Python handle over the C ABI: one Engine = one device, one HIP stream.

The method names follow the reference's i_saxs / i_rsasa interface
(resaxslib/include/saxs_algorithm.hpp:147-183, rsasa_algorithm.hpp:42-61):
calc_curve, recalc_curve, commit, revert, calc_rsasa.  Arrays are numpy.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import RsxError


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def make_bodies(pos, types):
    """float32 [N, 4] bodies: x, y, z and the type index in the bit pattern of w
    (the reference's cl_float4 layout, resaxslib/include/utils.hpp:61-88)."""
    pos = np.asarray(pos, dtype=np.float32).reshape(-1, 3)
    b = np.zeros((len(pos), 4), dtype=np.float32)
    b[:, :3] = pos
    b[:, 3] = np.asarray(types, dtype=np.uint32).view(np.float32)
    return b


class Engine:
    def __init__(self, device=0):
        self._L = _lib.lib()
        h = ctypes.c_void_p()
        self._check(self._L.rsx_create(int(device), ctypes.byref(h)), None)
        self.h = h
        self.nq = 0
        self.n = 0

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc, h="self"):
        if rc != 0:
            msg = ""
            if h is not None:
                m = self._L.rsx_last_error(self.h)
                msg = m.decode() if m else ""
            raise RsxError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            self._L.rsx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def state(self):
        return self._L.rsx_state(self.h)

    # -- parameters (saxs_params) ------------------------------------------
    def set_q(self, q):
        q = _f32(q)
        self._check(self._L.rsx_set_q(self.h, _ptr(q, ctypes.c_float), len(q)))
        self.nq = len(q)

    def set_factors(self, table, water_index=-1):
        table = _f32(table)
        if table.ndim != 2 or table.shape[0] != self.nq:
            raise ValueError("factor table must be [n_q, n_types]")
        self._check(self._L.rsx_set_factors(self.h, _ptr(table, ctypes.c_float), table.shape[1], int(water_index)))

    def set_atomic_factors(self, avg_radius):
        """The default atomic factor table [n_q][16] (water row OH2 = 14) at this mean radius."""
        self._check(self._L.rsx_set_atomic_factors(self.h, float(avg_radius)))

    def set_expansion_factor(self, c):
        self._check(self._L.rsx_set_expansion_factor(self.h, float(c)))

    def set_water_weight(self, w):
        self._check(self._L.rsx_set_water_weight(self.h, float(w)))

    def set_no_rsasa_recalc(self, flag):
        self._check(self._L.rsx_set_no_rsasa_recalc(self.h, int(bool(flag))))

    def set_sasa_params(self, radii, probe, kind=_lib.RSX_DOTS_FIBONACCI, n_half=128):
        radii = _f32(radii)
        self._check(self._L.rsx_set_sasa_params(self.h, _ptr(radii, ctypes.c_float), len(radii), float(probe),
                                                int(kind), int(n_half)))

    def set_rebaseline(self, every):
        self._check(self._L.rsx_set_rebaseline(self.h, int(every)))

    def set_tuning(self, k=0, col_range=0):
        self._check(self._L.rsx_set_tuning(self.h, int(k), int(col_range)))

    def set_variant(self, variant):
        """0: FP64 recurrence, one wave per job x q chunk; 1: FP64, shared pair geometry;
        2: FP32 difference-form recurrence with per-lane sums over class runs; 3: the FP32
        recurrence with each column class run reduced over the rows as it ends (K up to
        128; default)."""
        self._check(self._L.rsx_set_variant(self.h, int(variant)))

    def get_tuning(self):
        k, L, rec, sg = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.rsx_get_tuning(self.h, ctypes.byref(k), ctypes.byref(L), ctypes.byref(rec),
                                           ctypes.byref(sg)))
        return {"k": k.value, "col_range": L.value, "recurrence": bool(rec.value), "variant": sg.value}

    def batch_autotune(self, cache_path=None):
        """Tune the batch launch (cached per key in cache_path); returns batch_tuning()."""
        self._check(self._L.rsx_batch_autotune(self.h, cache_path.encode() if cache_path else None))
        return self.batch_tuning()

    def batch_tuning(self):
        v, k, L, pw = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.rsx_batch_get_tuning(self.h, ctypes.byref(v), ctypes.byref(k), ctypes.byref(L),
                                                 ctypes.byref(pw)))
        return {"variant": v.value, "k": k.value, "col_range": L.value, "persistent_waves": pw.value}

    def batch_launch(self):
        """Launch shape of the current batch graph: resident waves of the work-queue kernel
        (0: one wave per (job, q chunk) item) and the pair-job count."""
        pw, nj = ctypes.c_int(), ctypes.c_int()
        self._check(self._L.rsx_batch_launch(self.h, ctypes.byref(pw), ctypes.byref(nj)))
        return {"persistent_waves": pw.value, "n_jobs": nj.value}

    def autotune(self, cache_path=None):
        p = cache_path.encode() if cache_path else None
        self._check(self._L.rsx_autotune(self.h, p))
        return self.get_tuning()

    def initialize(self, n_bodies):
        self._check(self._L.rsx_initialize(self.h, int(n_bodies)))
        self.n = int(n_bodies)

    # -- Debye I(q) (i_saxs) ------------------------------------------------
    def calc_curve(self, bodies):
        b = _f32(bodies).reshape(-1, 4)
        out = np.zeros(self.nq, dtype=np.float64)
        self._check(self._L.rsx_calc_curve_f64(self.h, _ptr(b, ctypes.c_float), len(b), _ptr(out, ctypes.c_double)))
        return out

    def calc_curve_f32(self, bodies):
        b = _f32(bodies).reshape(-1, 4)
        out = np.zeros(self.nq, dtype=np.float32)
        self._check(self._L.rsx_calc_curve(self.h, _ptr(b, ctypes.c_float), len(b), _ptr(out, ctypes.c_float)))
        return out

    def recalc_curve(self, bodies, start, length):
        b = _f32(bodies).reshape(-1, 4)
        out = np.zeros(self.nq, dtype=np.float64)
        self._check(self._L.rsx_recalc_curve_f64(self.h, _ptr(b, ctypes.c_float), len(b), int(start), int(length),
                                                 _ptr(out, ctypes.c_double)))
        return out

    def commit(self):
        self._check(self._L.rsx_commit(self.h))

    def revert(self):
        self._check(self._L.rsx_revert(self.h))

    def get_rsasa(self):
        out = np.zeros(self.n, dtype=np.float32)
        self._check(self._L.rsx_get_rsasa(self.h, _ptr(out, ctypes.c_float), self.n))
        return out

    # -- SASA (i_rsasa) ------------------------------------------------------
    def calc_rsasa(self, bodies, radii, probe, dots):
        """Exposed-dot counts and fractions; dots [n_types, n_dots, 4]."""
        b = _f32(bodies).reshape(-1, 4)
        radii = _f32(radii)
        dots = _f32(dots)
        nd = dots.shape[1]
        counts = np.zeros(len(b), dtype=np.uint32)
        frac = np.zeros(len(b), dtype=np.float32)
        self._check(self._L.rsx_calc_rsasa(self.h, _ptr(b, ctypes.c_float), len(b), _ptr(radii, ctypes.c_float),
                                           len(radii), float(probe), _ptr(dots, ctypes.c_float), nd,
                                           _ptr(counts, ctypes.c_uint), _ptr(frac, ctypes.c_float)))
        return counts, frac

    def calc_curve_hist(self, bodies, bin_width=0.02):
        """Approximate I(q) (FP64) from the pair-distance histogram at bin_width (A)."""
        b = _f32(bodies).reshape(-1, 4)
        out = np.zeros(self.nq, dtype=np.float64)
        self._check(self._L.rsx_calc_curve_hist(self.h, _ptr(b, ctypes.c_float), len(b), float(bin_width),
                                                _ptr(out, ctypes.c_double)))
        return out

    # -- batched ensemble ----------------------------------------------------
    def batch_set_frames(self, bodies, offsets):
        b = _f32(bodies).reshape(-1, 4)
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        self._check(self._L.rsx_batch_set_frames(self.h, _ptr(b, ctypes.c_float), _ptr(off, ctypes.c_int),
                                                 len(off) - 1))
        self.n_frames = len(off) - 1
        self.n_batch_atoms = int(off[-1])
        self.batch_gen = getattr(self, "batch_gen", 0) + 1  # which frames the batch holds

    def set_persistent(self, on):
        """Batched launch: persistent work-queue grid (True, default) or one wave per item."""
        self._check(self._L.rsx_set_persistent(self.h, int(bool(on))))

    def batch_calc(self):
        frames = np.zeros((self.n_frames, self.nq), dtype=np.float64)
        mean = np.zeros(self.nq, dtype=np.float64)
        self._check(self._L.rsx_batch_calc(self.h, _ptr(frames, ctypes.c_double), _ptr(mean, ctypes.c_double)))
        return frames, mean

    def batch_vjp(self, w):
        """Gradients of every frame: grad [atoms of all frames][3] = sum_k w[f][k] dI_f(q_k)/dx_i
        for weights w [n_frames][n_q] (rsx_batch_vjp)."""
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(self.n_frames, self.nq)
        grad = np.zeros((self.n_batch_atoms, 3), dtype=np.float32)
        self._check(self._L.rsx_batch_vjp(self.h, _ptr(w, ctypes.c_double), self.n_frames, self.nq,
                                          _ptr(grad, ctypes.c_float)))
        return grad

    def time_batch_vjp(self, reps=3):
        ms = ctypes.c_float()
        self._check(self._L.rsx_time_batch_vjp(self.h, int(reps), ctypes.byref(ms)))
        return ms.value

    # -- differentiable fit ---------------------------------------------------
    def chi2_grad(self, bodies, i_exp, sigma):
        b = _f32(bodies).reshape(-1, 4)
        ie, sg = _f32(i_exp), _f32(sigma)
        chi2, scale = ctypes.c_double(), ctypes.c_double()
        grad = np.zeros((len(b), 3), dtype=np.float32)
        iq = np.zeros(self.nq, dtype=np.float64)
        self._check(self._L.rsx_chi2_grad(self.h, _ptr(b, ctypes.c_float), len(b), _ptr(ie, ctypes.c_float),
                                          _ptr(sg, ctypes.c_float), ctypes.byref(chi2), ctypes.byref(scale),
                                          _ptr(grad, ctypes.c_float), _ptr(iq, ctypes.c_double)))
        return chi2.value, scale.value, grad, iq

    def vjp(self, bodies, w):
        b = _f32(bodies).reshape(-1, 4)
        w = np.ascontiguousarray(w, dtype=np.float64)
        grad = np.zeros((len(b), 3), dtype=np.float32)
        self._check(self._L.rsx_debye_vjp(self.h, _ptr(b, ctypes.c_float), len(b), _ptr(w, ctypes.c_double),
                                          _ptr(grad, ctypes.c_float)))
        return grad

    def vjp_dev_async(self, bodies_dev, n, w_dev, grad_dev):
        """rsx_debye_vjp_dev_async on device pointers (ints or c_void_p)."""
        self._check(self._L.rsx_debye_vjp_dev_async(self.h, bodies_dev, int(n), w_dev, grad_dev))

    def chi2_grad_dev_async(self, bodies_dev, n, i_exp_dev, sigma_dev, out_dev, grad_dev=None, iq_dev=None):
        """rsx_chi2_grad_dev_async on device pointers: out_dev receives {chi2, scale} (FP64)."""
        self._check(self._L.rsx_chi2_grad_dev_async(self.h, bodies_dev, int(n), i_exp_dev, sigma_dev, out_dev,
                                                     grad_dev, iq_dev))

    # -- measurement -----------------------------------------------------------
    def time_forward(self, reps):
        tot, pair = ctypes.c_float(), ctypes.c_float()
        self._check(self._L.rsx_time_forward(self.h, int(reps), ctypes.byref(tot), ctypes.byref(pair)))
        return tot.value, pair.value

    def set_adjoint_tuning(self, fp32=True, k=0, col_range=0, per_lane=False):
        """fp32: FP32 difference-form adjoint kernel (default) or FP64 recurrence; per_lane:
        the FP32 kernel's per-lane coefficient form even where the class-pair form applies
        (plain layouts); k q per thread (0 = automatic); col_range columns per job (0 = automatic)."""
        mode = (2 if per_lane else 1) if fp32 else 0
        self._check(self._L.rsx_set_adjoint_tuning(self.h, mode, int(k), int(col_range)))

    def adjoint_layout(self, batch=False):
        """The last gradient call's adjoint layout: form ("fp64", "pair" or "per_lane"), k, col_range."""
        f, k, L = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.rsx_get_adjoint_layout(self.h, int(bool(batch)), ctypes.byref(f), ctypes.byref(k),
                                                   ctypes.byref(L)))
        return {"form": ("fp64", "pair", "per_lane")[f.value], "k": k.value, "col_range": L.value}

    def time_adjoint(self, reps=5):
        ms = ctypes.c_float()
        self._check(self._L.rsx_time_adjoint(self.h, int(reps), ctypes.byref(ms)))
        return ms.value

    def time_sasa(self, reps=5):
        """Average device ms of one SASA count of the engine's atoms (after a hydrated calc_curve)."""
        ms = ctypes.c_float()
        self._check(self._L.rsx_time_sasa(self.h, int(reps), ctypes.byref(ms)))
        return ms.value

    def time_copy(self, nbytes, reps=3):
        ms = ctypes.c_float()
        self._check(self._L.rsx_time_copy(self.h, int(nbytes), int(reps), ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        self._check(self._L.rsx_synchronize(self.h))

    def dev_alloc(self, nbytes):
        p = ctypes.c_void_p()
        self._check(self._L.rsx_dev_alloc(self.h, ctypes.byref(p), int(nbytes)))
        return p

    def dev_free(self, p):
        self._check(self._L.rsx_dev_free(self.h, p))

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        self._check(self._L.rsx_memcpy_h2d(self.h, dst, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes))

    def d2h(self, arr, src):
        self._check(self._L.rsx_memcpy_d2h(self.h, arr.ctypes.data_as(ctypes.c_void_p), src, arr.nbytes))
        return arr

    def calc_curve_dev_async(self, bodies_dev, n, out_dev=None):
        self._check(self._L.rsx_calc_curve_dev_async(self.h, bodies_dev, int(n), out_dev))

    def batch_calc_dev_async(self, frames_dev=None, mean_dev=None):
        self._check(self._L.rsx_batch_calc_dev_async(self.h, frames_dev, mean_dev))


def device_count():
    c = ctypes.c_int()
    rc = _lib.lib().rsx_device_count(ctypes.byref(c))
    if rc != 0:
        raise RsxError(rc, "device count")
    return c.value


def fibonacci_dots(n_half, radius, probe):
    out = np.zeros((2 * n_half, 4), dtype=np.float32)
    rc = _lib.lib().rsx_fibonacci_dots(int(n_half), float(radius), float(probe), _ptr(out, ctypes.c_float))
    if rc:
        raise RsxError(rc, "fibonacci_dots")
    return out


def dodeca_icosa_dots(radius, probe):
    out = np.zeros((32, 4), dtype=np.float32)
    rc = _lib.lib().rsx_dodeca_icosa_dots(float(radius), float(probe), _ptr(out, ctypes.c_float))
    if rc:
        raise RsxError(rc, "dodeca_icosa_dots")
    return out


def sphere_dots(radius, n_sect, probe):
    """Rings-of-dots sphere (the reference's create_sphere_dots); [n, 4] float32."""
    n = ctypes.c_int()
    L = _lib.lib()
    rc = L.rsx_sphere_dots(float(radius), float(n_sect), float(probe), None, 0, ctypes.byref(n))
    if rc:
        raise RsxError(rc, "sphere_dots")
    out = np.zeros((max(n.value, 1), 4), dtype=np.float32)
    rc = L.rsx_sphere_dots(float(radius), float(n_sect), float(probe), _ptr(out, ctypes.c_float), n.value,
                           ctypes.byref(n))
    if rc:
        raise RsxError(rc, "sphere_dots")
    return out[:n.value]


def density_dots(density, radius, probe):
    """The reference's create_dots: n_sect = 2 pi radius sqrt(density) (float arithmetic)."""
    s = np.sqrt(np.float32(density), dtype=np.float32)
    n_sect = np.float32(np.float32(np.float32(2.0) * np.float32(np.pi)) * np.float32(radius)) * s
    return sphere_dots(radius, n_sect, probe)


def device_name(device=0):
    buf = ctypes.create_string_buffer(256)
    rc = _lib.lib().rsx_device_name(int(device), buf, 256)
    if rc:
        raise RsxError(rc, "device_name")
    return buf.value.decode()


def atomic_form_factors(q, avg_radius=1.62, expansion=1.0):
    q = _f32(q)
    out = np.zeros((len(q), 16), dtype=np.float32)
    rc = _lib.lib().rsx_atomic_form_factors(_ptr(q, ctypes.c_float), len(q), float(avg_radius), float(expansion),
                                            _ptr(out, ctypes.c_float))
    if rc:
        raise RsxError(rc, "atomic_form_factors")
    return out


def atomic_radii():
    out = np.zeros(16, dtype=np.float32)
    rc = _lib.lib().rsx_atomic_radii(_ptr(out, ctypes.c_float))
    if rc:
        raise RsxError(rc, "atomic_radii")
    return out


def pdb_atom_type(atom_label, res_name, element=None):
    """United-atom form-factor type (0..15) of a PDB atom; element: a type index, a
    symbol, or None (the label's first letter)."""
    if isinstance(element, int):
        return _lib.lib().rsx_pdb_atom_type(atom_label.encode(), res_name.encode(), None, element)
    return _lib.lib().rsx_pdb_atom_type(atom_label.encode(), res_name.encode(),
                                        None if element is None else element.encode(), -1)


def read_pdb(path, all_models=True, verbose=0):
    """Bodies [n, 4] (float32, type bits in column 3) and model offsets [n_models + 1]
    of the ATOM / HETATM records of a PDB file."""
    L = _lib.lib()
    n, nm = ctypes.c_longlong(), ctypes.c_int()
    p = os.fsencode(path)
    rc = L.rsx_pdb_read(p, int(bool(all_models)), int(verbose), None, 0, ctypes.byref(n), None, 0, ctypes.byref(nm))
    if rc:
        raise RsxError(rc, f"read_pdb {path}")
    bodies = np.zeros((n.value, 4), dtype=np.float32)
    offsets = np.zeros(nm.value + 1, dtype=np.int32)
    rc = L.rsx_pdb_read(p, int(bool(all_models)), int(verbose), _ptr(bodies, ctypes.c_float), n.value,
                        ctypes.byref(n), _ptr(offsets, ctypes.c_int), len(offsets), ctypes.byref(nm))
    if rc:
        raise RsxError(rc, f"read_pdb {path}")
    return bodies, offsets


def fit_ensemble(models_bodies, offsets, q, i_exp, sigma, scale=None, water=(-2.0, 4.0), expansion=(0.95, 1.05),
                 device=0):
    """The calc_saxs ensemble fit on the GPU: returns (expansion factor, water weight,
    scale, chi, evaluations, scaled best I(q)).  scale None: fitted; a number: fixed.
    A range (lo, hi) with lo == hi fixes that parameter."""
    b = _f32(models_bodies).reshape(-1, 4)
    off = np.ascontiguousarray(offsets, dtype=np.int32)
    qq, ie, sg = _f32(q), _f32(i_exp), _f32(sigma)
    sr = np.array([1.0, 2.0] if scale is None else [scale, scale], dtype=np.float32)
    wr = np.array(water, dtype=np.float32)
    er = np.array(expansion, dtype=np.float32)
    out = np.zeros(5, dtype=np.float64)
    iq = np.zeros(len(qq), dtype=np.float32)
    rc = _lib.lib().rsx_fit_ensemble(_ptr(b, ctypes.c_float), _ptr(off, ctypes.c_int), len(off) - 1,
                                     _ptr(qq, ctypes.c_float), _ptr(ie, ctypes.c_float), _ptr(sg, ctypes.c_float),
                                     len(qq), _ptr(sr, ctypes.c_float), _ptr(wr, ctypes.c_float),
                                     _ptr(er, ctypes.c_float), int(device), _ptr(out, ctypes.c_double),
                                     _ptr(iq, ctypes.c_float))
    if rc:
        raise RsxError(rc, "fit_ensemble")
    return float(out[0]), float(out[1]), float(out[2]), float(out[3]), int(out[4]), iq
