"""TEST INFRASTRUCTURE ONLY -- ctypes front end of the CPU oracle (tdc_oracle.c).

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py may
import this module; the product path never does. Build the library with
``make -C oracle`` (done by ``__graft_entry__.build()``).

Functions
  image_f64(...)      FP64 direct sum of the reference formula (the oracle)
  refmethod_f32(...)  FP32 restatement of the reference method (CPU baseline)
  q_prod, q_conj, q_rot, norm3, dot3, doppler_centroid, doppler_freq,
  doppler_window, range_window, taylor_window, center_wavelength,
  chunk_processed     scalar helpers, exposed so the tests can pin them

Conventions for the plain (drop-in) form: the oracle evaluates the formula
in FP64 on the same float32 inputs, upcast, with c = 299792458 exactly and
the window parameter A = 25/46 exactly. For the precision form it runs on the
FP64 inputs. See tdc_oracle.c for the reference lines each part follows.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

C = 299792458.0
HAMMING_A = 25.0 / 46.0

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


class _Scene(ctypes.Structure):
    _fields_ = [
        ("npri", ctypes.c_int),
        ("ns", ctypes.c_int),
        ("raw", _dp),
        ("sample_times", _dp),
        ("pos", _dp),
        ("pos_per_sample", ctypes.c_int),
        ("vel", _dp),
        ("att", _dp),
        ("mod_rate", ctypes.c_double),
        ("start_freq", ctypes.c_double),
        ("lambda_c", ctypes.c_double),
        ("bore", ctypes.c_double * 3),
        ("range_win", _dp),
        ("doppler", ctypes.c_int),
        ("dop_center", ctypes.c_double),
        ("dop_bw", ctypes.c_double),
        ("centroid", ctypes.c_int),
        ("taper", ctypes.c_int),
        ("az_a", ctypes.c_double),
        ("c", ctypes.c_double),
    ]


def load():
    """Load liborc.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    lib = ctypes.CDLL(os.path.join(_HERE, "liborc.so"))
    lib.orc_norm3.restype = ctypes.c_double
    lib.orc_norm3.argtypes = [ctypes.c_double] * 3
    lib.orc_dot3.restype = ctypes.c_double
    lib.orc_dot3.argtypes = [_dp, _dp]
    for name in ("orc_q_conj",):
        getattr(lib, name).argtypes = [_dp, _dp]
    lib.orc_q_prod.argtypes = [_dp, _dp, _dp]
    lib.orc_q_rot.argtypes = [_dp, _dp, _dp]
    lib.orc_doppler_centroid.restype = ctypes.c_double
    lib.orc_doppler_centroid.argtypes = [_dp, _dp, ctypes.c_double]
    lib.orc_doppler_freq.restype = ctypes.c_double
    lib.orc_doppler_freq.argtypes = [_dp, _dp, _dp, ctypes.c_double]
    lib.orc_doppler_window.restype = ctypes.c_double
    lib.orc_doppler_window.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_double]
    lib.orc_range_window.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
    lib.orc_taylor_window.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                      _dp]
    lib.orc_center_wavelength.restype = ctypes.c_double
    lib.orc_center_wavelength.argtypes = [ctypes.c_double] * 4
    lib.orc_chunk_processed.restype = ctypes.c_int
    lib.orc_chunk_processed.argtypes = [ctypes.POINTER(_Scene), ctypes.c_int, _dp]
    lib.orc_image.argtypes = [ctypes.POINTER(_Scene), _dp, ctypes.c_int, _dp, ctypes.c_int]
    lib.orc_refmethod.argtypes = [_fp, _fp, _fp, _fp, _fp, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_float, ctypes.c_float, _fp, ctypes.c_float, _fp, _fp,
                                  ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, _fp, ctypes.c_int]
    _lib = lib
    return lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def _threads(n):
    return int(n) if n else min(16, os.cpu_count() or 1)


# ---- scalar helpers ---------------------------------------------------------

def norm3(x, y, z):
    return load().orc_norm3(x, y, z)


def dot3(a, b):
    a, pa = _d(a)
    b, pb = _d(b)
    return load().orc_dot3(pa, pb)


def q_conj(q):
    q, pq = _d(q)
    out = np.zeros(4)
    load().orc_q_conj(pq, out.ctypes.data_as(_dp))
    return out


def q_prod(a, b):
    a, pa = _d(a)
    b, pb = _d(b)
    out = np.zeros(4)
    load().orc_q_prod(pa, pb, out.ctypes.data_as(_dp))
    return out


def q_rot(q, v):
    q, pq = _d(q)
    v, pv = _d(v)
    out = np.zeros(3)
    load().orc_q_rot(pq, pv, out.ctypes.data_as(_dp))
    return out


def doppler_centroid(vel, ant, lam):
    v, pv = _d(vel)
    a, pa = _d(ant)
    return load().orc_doppler_centroid(pv, pa, lam)


def doppler_freq(pos, vel, target, lam):
    p, pp = _d(pos)
    v, pv = _d(vel)
    t, pt = _d(target)
    return load().orc_doppler_freq(pp, pv, pt, lam)


def doppler_window(fd, fc, bw, taper=True, a=HAMMING_A):
    return load().orc_doppler_window(fd, fc, bw, 1 if taper else 0, a)


def range_window(ns, a=HAMMING_A, apply=True):
    out = np.zeros(ns)
    load().orc_range_window(ns, a, 1 if apply else 0, out.ctypes.data_as(_dp))
    return out


def taylor_window(ns, nbar=4, sll=30.0, periodic=True):
    out = np.zeros(ns)
    load().orc_taylor_window(ns, nbar, sll, 1 if periodic else 0, out.ctypes.data_as(_dp))
    return out


def center_wavelength(t_first, t_last, mod_rate, start_freq):
    return load().orc_center_wavelength(t_first, t_last, mod_rate, start_freq)


# ---- image oracle ------------------------------------------------------------

def _scene(raw, sample_times, pos, mod_rate, start_freq, range_win, doppler, c):
    raw = np.asarray(raw)
    npri, ns = raw.shape
    keep = []
    rawd = np.ascontiguousarray(np.stack([raw.real, raw.imag], axis=-1), dtype=np.float64)
    keep.append(rawd)
    t, pt = _d(sample_times)
    keep.append(t)
    pos = np.asarray(pos, dtype=np.float64)
    per_sample = 1 if pos.ndim == 3 else 0
    pos, ppos = _d(pos)
    keep.append(pos)
    rw = np.ones(ns) if range_win is None else np.asarray(range_win, dtype=np.float64)
    rw, prw = _d(rw)
    keep.append(rw)
    s = _Scene()
    s.npri, s.ns = npri, ns
    s.raw = rawd.ctypes.data_as(_dp)
    s.sample_times = pt
    s.pos = ppos
    s.pos_per_sample = per_sample
    s.mod_rate, s.start_freq = float(mod_rate), float(start_freq)
    s.range_win = prw
    s.c = float(c)
    s.az_a = HAMMING_A
    if doppler:
        if not per_sample:
            raise ValueError("the Doppler window needs per-sample positions")
        vel, pv = _d(doppler["velocity"])
        att, pa = _d(doppler["attitude"])
        keep += [vel, att]
        s.vel, s.att = pv, pa
        s.doppler = 1
        s.dop_center = float(doppler["center"])
        s.dop_bw = float(doppler["bandwidth"])
        s.centroid = 1 if doppler.get("centroid") else 0
        s.taper = 1 if doppler.get("taper") else 0
        b = np.asarray(doppler["boresight"], dtype=np.float64)
        s.bore[0], s.bore[1], s.bore[2] = b
        s.lambda_c = float(doppler.get("lambda_c") or center_wavelength(
            float(np.asarray(sample_times)[0]), float(np.asarray(sample_times)[-1]), mod_rate,
            start_freq))
    return s, keep


def image_f64(raw, sample_times, pos, grid, mod_rate, start_freq, range_win=None, doppler=None,
              c=C, pixels=None, nthreads=None):
    """FP64 image. grid [..., 3]; pixels: optional flat indices into the grid.

    Returns complex128 of shape grid.shape[:-1] (or [len(pixels)]).
    """
    s, keep = _scene(raw, sample_times, pos, mod_rate, start_freq, range_win, doppler, c)
    g = np.asarray(grid, dtype=np.float64)
    shape = g.shape[:-1]
    gf = g.reshape(-1, 3)
    if pixels is not None:
        gf = gf[np.asarray(pixels, dtype=np.int64)]
        shape = (gf.shape[0],)
    gf, pg = _d(gf)
    out = np.zeros((gf.shape[0], 2))
    load().orc_image(ctypes.byref(s), pg, gf.shape[0], out.ctypes.data_as(_dp), _threads(nthreads))
    del keep
    return (out[:, 0] + 1j * out[:, 1]).reshape(shape)


def chunk_processed(raw, sample_times, pos, mod_rate, start_freq, doppler, chunk, target):
    s, keep = _scene(raw, sample_times, pos, mod_rate, start_freq, None, doppler, C)
    t, pt = _d(target)
    r = load().orc_chunk_processed(ctypes.byref(s), int(chunk), pt)
    del keep
    return bool(r)


def refmethod_f32(d, center, bandwidth, centroid, apply_range_win, taper, pixels=None,
                  nthreads=None, doppler=True, window_a=HAMMING_A):
    """Reference method in FP32 on the reference's float inputs.

    d: dict with rawData (complex64 [nc, ns]), sampleTimes, position, velocity,
    attitude, modRate, startFreq, bodyBoresight, focusGrid (float32).
    doppler=False leaves out the skip test and the Doppler window (valid when
    the window is identically one). Returns complex64 over the grid (or pixels).
    """
    raw = np.ascontiguousarray(d["rawData"], dtype=np.complex64)
    npri, ns = raw.shape
    rawf = raw.view(np.float32)
    t, pt = _f(d["sampleTimes"])
    pos, ppos = _f(d["position"])
    vel, pvel = _f(d["velocity"])
    att, patt = _f(d["attitude"])
    bore, pbore = _f(d["bodyBoresight"])
    g = np.asarray(d["focusGrid"], dtype=np.float32)
    shape = g.shape[:-1]
    gf = g.reshape(-1, 3)
    if pixels is not None:
        gf = gf[np.asarray(pixels, dtype=np.int64)]
        shape = (gf.shape[0],)
    gf, pg = _f(gf)
    mod_rate, start_freq = np.float32(d["modRate"]), np.float32(d["startFreq"])
    # tdcprocessor.cpp:204-208
    chirp_time = np.float64(np.float32(t[ns - 1]) - np.float32(t[0]))
    end_freq = np.float64(start_freq) + np.float64(mod_rate) * chirp_time
    lam = np.float32(C / ((np.float64(start_freq) + end_freq) / 2.0))
    rw = range_window(ns, window_a, apply_range_win).astype(np.float32)
    rw, prw = _f(rw)
    out = np.zeros((gf.shape[0], 2), dtype=np.float32)
    load().orc_refmethod(rawf.ctypes.data_as(_fp), pt, ppos, pvel, patt, npri, ns,
                         float(mod_rate), float(start_freq), pbore, float(lam), prw, pg,
                         gf.shape[0], float(center), float(bandwidth), 1 if centroid else 0,
                         1 if taper else 0, 1 if doppler else 0, out.ctypes.data_as(_fp),
                         _threads(nthreads))
    return (out[:, 0] + 1j * out[:, 1]).astype(np.complex64).reshape(shape)


def rel_l2(img, ref):
    img = np.asarray(img, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    return float(np.linalg.norm(img - ref) / np.linalg.norm(ref))


def peak_phase_error(img, ref):
    img = np.asarray(img, dtype=np.complex128).reshape(-1)
    ref = np.asarray(ref, dtype=np.complex128).reshape(-1)
    p = int(np.argmax(np.abs(ref)))
    return float(abs(np.angle(img[p] * np.conj(ref[p]))))
