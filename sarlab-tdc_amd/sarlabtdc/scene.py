"""Seeded synthetic FMCW SAR scenes for the TDC processor.

Generates point-scatterer scenes, their linear-FMCW beat signals (FP64,
stored complex64) and the focus grids for these workloads:

  line_small   64 x 128 samples, straight 0.3 m track, 32^2 grid at 5 mm,
               0.5 m range, 8 scatterers, 30 dB, Hann in range
  line_jitter  256 x 256, 0.5 m track with 0.2 mm jitter, 128^2 at 3 mm,
               1 m range, 200 scatterers, 25 dB, Hann x Hann
  arc          1024 x 512, 1 m arc at 2 m slant range with 0.5 mm jitter,
               512^2 at 2 mm, 2000 log-normal scatterers, 20 dB, Hann x Hann
  helix        2048 x 512, helical segment (radius 0.5 m), 768^2 at 2 mm on a
               plane tilted 30 degrees, 5000 scatterers (10% bright), 20 dB,
               Taylor (nbar 4, -30 dB) in range x Hann in azimuth
  stream_job   jobs of a serving stream: 64-2048 chirps, 128/256/512
               samples, 32^2-512^2 grids at 2-5 mm, drawn as line_jitter

Common radar: f0 = 77 GHz, B = 4 GHz, chirp 40 us (K = 1e14 Hz/s), samples
at t_m = m * Tc / Ns, c = 299792458 m/s. Positions are stop-and-hop (one
phase centre per chirp). Windows are applied to the raw data here, so the
processor runs with applyRangeWin = False and the Doppler window off
(dopplerBandwidth = 1e30, no centroid, no taper).

The beat signal follows the sign convention the correlator focuses
(reference tdc-lib/src/tdckernels.cu:206-215 and :270-272):
  raw[n, m] = sum_k a_k exp(-j 2 pi * 2 (f0 + K t_m) |track[n] - s_k| / c) + noise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

C = 299792458.0
F0 = 77e9
BANDWIDTH = 4e9
CHIRP_TIME = 40e-6
MOD_RATE = BANDWIDTH / CHIRP_TIME
PRI = 100e-6

_HERE = os.path.dirname(os.path.abspath(__file__))
_scene_lib = None


def _lib():
    global _scene_lib
    if _scene_lib is None:
        lib = ctypes.CDLL(os.path.join(_HERE, "libtdcscene.so"))
        dp = ctypes.POINTER(ctypes.c_double)
        lib.tdcsim_beat.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, dp, dp, ctypes.c_int, dp,
                                    ctypes.c_int]
        lib.tdcsim_beat.restype = ctypes.c_int
        _scene_lib = lib
    return _scene_lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def simulate_beat(track, sample_times, scatterers, amplitudes, f0=F0, mod_rate=MOD_RATE, c=C,
                  nthreads=None):
    """Noiseless beat signal, complex128 [n_chirps, n_samples]."""
    track = np.ascontiguousarray(track, dtype=np.float64)
    t = np.ascontiguousarray(sample_times, dtype=np.float64)
    s = np.ascontiguousarray(np.asarray(scatterers, dtype=np.float64).reshape(-1, 3))
    a = np.asarray(amplitudes, dtype=np.complex128).reshape(-1)
    amp = np.ascontiguousarray(np.stack([a.real, a.imag], axis=1))
    nc, ns = track.shape[0], t.shape[0]
    out = np.empty((nc, ns, 2), dtype=np.float64)
    nt = nthreads or min(16, os.cpu_count() or 1)
    rc = _lib().tdcsim_beat(_ptr(track), nc, ns, _ptr(t), float(f0), float(mod_rate), float(c),
                            _ptr(s), _ptr(amp), s.shape[0], _ptr(out), int(nt))
    if rc != 0:
        raise ValueError("tdcsim_beat rejected its arguments")
    return out[..., 0] + 1j * out[..., 1]


def hann(n):
    """Periodic Hann window (denominator n)."""
    m = np.arange(n)
    return 0.5 - 0.5 * np.cos(2.0 * np.pi * m / n)


def taylor(n, nbar=4, sll=30.0):
    """Periodic Taylor window (an (n+1)-point symmetric window without its last point)."""
    nf = n + 1
    b = 10.0 ** (sll / 20.0)
    a = np.arccosh(b) / np.pi
    s2 = nbar ** 2 / (a * a + (nbar - 0.5) ** 2)
    fm = np.zeros(max(nbar - 1, 0))
    for m in range(1, nbar):
        num = 1.0 if m % 2 == 1 else -1.0
        den = 2.0
        for i in range(1, nbar):
            num *= 1.0 - m * m / s2 / (a * a + (i - 0.5) ** 2)
            if i != m:
                den *= 1.0 - m * m / float(i * i)
        fm[m - 1] = num / den

    def w(x):
        x = np.asarray(x, dtype=np.float64)
        s = np.ones_like(x)
        for m in range(1, nbar):
            s = s + 2.0 * fm[m - 1] * np.cos(2.0 * np.pi * m * (x - nf / 2.0 + 0.5) / nf)
        return s

    return w(np.arange(n)) / w((nf - 1) / 2.0)


@dataclass
class Scene:
    """A synthetic acquisition plus its focus grid (FP64 master copies)."""

    name: str
    raw: np.ndarray            # complex64 [nc, ns]
    track: np.ndarray          # float64 [nc, 3], phase centre per chirp
    sample_times: np.ndarray   # float64 [ns]
    pri_times: np.ndarray      # float64 [nc]
    grid: np.ndarray           # float64 [rows, cols, 3]
    f0: float = F0
    mod_rate: float = MOD_RATE
    scatterers: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    amplitudes: np.ndarray = field(default_factory=lambda: np.zeros(0, np.complex128))
    velocity_speed: float = 0.0

    @property
    def n_chirps(self):
        return self.raw.shape[0]

    @property
    def n_samples(self):
        return self.raw.shape[1]

    @property
    def n_pixels(self):
        return self.grid.shape[0] * self.grid.shape[1]

    @property
    def terms(self):
        return float(self.n_pixels) * self.n_chirps * self.n_samples

    def dropin(self):
        """Arrays in the reference's float32 layout (tdcprocessor.h:57-107).

        position/velocity [nc, ns, 3] (stop-and-hop broadcast), attitude the
        identity quaternion, boresight (0, 1, 0), velocity along the track
        tangent with speed track length / (nc * 100 us).
        """
        nc, ns = self.raw.shape
        pos32 = self.track.astype(np.float32)
        position = np.ascontiguousarray(np.broadcast_to(pos32[:, None, :], (nc, ns, 3)))
        if nc > 1:
            tangent = np.gradient(self.track, axis=0)
        else:
            tangent = np.zeros_like(self.track)
        norms = np.linalg.norm(tangent, axis=1, keepdims=True)
        norms[norms == 0] = 1.0
        vel = (tangent / norms * self.velocity_speed).astype(np.float32)
        velocity = np.ascontiguousarray(np.broadcast_to(vel[:, None, :], (nc, ns, 3)))
        attitude = np.zeros((nc, ns, 4), dtype=np.float32)
        attitude[..., 0] = 1.0
        return dict(
            rawData=np.ascontiguousarray(self.raw, dtype=np.complex64),
            priTimes=self.pri_times.astype(np.float32),
            sampleTimes=self.sample_times.astype(np.float32),
            position=position,
            velocity=velocity,
            attitude=attitude,
            modRate=np.float32(self.mod_rate),
            startFreq=np.float32(self.f0),
            bodyBoresight=np.array([0.0, 1.0, 0.0], dtype=np.float32),
            focusGrid=np.ascontiguousarray(self.grid, dtype=np.float32),
        )

    def f64(self):
        """Arrays for the FP64 (precision-form) input API."""
        return dict(
            rawData=np.ascontiguousarray(self.raw, dtype=np.complex64),
            priTimes=np.ascontiguousarray(self.pri_times, dtype=np.float64),
            sampleTimes=np.ascontiguousarray(self.sample_times, dtype=np.float64),
            position=np.ascontiguousarray(self.track, dtype=np.float64),
            modRate=float(self.mod_rate),
            startFreq=float(self.f0),
            focusGrid=np.ascontiguousarray(self.grid, dtype=np.float64),
        )


def plane_grid(rows, cols, spacing, centre, row_axis=(0.0, 1.0, 0.0), col_axis=(1.0, 0.0, 0.0)):
    """rows x cols grid; rows step along row_axis, columns along col_axis."""
    r = (np.arange(rows) - (rows - 1) / 2.0) * spacing
    c = (np.arange(cols) - (cols - 1) / 2.0) * spacing
    ra, ca = np.asarray(row_axis, float), np.asarray(col_axis, float)
    return (np.asarray(centre, float)[None, None, :] + r[:, None, None] * ra[None, None, :]
            + c[None, :, None] * ca[None, None, :])


def _cn(rng, n, var):
    return (rng.normal(size=n) + 1j * rng.normal(size=n)) * np.sqrt(var / 2.0)


def _uniform_on_grid(rng, grid, n):
    """n points uniform over the grid's parallelogram."""
    rows, cols = grid.shape[:2]
    o = grid[0, 0]
    er = grid[-1, 0] - o
    ec = grid[0, -1] - o
    u, v = rng.uniform(size=n), rng.uniform(size=n)
    return o[None, :] + u[:, None] * er[None, :] + v[:, None] * ec[None, :]


def synthesize(name, track, ns, grid, scatterers, amplitudes, snr_db, rng, range_win=None,
               az_win=None, speed=0.0, f0=F0, mod_rate=MOD_RATE):
    nc = track.shape[0]
    t = np.arange(ns, dtype=np.float64) * (CHIRP_TIME / ns)
    sig = simulate_beat(track, t, scatterers, amplitudes, f0=f0, mod_rate=mod_rate)
    if snr_db is not None and len(amplitudes):
        p = np.mean(np.abs(sig) ** 2)
        sigma2 = p / 10.0 ** (snr_db / 10.0)
        sig = sig + _cn(rng, sig.size, sigma2).reshape(sig.shape)
    if range_win is not None:
        sig = sig * range_win[None, :]
    if az_win is not None:
        sig = sig * az_win[:, None]
    return Scene(name=name, raw=sig.astype(np.complex64), track=track, sample_times=t,
                 pri_times=np.arange(nc, dtype=np.float64) * PRI, grid=grid,
                 f0=f0, mod_rate=mod_rate, scatterers=np.asarray(scatterers),
                 amplitudes=np.asarray(amplitudes), velocity_speed=speed)


def _speed(track):
    length = float(np.sum(np.linalg.norm(np.diff(track, axis=0), axis=1)))
    return length / (track.shape[0] * PRI)


def line_small(seed=0):
    rng = np.random.default_rng([1, seed])
    nc, ns = 64, 128
    track = np.zeros((nc, 3))
    track[:, 0] = np.linspace(-0.15, 0.15, nc)
    grid = plane_grid(32, 32, 5e-3, (0.0, 0.5, 0.0))
    scat = _uniform_on_grid(rng, grid, 8)
    amp = _cn(rng, 8, 1.0)
    return synthesize("line_small", track, ns, grid, scat, amp, 30.0, rng, range_win=hann(ns),
                      speed=_speed(track))


def line_jitter(seed=0, nc=256, ns=256, side=128, spacing=3e-3, n_scat=200):
    rng = np.random.default_rng([2, seed])
    track = np.zeros((nc, 3))
    track[:, 0] = np.linspace(-0.25, 0.25, nc)
    speed = _speed(track)
    track = track + rng.normal(scale=0.2e-3, size=track.shape)
    grid = plane_grid(side, side, spacing, (0.0, 1.0, 0.0))
    scat = _uniform_on_grid(rng, grid, n_scat)
    amp = _cn(rng, n_scat, 1.0)
    return synthesize("line_jitter", track, ns, grid, scat, amp, 25.0, rng, range_win=hann(ns),
                      az_win=hann(nc), speed=speed)


def arc(seed=0, nc=1024, ns=512, side=512):
    rng = np.random.default_rng([3, seed])
    rho, height, length = 1.936, 0.5, 1.0
    phi = -np.pi / 2 + np.linspace(-length / (2 * rho), length / (2 * rho), nc)
    track = np.stack([rho * np.cos(phi), rho * np.sin(phi), np.full(nc, height)], axis=1)
    speed = _speed(track)
    track = track + rng.normal(scale=0.5e-3, size=track.shape)
    grid = plane_grid(side, side, 2e-3, (0.0, 0.0, 0.0))
    n = 2000
    scat = _uniform_on_grid(rng, grid, n)
    amp = np.exp(rng.normal(size=n)) * np.exp(1j * rng.uniform(0, 2 * np.pi, size=n))
    return synthesize("arc", track, ns, grid, scat, amp, 20.0, rng, range_win=hann(ns),
                      az_win=hann(nc), speed=speed)


def helix(seed=0, nc=2048, ns=512, side=768):
    rng = np.random.default_rng([4, seed])
    phi = np.deg2rad(np.linspace(30.0, 150.0, nc))
    radius, pitch = 0.5, 0.2
    track = np.stack([radius * np.cos(phi), radius * np.sin(phi), pitch * phi / (2 * np.pi)],
                     axis=1)
    speed = _speed(track)
    track = track + rng.normal(scale=0.3e-3, size=track.shape)
    tilt = np.deg2rad(30.0)
    grid = plane_grid(side, side, 2e-3, (0.0, 2.0, -0.5),
                      row_axis=(0.0, np.cos(tilt), np.sin(tilt)), col_axis=(1.0, 0.0, 0.0))
    n = 5000
    n_bright = n // 10
    scat = _uniform_on_grid(rng, grid, n)
    amp = np.concatenate([_cn(rng, n - n_bright, 0.1), _cn(rng, n_bright, 10.0)])
    return synthesize("helix", track, ns, grid, scat, amp, 20.0, rng,
                      range_win=taylor(ns, 4, 30.0), az_win=hann(nc), speed=speed)


def stream_shapes(n_jobs=1000, seed=0):
    """(chirps, samples, side, spacing) of each job of the serving stream."""
    rng = np.random.default_rng([5, seed, 0])
    nc = np.rint(np.exp(rng.uniform(np.log(64), np.log(2048), n_jobs))).astype(int)
    ns = rng.choice([128, 256, 512], size=n_jobs)
    side = np.rint(np.exp(rng.uniform(np.log(32), np.log(512), n_jobs))).astype(int)
    spacing = rng.uniform(2e-3, 5e-3, size=n_jobs)
    return [(int(a), int(b), int(c), float(d)) for a, b, c, d in zip(nc, ns, side, spacing)]


def stream_job(j, shape):
    """Scene of stream job j: a jittered 0.5 m line track as line_jitter, with the job's own
    shape and seed."""
    nc, ns, side, spacing = shape
    rng = np.random.default_rng([5, j])
    track = np.zeros((nc, 3))
    track[:, 0] = np.linspace(-0.25, 0.25, nc)
    speed = _speed(track)
    track = track + rng.normal(scale=0.2e-3, size=track.shape)
    grid = plane_grid(side, side, spacing, (0.0, 1.0, 0.0))
    scat = _uniform_on_grid(rng, grid, 200)
    amp = _cn(rng, 200, 1.0)
    return synthesize(f"job{j}", track, ns, grid, scat, amp, 25.0, rng, range_win=hann(ns),
                      az_win=hann(nc), speed=speed)


WORKLOADS = {"line_small": line_small, "line_jitter": line_jitter, "arc": arc, "helix": helix}


def make(name, seed=0, **kw):
    return WORKLOADS[name](seed=seed, **kw)
