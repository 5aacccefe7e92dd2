"""Writes the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py).

mathtest_vectors.json
    The reference's own known answers for the math helpers, restated as data
    from tdctest/mathtest.cpp:12-137 (inputs and expected outputs only).
line_small_plain.npz
    The line_small scene (64 chirps x 128 samples, 32 x 32 pixels) in the
    reference's float32 input layout, and its FP64 oracle image.
doppler_small.npz
    A 300-chirp (three chunks, the last one partial) x 64-sample scene with a
    moving platform, the Doppler-centroid window with taper and a band that
    skips some (pixel, chunk) pairs: inputs, the FP64 oracle image and the
    oracle's processed-chunk count.
precision_ragged.npz
    130 chirps x 100 samples, 13 x 21 pixels, FP64 inputs, and the FP64
    oracle image (precision-form fixture).

The images come from the oracle, whose formulas are pinned independently by
tests/test_oracle.py (closed forms, scipy windows, the known answers above,
an mpmath brute force); these files pin regressions and give the GPU tests
fixed inputs.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "sarlab-tdc_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def mathtest_vectors():
    h = float(np.sqrt(np.float32(2.0)) / np.float32(2.0))
    return {
        "source": "tdctest/mathtest.cpp:12-137",
        "norm3": [{"in": [2.0, 4.0, 6.0], "out": float(np.sqrt(56.0))}],
        "float2_add": [
            {"a": [1.0, 2.0], "b": [3.0, 5.0], "out": [4.0, 7.0]},
            {"a": [1.0, 2.0], "b": [-3.0, -5.0], "out": [-2.0, -3.0]},
            {"a": [0.0, 0.0], "b": [1.0, 1.0], "out": [1.0, 1.0]},
            {"a": [1.0, 1.0], "b": [0.0, 0.0], "out": [1.0, 1.0]},
        ],
        "float3_sub": [
            {"a": [1.0, 2.0, 3.0], "b": [4.0, 8.0, 12.0], "out": [-3.0, -6.0, -9.0]},
            {"a": [1.0, 2.0, 3.0], "b": [-4.0, -8.0, -12.0], "out": [5.0, 10.0, 15.0]},
            {"a": [1.0, 1.0, 1.0], "b": [0.0, 0.0, 0.0], "out": [1.0, 1.0, 1.0]},
            {"a": [0.0, 0.0, 0.0], "b": [1.0, 1.0, 1.0], "out": [-1.0, -1.0, -1.0]},
        ],
        # the reference test asserts a recomputed expression; these are the
        # values it states for v3_dot
        "dot3": [
            {"a": [1.0, 0.0, 0.0], "b": [4.0, 8.0, 12.0], "out": 4.0},
            {"a": [0.0, 1.0, 0.0], "b": [4.0, 8.0, 12.0], "out": 8.0},
            {"a": [0.0, 0.0, 1.0], "b": [4.0, 8.0, 12.0], "out": 12.0},
            {"a": [1.0, 1.0, 1.0], "b": [4.0, 8.0, 12.0], "out": 24.0},
        ],
        "q_conj": [{"q": [1.0, 2.0, 3.0, 4.0], "out": [1.0, -2.0, -3.0, -4.0]}],
        "q_prod": [{"a": [1.0, 2.0, 3.0, 4.0], "b": [8.0, 7.0, 6.0, 5.0],
                    "out": [-44.0, 14.0, 48.0, 28.0]}],
        "q_rot": [
            {"q": [h, 0.0, 0.0, h], "v": [1.0, 0.0, 0.0], "out": [0.0, 1.0, 0.0]},
            {"q": [h, 0.0, 0.0, -h], "v": [1.0, 0.0, 0.0], "out": [0.0, -1.0, 0.0]},
        ],
    }


def dropin_arrays(d):
    return {k: np.asarray(v) for k, v in d.items()}


def main():
    import oracle
    from sarlabtdc import scene

    with open(os.path.join(HERE, "mathtest_vectors.json"), "w") as fh:
        json.dump(mathtest_vectors(), fh, indent=1)

    # line_small scene, reference float32 layout (per-chirp track stored; the
    # per-sample arrays are broadcasts of it)
    s = scene.line_small()
    d = s.dropin()
    img = oracle.image_f64(d["rawData"], d["sampleTimes"], d["position"], d["focusGrid"],
                           float(d["modRate"]), float(d["startFreq"]))
    np.savez_compressed(
        os.path.join(HERE, "line_small_plain.npz"),
        rawData=d["rawData"], priTimes=d["priTimes"], sampleTimes=d["sampleTimes"],
        track=d["position"][:, 0, :], velocity=d["velocity"][:, 0, :],
        modRate=d["modRate"], startFreq=d["startFreq"], bodyBoresight=d["bodyBoresight"],
        focusGrid=d["focusGrid"], image_f64=img)

    # Doppler window + skip
    rng = np.random.default_rng(5)
    npri, ns = 300, 64
    track = np.zeros((npri, 3))
    track[:, 0] = np.linspace(-0.5, 0.5, npri)
    grid = scene.plane_grid(16, 16, 5e-3, (0.0, 0.5, 0.0))
    scat = scene._uniform_on_grid(rng, grid, 5)
    amp = scene._cn(rng, 5, 1.0)
    sd = scene.synthesize("doppler", track, ns, grid, scat, amp, 30.0, rng, speed=40.0)
    dd = sd.dropin()
    dop = dict(velocity=dd["velocity"], attitude=dd["attitude"], boresight=dd["bodyBoresight"],
               center=0.0, bandwidth=8e3, centroid=True, taper=True)
    rw = oracle.range_window(ns)
    img = oracle.image_f64(dd["rawData"], dd["sampleTimes"], dd["position"], dd["focusGrid"],
                           float(dd["modRate"]), float(dd["startFreq"]), range_win=rw, doppler=dop)
    g = dd["focusGrid"].reshape(-1, 3).astype(np.float64)
    processed = sum(oracle.chunk_processed(dd["rawData"], dd["sampleTimes"], dd["position"],
                                           float(dd["modRate"]), float(dd["startFreq"]), dop, c,
                                           g[i]) for i in range(g.shape[0]) for c in range(3))
    np.savez_compressed(
        os.path.join(HERE, "doppler_small.npz"),
        rawData=dd["rawData"], priTimes=dd["priTimes"], sampleTimes=dd["sampleTimes"],
        track=dd["position"][:, 0, :], velocity=dd["velocity"][:, 0, :],
        modRate=dd["modRate"], startFreq=dd["startFreq"], bodyBoresight=dd["bodyBoresight"],
        focusGrid=dd["focusGrid"], dopplerWinCenter=np.float32(0.0),
        dopplerBandwidth=np.float32(8e3), dopCentroidWin=True, applyRangeWin=True,
        dopplerWinTaper=True, image_f64=img, processed_chunks=processed)

    # precision form, ragged shapes, FP64 inputs
    rng = np.random.default_rng(3)
    npri, ns = 130, 100
    track = np.zeros((npri, 3))
    track[:, 0] = np.linspace(-0.2, 0.2, npri)
    track += rng.normal(scale=1e-4, size=track.shape)
    grid = scene.plane_grid(13, 21, 4e-3, (0.01, 0.6, 0.0))
    scat = scene._uniform_on_grid(rng, grid, 6)
    amp = scene._cn(rng, 6, 1.0)
    sp = scene.synthesize("ragged", track, ns, grid, scat, amp, 30.0, rng, speed=10.0)
    img = oracle.image_f64(sp.raw, sp.sample_times, sp.track, sp.grid, sp.mod_rate, sp.f0)
    np.savez_compressed(
        os.path.join(HERE, "precision_ragged.npz"),
        rawData=sp.raw, priTimes=sp.pri_times, sampleTimes=sp.sample_times, position=sp.track,
        modRate=sp.mod_rate, startFreq=sp.f0, focusGrid=sp.grid, image_f64=img)
    print("golden fixtures written to", HERE)


def load_dropin(name):
    """Expand a float32 fixture to the reference's per-sample layout."""
    z = np.load(os.path.join(HERE, name))
    nc, ns = z["rawData"].shape
    pos = np.ascontiguousarray(np.broadcast_to(z["track"][:, None, :], (nc, ns, 3)))
    vel = np.ascontiguousarray(np.broadcast_to(z["velocity"][:, None, :], (nc, ns, 3)))
    att = np.zeros((nc, ns, 4), np.float32)
    att[..., 0] = 1.0
    d = dict(rawData=z["rawData"], priTimes=z["priTimes"], sampleTimes=z["sampleTimes"],
             position=pos, velocity=vel, attitude=att, modRate=z["modRate"][()],
             startFreq=z["startFreq"][()], bodyBoresight=z["bodyBoresight"],
             focusGrid=z["focusGrid"])
    return d, z


if __name__ == "__main__":
    main()
