"""GPU results against the committed golden fixtures (tests/golden/*.npz)."""
import os

import numpy as np
import pytest

from golden.make_golden import load_dropin

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _proc(tdc, d):
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    p.setRawData(d["rawData"], d["priTimes"], d["sampleTimes"], d["position"], d["velocity"],
                 d["attitude"], d["modRate"], d["startFreq"], d["bodyBoresight"])
    p.setFocusGrid(d["focusGrid"])
    return p


def test_golden_line_small_plain(tdc, orc):
    d, z = load_dropin("line_small_plain.npz")
    p = _proc(tdc, d)
    p.start(0.0, 1e30, False, False, False)
    img = p.getFocusedImage()
    assert orc.rel_l2(img, z["image_f64"]) < 1e-4
    assert orc.peak_phase_error(img, z["image_f64"]) < 1e-4


def test_golden_doppler_small(tdc, orc):
    d, z = load_dropin("doppler_small.npz")
    p = _proc(tdc, d)
    p.start(float(z["dopplerWinCenter"]), float(z["dopplerBandwidth"]), bool(z["dopCentroidWin"]),
            bool(z["applyRangeWin"]), bool(z["dopplerWinTaper"]))
    img = p.getFocusedImage()
    st = p.stats()
    assert st["doppler_active"]
    assert st["processed_chunks"] == int(z["processed_chunks"])
    assert orc.rel_l2(img, z["image_f64"]) < 5e-3


@pytest.mark.parametrize("form", ["precision", "plain"])
def test_golden_precision_ragged(tdc, orc, form):
    z = np.load(os.path.join(HERE, "golden", "precision_ragged.npz"))
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    if form == "precision":
        p.setRawDataF64(z["rawData"], z["priTimes"], z["sampleTimes"], z["position"],
                        float(z["modRate"]), float(z["startFreq"]))
        p.setFocusGridF64(z["focusGrid"])
        tol = 1e-4
    else:
        nc, ns = z["rawData"].shape
        pos = np.ascontiguousarray(np.broadcast_to(z["position"].astype(np.float32)[:, None, :],
                                                   (nc, ns, 3)))
        att = np.zeros((nc, ns, 4), np.float32)
        att[..., 0] = 1
        p.setRawData(z["rawData"], z["priTimes"].astype(np.float32),
                     z["sampleTimes"].astype(np.float32), pos, np.zeros_like(pos), att,
                     float(z["modRate"]), float(z["startFreq"]), np.array([0, 1, 0], np.float32))
        p.setFocusGrid(z["focusGrid"].astype(np.float32))
        tol = 2e-3  # float32 inputs against the FP64-input image
    p.start(0.0, 1e30, False, False, False)
    img = p.getFocusedImage()
    assert orc.rel_l2(img, z["image_f64"]) < tol
