"""The range-profile engine's kernel constants: the device literals of every
built kernel (taps, oversampling) in csrc/tdc_tiles.h are exp(-b u^2) at the
b = k pi / (taps / 2) the host's correction table uses (profile_bfactor in
csrc/tdc_internal.h), so the interpolation weights and the deconvolution that
undoes them describe the same kernel."""
import math
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "sarlab-tdc_amd", "csrc")


def _kernels():
    text = open(os.path.join(CSRC, "tdc_tiles.h")).read()
    out = {}
    for m in re.finditer(r"template <> struct Kernel<(\d+), (\d+)> \{(.*?)\n\};", text, re.S):
        taps, over, body = int(m.group(1)), int(m.group(2)), m.group(3)
        e0 = float(re.search(r"kE0 = (-?[0-9.e+-]+)F", body).group(1))
        e1 = float(re.search(r"kE1 = (-?[0-9.e+-]+)F", body).group(1))
        arr = re.search(r"kC\[\d+\] = \{(.*?)\}", body, re.S).group(1)
        kc = [float(x.strip().rstrip("F")) for x in arr.split(",") if x.strip()]
        out[(taps, over)] = (e0, e1, kc)
    return out


def _bfactor(taps, over):
    text = open(os.path.join(CSRC, "tdc_internal.h")).read()
    body = re.search(r"inline double profile_bfactor\(int taps, int over\)\s*\{(.*?)\n\}", text,
                     re.S).group(1)
    m6, m4 = map(float, re.search(r"if \(over == 4\) return taps == 6 \? ([0-9.]+) : ([0-9.]+);",
                                  body).groups())
    m12, mother = map(float, re.search(r"return taps == 12 \? ([0-9.]+) : ([0-9.]+);",
                                       body).groups())
    if over == 4:
        return m6 if taps == 6 else m4
    return m12 if taps == 12 else mother


def test_every_built_kernel_has_constants():
    assert set(_kernels()) == {(8, 2), (10, 2), (12, 2), (6, 4), (8, 4)}


def test_constants_follow_the_host_kernel():
    for (taps, over), (e0, e1, kc) in _kernels().items():
        b = _bfactor(taps, over) * math.pi / (taps / 2)
        lead = taps // 2 - 1
        assert math.isclose(e0, -b * math.log2(math.e), rel_tol=1e-12), (taps, over)
        assert math.isclose(e1, 2 * b * math.log2(math.e), rel_tol=1e-12), (taps, over)
        assert len(kc) == taps
        for t, c in enumerate(kc):
            assert math.isclose(c, math.exp(-b * (lead - t) ** 2), rel_tol=1e-12), (taps, over, t)
        # symmetric about the centre tap, which weighs one
        assert kc[lead] == 1.0 and kc[:lead] == kc[lead + 1:2 * lead + 1][::-1]
