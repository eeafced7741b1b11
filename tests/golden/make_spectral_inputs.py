"""Writes spectral_extension_inputs.json: the spectral inputs of the reference's spectral-extension
test (testing/src/projectortest.cpp:41-147), for the GPU test that reproduces its stored outputs
(testing/testData/spectralExtension/spectral_lin_{simple,composite}.den, copied next to this file).

The reference derives them from physics models that are not on the projection path; this script
restates those models in the order and precision the reference evaluates them, from the
reference's own data:

  * the X-ray tube's TASMIP spectrum (XrayTube, xraytube.cpp:21, 76, 152): a table over 0 .. 139
    keV whose entries are cubic polynomials in the tube voltage, zero above it and where negative
    (TASMIPSpectrumModel::TASMIPtable, xrayspectrummodels.cpp:353-...). The polynomial
    coefficients (Boone and Seibert's published table) are read from that source file;
  * the materials' mass attenuation coefficients from the database files water.json
    (Composite::Water = 1048) and bone.json (Composite::Bone_Cortical = 1008;
    ctldatabase.cpp:100-140), tabulated data models (type id 30);
  * TabulatedDataModel::valueAt / binIntegral (tabulateddatamodel.cpp:106-200): linear
    interpolation and its trapezoid integral, with the reference's edge cases;
  * IntervalDataSeries::sampledFromModel / normalizeByIntegral (intervaldataseries.cpp:5-60) and
    AbstractSource::spectrum(range, n) (abstractsource.h:234-247);
  * RadiationEncoder::spectralInformation / SpectralInformation::extractViewSpectrum
    (radiationencoder.cpp:168-280): one energy range covering every view's [0, kV], bins of the
    requested width (15 keV here), per-view intensity = normalised spectrum x bin energy, the total
    intensity their sum (no detector response model);
  * AbstractIntegrableDataModel::meanValue (abstractdatamodel.cpp:122-126): the bin's mean mass
    attenuation coefficient = binIntegral / bin width (cm^2/g).

The test's five views run the tube at 80 + 20/5 * v kV (projectortest.cpp:71-77). Floating-point
steps are evaluated in float32 where the reference uses float.

Run in the build container (it reads /root/reference): python tests/golden/make_spectral_inputs.py
"""
from __future__ import annotations

import bisect
import json
import math
import re
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
HERE = Path(__file__).resolve().parent
f32 = np.float32


def tasmip_coefficients() -> np.ndarray:
    src = (REF / "modules/src/models/xrayspectrummodels.cpp").read_text()
    body = src[src.index("TASMIPSpectrumModel::TASMIPtable"):]
    body = body[body.index("coeff[nbBins][4] = {"):body.index("static_assert")]
    num = r"([+-]\d+\.\d+e[+-]\d+)f"
    rows = re.findall(r"\{\s*" + r"\s*,\s*".join([num] * 4) + r"\s*\}", body)
    c = np.array(rows, dtype=np.float32)
    assert c.shape == (140, 4), c.shape
    return c


def tasmip_table(coeff: np.ndarray, kv: float) -> tuple[list[float], list[float]]:
    kv = f32(kv)
    keys, vals = [], []
    for e in range(140):
        if f32(e) - f32(0.5) > kv:
            v = f32(0.0)
        else:
            a = coeff[e]
            k2 = kv * kv
            v = a[0] + a[1] * kv + a[2] * k2 + a[3] * k2 * kv
            v = v if v > 0 else f32(0.0)
        keys.append(f32(e))
        vals.append(f32(v))
    return keys, vals


def fuzzy_compare(a, b) -> bool:  # qFuzzyCompare for float
    return abs(float(a) - float(b)) * 100000.0 <= min(abs(float(a)), abs(float(b)))


class Tabulated:
    """TabulatedDataModel: sorted keys, linear interpolation, zero outside the keys."""

    def __init__(self, keys, vals):
        order = np.argsort(np.asarray(keys, np.float32), kind="stable")
        self.k = [f32(keys[i]) for i in order]
        self.v = [f32(vals[i]) for i in order]

    def value_at(self, pos):
        pos = f32(pos)
        i = bisect.bisect_left(self.k, pos)
        if i < len(self.k) and self.k[i] == pos:
            return self.v[i]
        if i == 0 or i == len(self.k):
            return f32(0.0)
        w = (self.k[i] - pos) / (self.k[i] - self.k[i - 1])
        return f32(self.v[i - 1] * w + self.v[i] * (f32(1.0) - w))

    def bin_integral(self, position, width):
        position, width = f32(position), f32(width)
        frm = f32(position - f32(0.5) * width)
        to = f32(position + f32(0.5) * width)
        lo = bisect.bisect_left(self.k, frm)
        up = bisect.bisect_left(self.k, to)
        if lo == len(self.k) or to < self.k[0]:
            return f32(0.0)
        if lo == up and up != 0:
            return f32(self.value_at(position) * width)
        ret = f32(f32(0.5) * (self.value_at(frm) + self.v[lo]) * (self.k[lo] - frm))
        if up == 0:
            return ret
        cur = lo
        while cur != up:
            if cur + 1 == len(self.k) or self.k[cur + 1] > to:
                break
            ret = f32(ret + (self.k[cur + 1] - self.k[cur]) * (f32(0.5) * (self.v[cur + 1] + self.v[cur])))
            cur += 1
        if fuzzy_compare(self.k[cur], to):
            return ret
        last = up - 1
        return f32(ret + f32(0.5) * (self.v[last] + self.value_at(to)) * (to - self.k[last]))

    def mean_value(self, position, width):
        return f32(self.bin_integral(position, width) / f32(width))


def spectrum(model: Tabulated, frm: float, to: float, n: int):
    """AbstractSource::spectrum(range, n): bin integrals at the bins' centres, normalised."""
    width = f32((f32(to) - f32(frm)) / f32(n))
    pt = f32(f32(frm) + f32(0.5) * width)
    pts, vals = [], []
    for _ in range(n):
        pts.append(pt)
        vals.append(model.bin_integral(pt, width))
        pt = f32(pt + width)
    total = f32(0.0)
    for v in vals:
        total = f32(total + v)
    return pts, [f32(v / total) for v in vals], width


def material(name: str) -> Tabulated:
    d = json.loads((REF / "database/attenuation_spectra" / f"{name}.json").read_text())
    assert d["type-id"] == 30, d["type-id"]
    data = d["parameters"]["data"]
    return Tabulated([e for e, _ in data], [m for _, m in data])


def main() -> None:
    coeff = tasmip_coefficients()
    views = 5
    kv = [80.0 + 20.0 / views * v for v in range(views)]  # projectortest.cpp:71-77
    bin_width = 15.0  # setSpectralSamplingResolution(15.0f), projectortest.cpp:86
    lo, hi = min(0.0, *[0.0] * views), max(kv)  # every view's energy range is [0, kV]
    n_bins = max(int(math.ceil(f32(hi - lo) / f32(bin_width))), 1)
    end = lo + n_bins * bin_width
    energies, intens = None, np.zeros((n_bins, views))
    for v, k in enumerate(kv):
        pts, spec, width = spectrum(Tabulated(*tasmip_table(coeff, k)), lo, end, n_bins)
        energies = pts
        for b in range(n_bins):
            intens[b, v] = float(spec[b]) * float(pts[b])  # intensity = spectrum value x energy
    water, bone = material("water"), material("bone")
    out = {
        "source": "derived from the reference's TASMIP table, database/attenuation_spectra/{water,bone}.json "
                  "and the spectral-information steps of RadiationEncoder (see make_spectral_inputs.py)",
        "tube_voltage_kV": kv,
        "bin_width_keV": float(width),
        "energies_keV": [float(e) for e in energies],
        "intensities": intens.tolist(),  # [bin][view]
        "mass_attenuation_cm2_g": {
            "water": [float(water.mean_value(e, width)) for e in energies],
            "bone": [float(bone.mean_value(e, width)) for e in energies],
        },
    }
    (HERE / "spectral_extension_inputs.json").write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
