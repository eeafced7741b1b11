"""The HIP path against the reference's stored spectral-extension projections
(spectral_lin_simple.den, spectral_lin_composite.den; projectortest.cpp:41-147): the ray caster
projects the ball densities of the reference's case on the GPU and the spectral extension
combines them with the spectral inputs derived from the reference's TASMIP table and attenuation
database. Bars: the reference's (|mean difference| < 0.01, variance < 0.01) and, on every pixel,
2e-4 against the stored values and 2e-5 against the oracle's restatement of the same chain."""
import numpy as np
import pytest

import oracle
import spectral_case as S
from ctl_amd import RayCasterProjector, VoxelVolume
from ctl_amd.extensions import SpectralEffectsExtension, SpectralInformation, SpectralVolume

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["lin_simple", "lin_composite"])
def test_spectral_extension_reproduces_reference_projections(name):
    d = S.inputs()
    pm = S.geometry()
    mats = ["water"] if name == "lin_simple" else ["water", "bone"]
    balls = {"water": S.ball(40.0, 0.5, 1.0), "bone": S.ball(30.0, 0.5, 0.3)}
    proj = RayCasterProjector()  # default settings: ray sampling 0.3, interpolation
    proj.configure((pm, 50, 50))
    ext = SpectralEffectsExtension(proj, SpectralInformation(d["energies_keV"], d["intensities"]))
    vols = [SpectralVolume(VoxelVolume(balls[m], (0.5, 0.5, 0.5)), d["mass_attenuation_cm2_g"][m], m) for m in mats]
    got = ext.projectComposite(vols).data[:, 0]
    want = S.expected(name)
    mean, var, worst = S.stats(got, want)
    mu = np.stack([np.array(d["mass_attenuation_cm2_g"][m], np.float32) * np.float32(0.1) for m in mats], axis=1)
    dens = [oracle.project(balls[m], (0.5, 0.5, 0.5), (0.0, 0.0, 0.0), pm, 50, 50, 0.3)[0][:, 0] for m in mats]
    ref = oracle.spectral_extinction(np.stack(dens), mu, np.array(d["intensities"]))
    vs_oracle = float(np.abs(got - ref).max())
    print(f"{name}: vs stored mean {mean:.2e}, variance {var:.2e}, max |diff| {worst:.2e}; "
          f"vs oracle max |diff| {vs_oracle:.2e}")
    assert got.shape == want.shape == (5, 50, 50)
    assert abs(mean) < 0.01 and var < 0.01  # projectortest.cpp:96-144
    assert worst < 2e-4 and vs_oracle < 2e-5
