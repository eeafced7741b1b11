"""The oracle against the reference's only stored projector outputs: the linear spectral-extension
projections spectral_lin_simple.den and spectral_lin_composite.den (testing/testData/
spectralExtension, projectortest.cpp:41-147), made by the reference's RayCasterProjector on a GPU
OpenCL runtime. ref32 projects the ball densities of the reference's case, and the oracle's
restatement of SpectralEffectsExtension::projectLinear combines them with the spectral inputs
derived from the reference's TASMIP table and attenuation database (golden/
make_spectral_inputs.py).

Bars: the reference's own (|mean difference| < 0.01, variance < 0.01), and a far tighter one on
the largest difference: the extinctions (about 1.6-2.9) agree to within 2e-4 on every pixel
(measured 7.2e-5 and 9.0e-5). A wrong spectrum or attenuation table, a wrong ray step or a
voxelisation off by half a voxel moves the centre pixels by 1e-3 or more."""
import numpy as np
import pytest

import oracle
import spectral_case as S


@pytest.fixture(scope="module")
def density_projs():
    pm = S.geometry()
    water = S.ball(40.0, 0.5, 1.0)
    bone = S.ball(30.0, 0.5, 0.3)
    out = []
    for vol in (water, bone):
        p, _ = oracle.project(vol, (0.5, 0.5, 0.5), (0.0, 0.0, 0.0), pm, 50, 50, 0.3)
        out.append(p[:, 0])
    return out


def test_ball_voxelisation():
    b = S.ball(40.0, 0.5, 1.0)
    assert b.shape == (160, 160, 160) and b.max() == 1.0
    # a digital ball of 80 voxels radius: the volume is close to 4/3 pi r^3
    assert abs(b.sum() / (4.0 / 3.0 * np.pi * 80 ** 3) - 1.0) < 0.01
    # squared index distances from the centre 79.5: 3 x 44.5^2 = 5940.75 <= 6400 < 3 x 49.5^2
    assert b[80, 80, 0] == 1.0 and b[35, 35, 35] == 1.0 and b[30, 30, 30] == 0.0 and b[0, 0, 0] == 0.0


def test_spectral_inputs_are_consistent():
    d = S.inputs()
    e = np.array(d["energies_keV"])
    inten = np.array(d["intensities"])
    assert d["tube_voltage_kV"] == [80.0, 84.0, 88.0, 92.0, 96.0]
    assert np.allclose(e, 7.5 + 15.0 * np.arange(7)) and d["bin_width_keV"] == 15.0
    # intensities are the normalised spectrum times the bin energy: their energy-weighted mean rises
    # with the tube voltage and no bin above a view's voltage holds photons
    spec = inten / e[:, None]
    assert np.allclose(spec.sum(axis=0), 1.0, atol=1e-6)
    mean_e = (spec * e[:, None]).sum(axis=0)
    assert np.all(np.diff(mean_e) > 0)
    assert spec[6, 0] == 0.0  # 90-105 keV at 80 kV
    # water's mass attenuation near 50 keV (NIST XCOM: 0.2269 cm^2/g at 50 keV, 0.2059 at 60 keV)
    mw = np.array(d["mass_attenuation_cm2_g"]["water"])
    mb = np.array(d["mass_attenuation_cm2_g"]["bone"])
    assert 0.20 < mw[3] < 0.24 and np.all(np.diff(mw) < 0) and np.all(mb > mw)


@pytest.mark.parametrize("name", ["lin_simple", "lin_composite"])
def test_oracle_reproduces_reference_spectral_projections(density_projs, name):
    d = S.inputs()
    mats = ["water"] if name == "lin_simple" else ["water", "bone"]
    mu = np.stack([np.array(d["mass_attenuation_cm2_g"][m], np.float32) * np.float32(0.1) for m in mats],
                  axis=1)  # [bin][material], 1/cm -> 1/mm (spectraleffectsextension.cpp:286)
    got = oracle.spectral_extinction(np.stack(density_projs[:len(mats)]), mu, np.array(d["intensities"]))
    want = S.expected(name)
    mean, var, worst = S.stats(got, want)
    print(f"{name}: mean diff {mean:.2e}, variance {var:.2e}, max |diff| {worst:.2e}")
    assert abs(mean) < 0.01 and var < 0.01  # projectortest.cpp:96-144
    assert worst < 2e-4


@pytest.mark.parametrize("mistake", ["bin_shift", "no_energy_weight", "step", "ball_centre"])
def test_planted_mistakes_fail_the_fixture_bar(density_projs, mistake):
    """Each plausible mistake in the chain moves some pixel of spectral_lin_simple.den by more than
    the 2e-4 bar: attenuation coefficients taken one energy bin off, photon counts instead of
    energy-weighted intensities, a ray step of 0.33 instead of 0.3 voxel, a ball centred at n / 2
    instead of (n - 1) / 2."""
    d = S.inputs()
    mu = np.array(d["mass_attenuation_cm2_g"]["water"], np.float32)[:, None] * np.float32(0.1)
    inten = np.array(d["intensities"])
    proj = density_projs[0]
    if mistake == "bin_shift":
        mu = np.roll(mu, 1, axis=0)
    elif mistake == "no_energy_weight":
        inten = inten / np.array(d["energies_keV"])[:, None]
    elif mistake == "step":
        proj = oracle.project(S.ball(40.0, 0.5, 1.0), (0.5, 0.5, 0.5), (0.0, 0.0, 0.0), S.geometry(), 50, 50,
                              0.33)[0][:, 0]
    elif mistake == "ball_centre":
        n = 160
        i = np.arange(n, dtype=np.float32) - np.float32(n / 2)
        d2 = i[:, None, None] ** 2 + i[None, :, None] ** 2 + i[None, None, :] ** 2
        ball = np.where(d2 > np.float32(80.0 ** 2), np.float32(0.0), np.float32(1.0)).astype(np.float32)
        proj = oracle.project(ball, (0.5, 0.5, 0.5), (0.0, 0.0, 0.0), S.geometry(), 50, 50, 0.3)[0][:, 0]
    got = oracle.spectral_extinction(proj[None], mu, inten)
    worst = S.stats(got, S.expected("lin_simple"))[2]
    print(f"{mistake}: max |diff| {worst:.2e}")
    assert worst > 2e-4
