"""P-matrix encoder against CTL's golden fixtures (geometrytest.cpp:35-56, tolerance :104-124)
and closed forms (projectionmatrixtest.cpp:34-52). Host code only, no GPU."""
import numpy as np
import pytest

import oracle
from ctl_amd import (DEG, AcquisitionSetup, CarmGantry, CTSystem, CylindricalDetector,
                     FlatPanelDetector, GeometryEncoder, HelicalTrajectory, ShortScanTrajectory,
                     TubularGantry, WobbleTrajectory)


def read_den_pmats(path, modules):
    raw = path.read_bytes()
    rows, cols, count = np.frombuffer(raw[:6], "<u2")
    assert (rows, cols) == (3, 4)
    P = np.frombuffer(raw[6:], "<f8").reshape(count, 3, 4)
    return P.reshape(count // modules, modules, 3, 4)


def normalized(P):
    n = np.linalg.norm(P[..., 2, :3], axis=-1)[..., None, None]
    s = np.sign(np.linalg.det(P[..., :3]))[..., None, None]
    return P / (n * s)


def mean_frobenius_diff(a, b):
    d = normalized(a) - normalized(b)
    return np.mean(np.sqrt(np.sum(d * d, axis=(-2, -1))))


def test_det_geometry_fixture_tubular(golden):
    det = CylindricalDetector((16, 64), (1.0, 1.0), 40, 1.0 * DEG, 0.2)
    setup = AcquisitionSetup(CTSystem(TubularGantry(1000.0, 550.0, 0.0, 90.0 * DEG, 0.0), det), 10)
    setup.applyPreparationProtocol(HelicalTrajectory(3.6 * DEG, 1.0))
    got = GeometryEncoder.encodeFullGeometry(setup)
    ref = read_den_pmats(golden / "tubeGeo.den", 40)
    assert got.shape == ref.shape == (10, 40, 3, 4)
    assert mean_frobenius_diff(ref, got) < 1e-8


def test_det_geometry_fixture_carm_wobble(golden):
    det = CylindricalDetector((16, 64), (1.0, 1.0), 40, 1.0 * DEG, 0.2)
    setup = AcquisitionSetup(CTSystem(CarmGantry(1200.0), det), 30)
    setup.applyPreparationProtocol(WobbleTrajectory(200.0 * DEG, 750.0, 0.0, 20.0 * DEG, 5.0))
    got = GeometryEncoder.encodeFullGeometry(setup)
    ref = read_den_pmats(golden / "cArmGeo.den", 40)
    assert got.shape == ref.shape == (30, 40, 3, 4)
    assert mean_frobenius_diff(ref, got) < 1e-8


def test_a_perturbed_encoder_fails_the_fixture(golden):
    """The fixture check is sharp: a 0.01 deg change of the wobble angle is caught."""
    det = CylindricalDetector((16, 64), (1.0, 1.0), 40, 1.0 * DEG, 0.2)
    setup = AcquisitionSetup(CTSystem(CarmGantry(1200.0), det), 30)
    setup.applyPreparationProtocol(WobbleTrajectory(200.0 * DEG, 750.0, 0.0, 20.01 * DEG, 5.0))
    got = GeometryEncoder.encodeFullGeometry(setup)
    assert mean_frobenius_diff(read_den_pmats(golden / "cArmGeo.den", 40), got) > 1e-6


def test_carm_circle_source_positions_and_projection():
    setup = AcquisitionSetup(CTSystem(CarmGantry(1200.0), FlatPanelDetector((256, 128), (1.0, 1.0))), 36)
    setup.applyPreparationProtocol(ShortScanTrajectory(750.0, 0.0, 350.0 * DEG))
    P = GeometryEncoder.encodeFullGeometry(setup)[:, 0]
    for v in range(36):
        c = oracle.source_position(P[v])
        assert np.linalg.norm(c) == pytest.approx(750.0, rel=1e-12)
        assert abs(c[2]) < 1e-9  # circle in the z = 0 plane
        assert np.allclose(P[v] @ np.append(c, 1.0), 0.0, atol=1e-9)
        # the isocentre projects onto the detector centre ((C-1)/2, (R-1)/2)
        h = P[v] @ np.array([0.0, 0.0, 0.0, 1.0])
        assert np.allclose(h[:2] / h[2], [127.5, 63.5], atol=1e-9)
        # magnification: 1 mm at the isocentre spans 1200/750 pixels
    ang = [np.arctan2(*oracle.source_position(P[v])[[1, 0]]) for v in range(36)]
    steps = np.diff(np.unwrap(ang))
    assert np.allclose(steps, 10.0 * DEG, atol=1e-12)
