"""Multi-projection callers on resident volumes: projectComposite (abstractprojector.h:99-112)
accumulated on the device; ArealFocalSpotExtension (arealfocalspotextension.cpp:82-166, 257-292)
as one launch over all focal-spot samples plus a device average; SpectralEffectsExtension's
linear path (spectraleffectsextension.cpp:232-352) as one projection per material plus a device
sum over energy bins; DynamicProjectorExtension (dynamicprojectorextension.cpp:39-66) on linear
and general dynamic volumes. Checked against per-volume projections and against the oracle."""
import numpy as np
import pytest

import oracle
from ctl_amd import (AcquisitionSetup, ArealFocalSpotExtension, CarmGantry, CTSystem, DynamicProjectorExtension,
                     FlatPanelDetector, GeometryEncoder, LinearDynamicVolume, RayCasterProjector, ResidentVolume,
                     ShortScanTrajectory, SpectralEffectsExtension, SpectralInformation, SpectralVolume,
                     VoxelVolume)
from ctl_amd import workloads as W

pytestmark = pytest.mark.gpu


def small_setup(views=4, spot=(0.0, 0.0)):
    setup = AcquisitionSetup(CTSystem(CarmGantry(1200.0), FlatPanelDetector((256, 128), (1.0, 1.0)), spot), views)
    setup.applyPreparationProtocol(ShortScanTrajectory(750.0, 0.0, 270.0 * np.pi / 180.0))
    return setup


def test_project_composite_is_the_sum_of_the_projections():
    w = W.small(views=4)
    a = VoxelVolume(w.volume, (1.0, 1.0, 1.0))
    b = VoxelVolume(np.ascontiguousarray(w.volume[::-1, ::-1, :] * 0.5), (0.8, 0.8, 0.8), (3.0, -2.0, 1.0))
    ra, rb = ResidentVolume(a), ResidentVolume(b, "f16")
    p = RayCasterProjector()
    p.settings().raySampling = 0.1
    p.configure(small_setup())
    pa, pb = p.project(ra).data, p.project(rb).data
    comp = p.projectComposite([ra, rb]).data
    assert np.array_equal(comp, pa + pb)  # FP32 sum in volume order
    assert p.isLinear()


def test_project_composite_against_oracle_sum():
    """projectComposite (abstractprojector.h:99-112) of three sub-volumes with their own sizes,
    voxel sizes and offsets (an FP32 phantom, an FP16 insert of 0.8 mm voxels off centre, an FP32
    slab of anisotropic voxels) against the sum of the oracle's projections of each: ref32 for the
    FP32 parts, ref64 for the FP16 insert (FP16 storage is measured against ref64), summed in FP64."""
    w = W.small(views=36)
    pm = w.pmats[::6]
    a = np.ascontiguousarray(w.volume)
    b = np.ascontiguousarray(W.ellipsoid_phantom((40, 36, 30), (0.8, 0.8, 0.8), 6, 91, semi_axes=(2.0, 8.0)))
    c = np.ascontiguousarray(W.ellipsoid_phantom((48, 20, 52), (1.1, 1.5, 0.9), 5, 92, semi_axes=(3.0, 9.0)))
    parts = [(a, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), "f32"), (b, (0.8, 0.8, 0.8), (4.0, -3.0, 2.5), "f16"),
             (c, (1.1, 1.5, 0.9), (-6.0, 2.0, -1.5), "f32")]
    p = RayCasterProjector()
    p.settings().raySampling = w.ray_sampling
    p.configure((pm, w.channels, w.rows))
    comp = p.projectComposite([ResidentVolume(VoxelVolume(v, vs, off), st) for v, vs, off, st in parts]).data
    ref = np.zeros(comp.shape)
    for v, vs, off, st in parts:
        r, _ = oracle.project(v, vs, off, pm, w.channels, w.rows, w.ray_sampling, fp64=(st != "f32"))
        ref += r
    e, m = oracle.rel_l2(ref, comp), oracle.max_abs_rel(ref, comp)
    print(f"composite of 3 sub-volumes vs oracle sum: rel L2 {e:.3e}, max abs rel {m:.3e}")
    assert e <= 5e-4  # the FP16 part's storage rounding; the FP32 parts alone stay below 1e-4
    fp32_only = p.projectComposite([ResidentVolume(VoxelVolume(v, vs, off), st) for v, vs, off, st in parts
                                    if st == "f32"]).data
    ref32 = sum(oracle.project(v, vs, off, pm, w.channels, w.rows, w.ray_sampling)[0].astype(np.float64)
                for v, vs, off, st in parts if st == "f32")
    e32, m32 = oracle.rel_l2(ref32, fp32_only), oracle.max_abs_rel(ref32, fp32_only)
    print(f"composite of the FP32 sub-volumes vs ref32 sum: rel L2 {e32:.3e}, max abs rel {m32:.3e}")
    assert e32 <= 1e-4 and m32 <= 1e-3


def test_fp32_drop_in_against_the_reference_cpu_projector():
    """The FP32 drop-in against refcpu, the restatement of the reference's RayCasterProjectorCPU
    (raycasterprojectorcpu.cpp:132-285: FP64 positions, FP32 weights), under the drop-in bars, on
    the small workload and on two views of the 256^3 / 1024 x 768 one."""
    w = W.small(views=36)
    rc, _ = oracle.project(w.volume, w.voxel_mm, w.offset_mm, w.pmats, w.channels, w.rows, w.ray_sampling,
                           cpu_projector=True)
    p = RayCasterProjector()
    p.settings().raySampling = w.ray_sampling
    p.configure((w.pmats, w.channels, w.rows))
    got = p.project(ResidentVolume(VoxelVolume(w.volume, w.voxel_mm, w.offset_mm))).data
    e, m = oracle.rel_l2(rc, got), oracle.max_abs_rel(rc, got)
    print(f"small f32 vs refcpu: rel L2 {e:.3e}, max abs rel {m:.3e}")
    assert e <= 1e-4 and m <= 1e-3
    real = W.real(views=360)
    pick = [0, 137]
    rc, _ = oracle.project(real.volume, real.voxel_mm, real.offset_mm, real.pmats, real.channels, real.rows,
                           real.ray_sampling, views=pick, threads=16, cpu_projector=True)
    p.settings().raySampling = real.ray_sampling
    p.configure((real.pmats[pick], real.channels, real.rows))
    got = p.project(ResidentVolume(VoxelVolume(real.volume, real.voxel_mm, real.offset_mm))).data
    e, m = oracle.rel_l2(rc, got), oracle.max_abs_rel(rc, got)
    print(f"256^3 views {pick} f32 vs refcpu: rel L2 {e:.3e}, max abs rel {m:.3e}")
    assert e <= 1e-4 and m <= 1e-3


@pytest.mark.parametrize("low_extinction", [False, True])
def test_areal_focal_spot_against_oracle(low_extinction):
    w = W.small(views=4)
    spot = (2.0, 1.0)
    setup = small_setup(spot=spot)
    nested = RayCasterProjector()
    nested.settings().raySampling = 0.1
    ext = ArealFocalSpotExtension(nested, (2, 2), low_extinction)
    ext.configure(setup)
    got = ext.project(VoxelVolume(w.volume, (1.0, 1.0, 1.0))).data
    assert got.shape == (4, 1, 128, 256) and ext.isLinear() == low_extinction
    subs = []
    for x, y in ext.discretizationGrid():
        pm = GeometryEncoder.encodeFullGeometry(setup, (x * spot[0], y * spot[1], 0.0))
        r32, _ = oracle.project(w.volume, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), pm, 256, 128, 0.1)
        subs.append(r32.astype(np.float64))
    subs = np.stack(subs)
    ref = subs.mean(0) if low_extinction else -np.log(np.exp(-subs).mean(0))
    e = oracle.rel_l2(ref, got)
    print(f"areal focal spot (low extinction {low_extinction}): rel L2 vs ref32 {e:.3e}")
    assert e <= 1e-4
    # the samples differ: the extension is not the projection of the centre geometry
    p0 = RayCasterProjector()
    p0.settings().raySampling = 0.1
    p0.configure(setup)
    assert oracle.rel_l2(p0.project(VoxelVolume(w.volume, (1.0, 1.0, 1.0))).data, got) > 1e-4


def test_spectral_effects_against_oracle():
    """SpectralEffectsExtension's linear path: one projection per material density and the bins
    summed on the device, against ref32 projections combined in FP64 (oracle.spectral_extinction)."""
    w = W.small(views=4)
    setup = small_setup()
    d1 = np.ascontiguousarray(w.volume / w.volume.max() * 1.2, np.float32)           # water-like
    d2 = np.ascontiguousarray(w.volume[::-1, :, ::-1] / w.volume.max() * 0.8, np.float32)  # bone-like
    energies = [30.0, 50.0, 70.0, 90.0, 110.0]
    rng = np.random.default_rng(21)
    inten = rng.uniform(0.5, 2.0, (5, 4))
    inten[3] = 0.0  # a bin without photons is skipped
    ma1 = [0.375, 0.227, 0.193, 0.171, 0.160]  # cm^2/g
    ma2 = [1.33, 0.425, 0.253, 0.196, 0.172]
    resp = [0.7, 0.9, 1.0, 1.0, 0.95]
    nested = RayCasterProjector()
    nested.settings().raySampling = 0.1
    ext = SpectralEffectsExtension(nested, SpectralInformation(energies, inten))
    ext.setDetectorResponse(resp)
    ext.configure(setup)
    got = ext.projectComposite([SpectralVolume(VoxelVolume(d1, (1.0, 1.0, 1.0)), ma1),
                                SpectralVolume(VoxelVolume(d2, (1.0, 1.0, 1.0)), ma2)]).data
    pm = GeometryEncoder.encodeFullGeometry(setup)
    p1, _ = oracle.project(d1, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), pm, 256, 128, 0.1)
    p2, _ = oracle.project(d2, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), pm, 256, 128, 0.1)
    mu = 0.1 * np.array([ma1, ma2]).T
    ref = oracle.spectral_extinction(np.stack([p1, p2]), mu, inten, resp)
    e = oracle.rel_l2(ref, got)
    print(f"spectral effects: rel L2 vs ref32 projections + FP64 spectrum {e:.3e}, max extinction {ref.max():.2f}")
    assert got.shape == (4, 1, 128, 256) and e <= 1e-4
    assert not ext.isLinear()


def test_spectral_effects_nonlinear_nested_projector():
    """The non-linear path (spectraleffectsextension.cpp:292-395) with ArealFocalSpotExtension
    averaging in the intensity domain as the nested projector: per bin and material the areal
    projection of the attenuation volume, against the oracle (ref32 at each focal-spot position,
    averaged in the intensity domain in FP64, summed over materials, combined over bins in FP64).
    With the linear ray caster forced onto the same path the result equals the linear path's to
    FP32 rounding, since F(m rho) = m F(rho)."""
    w = W.small(views=4)
    spot = (2.0, 1.0)
    setup = small_setup(spot=spot)
    d1 = np.ascontiguousarray(w.volume / w.volume.max() * 1.2, np.float32)
    d2 = np.ascontiguousarray(w.volume[::-1, :, ::-1] / w.volume.max() * 0.8, np.float32)
    energies = [30.0, 60.0, 90.0]
    rng = np.random.default_rng(23)
    inten = rng.uniform(0.5, 2.0, (3, 4))
    ma1, ma2 = [0.375, 0.206, 0.169], [1.33, 0.315, 0.186]
    resp = [0.8, 1.0, 0.9]
    vols = [SpectralVolume(VoxelVolume(d1, (1.0, 1.0, 1.0)), ma1), SpectralVolume(VoxelVolume(d2, (1.0, 1.0, 1.0)), ma2)]
    nested = RayCasterProjector()
    nested.settings().raySampling = 0.1
    areal = ArealFocalSpotExtension(nested, (2, 2), False)
    assert not areal.isLinear()
    ext = SpectralEffectsExtension(areal, SpectralInformation(energies, inten))
    ext.setDetectorResponse(resp)
    ext.configure(setup)
    got = ext.projectComposite(vols).data
    grid = areal.discretizationGrid()
    pms = [GeometryEncoder.encodeFullGeometry(setup, (x * spot[0], y * spot[1], 0.0)) for x, y in grid]
    total = 0.0
    for b in range(3):
        q = 0.0
        for d, ma in ((d1, ma1), (d2, ma2)):
            mu = np.ascontiguousarray(d * np.float32(ma[b] * np.float32(0.1)), np.float32)
            subs = np.stack([oracle.project(mu, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), pm, 256, 128, 0.1)[0]
                             .astype(np.float64) for pm in pms])
            q = q + -np.log(np.exp(-subs).mean(0))
        total = total + inten[b].reshape(4, 1, 1, 1) * np.exp(-q) * resp[b]
    ref = np.log(inten.sum(0).reshape(4, 1, 1, 1) / total)
    e = oracle.rel_l2(ref, got)
    print(f"spectral effects, non-linear nested projector: rel L2 vs oracle {e:.3e}")
    assert got.shape == (4, 1, 128, 256) and e <= 1e-4
    # the linear nested projector forced onto the non-linear path agrees with the linear path
    lin = SpectralEffectsExtension(nested, SpectralInformation(energies, inten))
    lin.setDetectorResponse(resp)
    lin.configure(setup)
    a = lin.projectComposite(vols).data
    b_ = lin._project_nonlinear(vols).data
    d = oracle.rel_l2(a, b_)
    print(f"linear vs non-linear path with the linear ray caster: rel L2 {d:.3e}")
    assert d <= 1e-5


def test_dynamic_linear_volume_against_oracle():
    """DynamicProjectorExtension on a LinearDynamicVolume: t_v * P_v(slope) + P_v(offset) on the
    device against ref32 projections of slope * t_v + offset, one view at a time."""
    w = W.small(views=4)
    setup = small_setup()
    for v, t in enumerate([0.0, 12.5, 30.0, 47.0]):
        setup.setTimeStamp(v, t)
    slope = VoxelVolume(np.ascontiguousarray(w.volume * 0.02, np.float32), (1.0, 1.0, 1.0))
    offset = VoxelVolume(np.ascontiguousarray(w.volume[:, ::-1, :], np.float32), (1.0, 1.0, 1.0))
    dyn = LinearDynamicVolume(slope, offset)
    nested = RayCasterProjector()
    nested.settings().raySampling = 0.1
    ext = DynamicProjectorExtension(nested)
    ext.configure(setup)
    got = ext.project(dyn).data
    pm = GeometryEncoder.encodeFullGeometry(setup)
    ref = np.concatenate([oracle.project(dyn.at(setup.timeStamp(v)).data, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0),
                                         pm[v:v + 1], 256, 128, 0.1)[0] for v in range(4)])
    e = oracle.rel_l2(ref, got)
    print(f"dynamic (linear): rel L2 vs ref32 {e:.3e}")
    assert got.shape == (4, 1, 128, 256) and e <= 1e-4
    # the views differ by more than the tolerance: the time stamps matter
    assert oracle.rel_l2(ref[3], ref[0]) > 1e-2


def test_dynamic_general_volume_view_by_view():
    """Any volume with setTime(): one projection per view at that view's time stamp."""
    w = W.small(views=4)
    setup = small_setup()

    class Pulsing:
        def __init__(self, base):
            self.base = base
            self.setTime(0.0)

        def setTime(self, t):
            self.volume = VoxelVolume(np.ascontiguousarray(self.base * np.float32(1.0 + 0.5 * np.sin(t))),
                                      (1.0, 1.0, 1.0))

    nested = RayCasterProjector()
    nested.settings().raySampling = 0.1
    ext = DynamicProjectorExtension(nested)
    ext.configure(setup)  # default time stamps 0, 1, 2, 3
    got = ext.project(Pulsing(w.volume)).data
    pm = GeometryEncoder.encodeFullGeometry(setup)
    ref = np.concatenate([oracle.project(w.volume * np.float32(1.0 + 0.5 * np.sin(float(v))), (1.0, 1.0, 1.0),
                                         (0.0, 0.0, 0.0), pm[v:v + 1], 256, 128, 0.1)[0] for v in range(4)])
    assert oracle.rel_l2(ref, got) <= 1e-4
    # a static volume goes straight to the nested projector
    static = ext.project(VoxelVolume(w.volume, (1.0, 1.0, 1.0))).data
    full, _ = oracle.project(w.volume, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), pm, 256, 128, 0.1)
    assert oracle.rel_l2(full, static) <= 1e-4


def test_spectral_effects_bypass_and_errors():
    """Without spectral information anywhere the extension is the nested projector
    (canBypassExtension, spectraleffectsextension.cpp:169-190); a spectral detector response with
    non-spectral volumes is an error, as in the reference; parameters keep the reference's key."""
    w = W.small(views=4)
    setup = small_setup()
    nested = RayCasterProjector()
    nested.settings().raySampling = 0.1
    ext = SpectralEffectsExtension(nested)
    ext.configure(setup)
    plain = VoxelVolume(w.volume, (1.0, 1.0, 1.0))
    got = ext.project(plain).data
    ref = RayCasterProjector()
    ref.settings().raySampling = 0.1
    ref.configure(setup)
    assert np.array_equal(got, ref.project(plain).data)
    ext.setDetectorResponse([1.0])
    with pytest.raises(RuntimeError):
        ext.project(plain)
    assert ext.parameter() == {"Sampling resolution": 10.0}
    ext.setParameter({"Sampling resolution": 2.5})
    assert ext.parameter() == {"Sampling resolution": 2.5}
    # a spectrum for the wrong number of views is rejected
    ext.setDetectorResponse(None)
    ext.setSpectralInformation(SpectralInformation([50.0], np.ones((1, 3))))
    with pytest.raises(ValueError):
        ext.project(SpectralVolume(plain, [0.2]))


def test_cpp_mirror_extensions_program(tmp_path):
    """The C++ mirrors of SpectralEffectsExtension and DynamicProjectorExtension
    (include/ctlproj.hpp over ctlproj_project_spectral / ctlproj_project_linear_dynamic), driven
    like CTL code, against the oracle."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent.parent / "ctl_amd" / "lib" / "extensions_demo"
    out = tmp_path / "ext.f32"
    res = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    raw = np.fromfile(out, np.float32)
    nv = 64 ** 3
    n_proj = 4 * 128 * 256
    vol = raw[:nv].reshape(64, 64, 64)
    spec = raw[nv:nv + n_proj].reshape(4, 1, 128, 256)
    dyn = raw[nv + n_proj:].reshape(4, 1, 128, 256)
    bone = np.ascontiguousarray(vol[:, :, ::-1] * np.float32(0.5))
    slope = vol * np.float32(0.02)
    pm = GeometryEncoder.encodeFullGeometry(small_setup())
    one = ((1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    p1, _ = oracle.project(vol, *one, pm, 256, 128, 0.1)
    p2, _ = oracle.project(bone, *one, pm, 256, 128, 0.1)
    inten = np.array([[1.0, 1.1, 0.9, 1.2], [2.0, 2.0, 1.8, 2.1], [0.5, 0.4, 0.6, 0.5]])
    mu = np.float32(0.1) * np.array([[0.4, 1.3], [0.25, 0.5], [0.18, 0.3]], np.float32)
    ref = oracle.spectral_extinction(np.stack([p1, p2]), mu, inten, [0.8, 1.0, 0.95])
    assert oracle.rel_l2(ref, spec) <= 1e-4
    ts = [0.0, 10.0, 25.0, 40.0]
    ref_d = np.concatenate([oracle.project(slope * np.float32(t) + vol, *one, pm[v:v + 1], 256, 128, 0.1)[0]
                            for v, t in enumerate(ts)])
    assert oracle.rel_l2(ref_d, dyn) <= 1e-4
