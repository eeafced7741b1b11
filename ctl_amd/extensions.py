"""SpectralEffectsExtension and DynamicProjectorExtension on resident volumes.

Both extensions call the projection many times in the reference (once per energy bin or material,
once per view). Here the projections are written to device memory and combined there:

  SpectralEffectsExtension (spectraleffectsextension.cpp:66-156, 232-352): with a linear nested
    projector (the ray caster), one projection per material density and one kernel that sums the
    intensities of all energy bins and returns the extinction (ctlproj_spectral_extinction); with a
    non-linear one (ArealFocalSpotExtension averaging in the intensity domain), one projection per
    energy bin and material of the bin's attenuation volume (density x mean mass attenuation),
    summed per bin, as projectNonLinear does (:292-395).
  DynamicProjectorExtension (dynamicprojectorextension.cpp:39-66): for a LinearDynamicVolume
    (lineardynamicvolume.cpp:194-198, value = slope * t + offset) two full-scan projections and a
    per-view combination t_v * P_v(slope) + P_v(offset) (ctlproj_combine_views); any other dynamic
    volume (an object with setTime() and a `volume` attribute) is projected view by view at each
    view's time stamp, as the reference does.

The X-ray spectrum and the materials' attenuation (RadiationEncoder, AttenuationModel) are inputs
here: SpectralInformation carries the bins' energies and per-view intensities, SpectralVolume a
density volume with its mass attenuation coefficient in each bin.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .projector import (AcquisitionSetup, GeometryEncoder, ProjectionData, RayCasterProjector, ResidentVolume,
                        VoxelVolume, _module_pixels)


def _resident(volume) -> ResidentVolume:
    return volume if isinstance(volume, ResidentVolume) else ResidentVolume(volume)


class SpectralInformation:
    """SpectralInformation (radiationencoder.h:12-41): per energy bin its energy (keV) and the
    initial intensity of every view; the total intensity of a view defaults to the sum over bins."""

    def __init__(self, energies, intensities, total_intensity=None, bin_width: float = 0.0):
        self.energies = np.asarray(energies, np.float32).reshape(-1)
        self.intensities = np.ascontiguousarray(intensities, np.float64).reshape(len(self.energies), -1)
        tot = self.intensities.sum(axis=0) if total_intensity is None else total_intensity
        self.total_intensity = np.ascontiguousarray(tot, np.float64).reshape(-1)
        self.bin_width = float(bin_width)
        if self.total_intensity.shape[0] != self.intensities.shape[1]:
            raise ValueError("one total intensity per view")

    def nbEnergyBins(self) -> int:
        return len(self.energies)

    def binWidth(self) -> float:
        return self.bin_width

    def nbViews(self) -> int:
        return self.intensities.shape[1]


class SpectralVolume:
    """A material density volume (g/cm^3) with its mean mass attenuation coefficient (cm^2/g) in
    each energy bin (SpectralVolumeData::densityVolume / meanMassAttenuationCoeff,
    spectralvolumedata.cpp)."""

    def __init__(self, density, mass_attenuation, material_name: str = ""):
        self.density = density
        self.mass_attenuation = np.asarray(mass_attenuation, np.float32).reshape(-1)
        self.material_name = material_name


class SpectralEffectsExtension:
    """SpectralEffectsExtension: per pixel ln(I0 / sum_E i0(E) r(E) exp(-sum_k m_k(E) F(rho_k)))
    with a linear nested projector F, ln(I0 / sum_E i0(E) r(E) exp(-sum_k F(m_k(E) rho_k))) with a
    non-linear one (spectraleffectsextension.cpp:60-98)."""

    def __init__(self, projector: RayCasterProjector | None = None,
                 spectral_info: SpectralInformation | None = None, energyBinWidth: float = 10.0):
        self._nested = projector or RayCasterProjector()
        self._info = spectral_info
        self._delta_e = float(energyBinWidth)
        self._response = None

    def use(self, projector: RayCasterProjector) -> None:
        self._nested = projector

    def setSpectralInformation(self, info: SpectralInformation) -> None:
        self._info = info

    def setSpectralSamplingResolution(self, energyBinWidth: float) -> None:
        self._delta_e = float(energyBinWidth)

    def setDetectorResponse(self, response) -> None:
        """Fraction of the radiation detected in each bin (AbstractDetectorResponse::valueAt at the
        bins' energies); None: no response model."""
        self._response = None if response is None else np.asarray(response, np.float32).reshape(-1)

    def parameter(self) -> dict:
        return {"Sampling resolution": self._delta_e}

    def setParameter(self, parameter: dict) -> None:
        self.setSpectralSamplingResolution(float(parameter.get("Sampling resolution", 0.0)))

    def isLinear(self) -> bool:
        return False

    def configure(self, setup) -> None:
        self._nested.configure(setup)

    def project(self, volume) -> ProjectionData:
        return self.projectComposite([volume])

    def projectComposite(self, volumes) -> ProjectionData:
        spectral = [v for v in volumes if isinstance(v, SpectralVolume)]
        if not spectral:  # canBypassExtension: no spectral information anywhere
            if self._response is not None:
                raise RuntimeError("SpectralEffectsExtension: cannot simulate a spectral detector response "
                                   "for volume data without spectral information")
            return self._nested.projectComposite([_resident(v) for v in volumes])
        if len(spectral) != len(volumes):
            raise RuntimeError("SpectralEffectsExtension: mixing volumes with and without spectral "
                               "information is not supported")
        if self._info is None:
            raise RuntimeError("SpectralEffectsExtension: no spectral information set")
        if not self._nested.isLinear():
            return self._project_nonlinear(spectral)
        import torch  # device buffers only
        info = self._info
        V, M, R, C_ = self._nested._dims
        if info.nbViews() != V:
            raise ValueError(f"spectral information holds {info.nbViews()} views, the projector {V}")
        n = M * R * C_
        B = info.nbEnergyBins()
        for v in spectral:
            if len(v.mass_attenuation) != B:
                raise ValueError("one mass attenuation coefficient per energy bin and material")
        projs = torch.empty(len(spectral) * V * n, dtype=torch.float32, device="cuda")
        out = torch.empty(V * n, dtype=torch.float32, device="cuda")
        keep = []
        for k, v in enumerate(spectral):
            rv = _resident(v.density)
            keep.append(rv)
            self._nested.project_device(rv, projs.data_ptr() + k * V * n * 4, V * n)
        self._nested.synchronize()
        mu = np.ascontiguousarray(np.stack([v.mass_attenuation * np.float32(0.1) for v in spectral], axis=1),
                                  np.float32)  # [B][materials], 1/cm -> 1/mm (spectraleffectsextension.cpp:286)
        resp = None if self._response is None else np.ascontiguousarray(self._response, np.float32)
        if resp is not None and len(resp) != B:
            raise ValueError("one detector response value per energy bin")
        check(lib().ctlproj_spectral_extinction(
            self._nested.ctx.h, C.c_void_p(projs.data_ptr()), len(spectral), V, n, _lib.fptr(mu),
            _lib.dptr(info.intensities), _lib.fptr(resp) if resp is not None else None, B,
            _lib.dptr(info.total_intensity), C.c_void_p(out.data_ptr())))
        return ProjectionData(out.cpu().numpy().reshape(V, M, R, C_))


    def _project_nonlinear(self, spectral) -> ProjectionData:
        """projectNonLinear / singleBinIntensityNonLinear (spectraleffectsextension.cpp:292-395):
        for each energy bin with photons, the nested projector projects every material's
        attenuation volume (density x mean mass attenuation coefficient of the bin, 1/cm -> 1/mm),
        the projections add up, turn into intensity with the bin's per-view intensities and the
        detector response; the bins add up and the sum turns into extinction against I0."""
        info = self._info
        B, V = info.nbEnergyBins(), info.nbViews()
        for v in spectral:
            if len(v.mass_attenuation) != B:
                raise ValueError("one mass attenuation coefficient per energy bin and material")
            if not isinstance(v.density, VoxelVolume):
                raise ValueError("the non-linear spectral path needs host densities (VoxelVolume)")
        if self._response is not None and len(self._response) != B:
            raise ValueError("one detector response value per energy bin")
        total = None
        for b in range(B):
            if abs(float(info.intensities[b].sum())) <= 1e-12:  # qFuzzyIsNull: no photons in this bin
                continue
            q = None
            for v in spectral:
                mu = VoxelVolume(v.density.data * np.float32(v.mass_attenuation[b] * np.float32(0.1)),
                                 v.density.voxel_size, v.density.offset)
                pb = self._nested.project(mu).data.astype(np.float64)
                q = pb if q is None else q + pb
            if q.shape[0] != V:
                raise ValueError(f"spectral information holds {V} views, the projector {q.shape[0]}")
            inten = info.intensities[b].reshape((V,) + (1,) * (q.ndim - 1)) * np.exp(-q)
            if self._response is not None:
                inten = inten * float(self._response[b])
            total = inten if total is None else total + inten
        if total is None:
            raise RuntimeError("SpectralEffectsExtension: no energy bin holds photons")
        i0 = info.total_intensity.reshape((V,) + (1,) * (total.ndim - 1))
        return ProjectionData(np.log(i0 / total).astype(np.float32))


class LinearDynamicVolume:
    """LinearDynamicVolume (lineardynamicvolume.cpp:5-198): every voxel is slope * t + offset at the
    time t set with setTime (FP32 arithmetic)."""

    def __init__(self, slope: VoxelVolume, offset: VoxelVolume, voxel_size=None):
        if slope.data.shape != offset.data.shape:
            raise ValueError("LinearDynamicVolume: offset and slope volumes have different sizes")
        vs = slope.voxel_size if voxel_size is None else tuple(float(v) for v in voxel_size)
        self.slope = VoxelVolume(slope.data, vs, slope.offset)
        self.offset = VoxelVolume(offset.data, vs, slope.offset)
        self._t = 0.0

    def setTime(self, t: float) -> None:
        self._t = float(t)

    def time(self) -> float:
        return self._t

    @property
    def volume(self) -> VoxelVolume:
        return self.at(self._t)

    def at(self, t: float) -> VoxelVolume:
        data = self.slope.data * np.float32(t) + self.offset.data
        return VoxelVolume(data, self.slope.voxel_size, self.slope.offset)


class DynamicProjectorExtension:
    """DynamicProjectorExtension (dynamicprojectorextension.cpp:39-66): view v sees the volume at
    the view's time stamp (AcquisitionSetup.timeStamp)."""

    def __init__(self, projector: RayCasterProjector | None = None):
        self._nested = projector or RayCasterProjector()
        self._setup = None

    def use(self, projector: RayCasterProjector) -> None:
        self._nested = projector

    def configure(self, setup: AcquisitionSetup) -> None:
        self._setup = setup

    def isLinear(self) -> bool:
        return self._nested.isLinear()

    def project(self, volume) -> ProjectionData:
        setup = self._setup
        if setup is None:
            raise _lib.CtlProjError(_lib.ERR_NOT_CONFIGURED, "project() called before configure()")
        if isinstance(volume, LinearDynamicVolume):
            return self._project_linear(volume)
        if not hasattr(volume, "setTime"):
            self._nested.configure(setup)
            return self._nested.project(volume)
        pm = GeometryEncoder.encodeFullGeometry(setup)
        ch, rows = _module_pixels(setup)
        views = []
        for v in range(setup.nbViews()):
            volume.setTime(setup.timeStamp(v))
            self._nested.configure((pm[v:v + 1], ch, rows))
            views.append(self._nested.project(volume.volume).data[0])
        return ProjectionData(np.stack(views) if views else np.zeros((0, pm.shape[1], rows, ch), np.float32))

    def _project_linear(self, volume: LinearDynamicVolume) -> ProjectionData:
        import torch  # device buffers only
        setup = self._setup
        self._nested.configure(setup)
        V, M, R, C_ = self._nested._dims
        n = M * R * C_
        a = torch.empty(V * n, dtype=torch.float32, device="cuda")
        b = torch.empty(V * n, dtype=torch.float32, device="cuda")
        rs, ro = ResidentVolume(volume.slope), ResidentVolume(volume.offset)
        self._nested.project_device(rs, a.data_ptr(), V * n)
        self._nested.project_device(ro, b.data_ptr(), V * n)
        self._nested.synchronize()
        ca = np.array([setup.timeStamp(v) for v in range(V)], np.float32)
        cb = np.ones(V, np.float32)
        check(lib().ctlproj_combine_views(self._nested.ctx.h, C.c_void_p(a.data_ptr()), _lib.fptr(ca),
                                          C.c_void_p(b.data_ptr()), _lib.fptr(cb), V, n, C.c_void_p(a.data_ptr())))
        return ProjectionData(a.cpu().numpy().reshape(V, M, R, C_))
