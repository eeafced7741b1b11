"""MI355X-native forward projector with CTL's RayCasterProjector interface."""
from ._lib import CtlProjError, lib  # noqa: F401
from .projector import (  # noqa: F401
    DEG, AcquisitionSetup, ArealFocalSpotExtension, Batch, Server, autotune_table, AxialScanTrajectory, CarmGantry, CirclePlusLineTrajectory, Context,
    CTSystem, CylindricalDetector, DecodedGeometry, FlatPanelDetector, GeometryDecoder, GeometryEncoder,
    HelicalTrajectory,
    ProjectionData, RayCasterProjector, ResidentVolume, Settings, ShortScanTrajectory,
    TubularGantry, VolumeData, VoxelVolume, WobbleTrajectory, ellipsoid_phantom,
)
from .processing import RadonTransform3D, SamplingRange, VolumeResampler, VolumeSlicer  # noqa: F401
from .extensions import (DynamicProjectorExtension, LinearDynamicVolume, SpectralEffectsExtension,  # noqa: F401
                         SpectralInformation, SpectralVolume)
