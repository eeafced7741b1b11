// Drives the C++ mirrors of SpectralEffectsExtension and DynamicProjectorExtension
// (include/ctlproj.hpp). 64^3 x 1 mm phantom (seed 1000, 10 ellipsoids), flat panel 256x128 @ 1 mm,
// C-arm 1200 mm, 4 views over 270 degrees, ray step 0.1 voxel. Writes, as raw float32 to argv[1]:
// the phantom, the spectral extinction of two materials (the phantom and 0.5 x the phantom
// mirrored in x) over three energy bins, and the projection of the linear dynamic volume
// slope = 0.02 x phantom, offset = phantom at time stamps 0, 10, 25, 40. Prints "ok".
#include "ctlproj.hpp"

#include <cstdio>
#include <vector>

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s out.f32\n", argv[0]);
        return 2;
    }
    const uint32_t n = 64;
    ctl_amd::VoxelVolume vol;
    vol.dims = { { n, n, n } };
    vol.data.resize(size_t(n) * n * n);
    ctl_amd::check(ctlproj_phantom_ellipsoids(vol.dims.data(), vol.voxelSize.data(), 10, 1000, 4.0f, 24.0f, 0.005f,
                                              0.03f, 0.02f, vol.data.data()));
    ctl_amd::VoxelVolume bone = vol, slope = vol;
    for (uint32_t z = 0; z < n; ++z)
        for (uint32_t y = 0; y < n; ++y)
            for (uint32_t x = 0; x < n; ++x) {
                const size_t i = (size_t(z) * n + y) * n + x;
                bone.data[i] = 0.5f * vol.data[(size_t(z) * n + y) * n + (n - 1 - x)];
                slope.data[i] = 0.02f * vol.data[i];
            }
    ctlproj_detector det{};
    det.pixels[0] = 256;
    det.pixels[1] = 128;
    det.pixel_mm[0] = det.pixel_mm[1] = 1.0;
    det.modules = 1;
    ctlproj_gantry g{};
    g.type = CTLPROJ_CARM;
    g.carm_span = 1200.0;
    ctlproj_trajectory t{};
    t.type = CTLPROJ_SHORT_SCAN;
    t.p[0] = 750.0;
    t.p[1] = 0.0;
    t.p[2] = 270.0 * 3.14159265358979323846 / 180.0;
    std::vector<double> pmats(4 * 12);
    ctl_amd::check(ctlproj_encode_trajectory(&g, &det, &t, 4, pmats.data(), pmats.size()));

    ctl_amd::RayCasterProjector proj;
    proj.settings().raySampling = 0.1f;
    proj.configure(pmats, { 256, 128, 1 });
    ctl_amd::ResidentVolume rv(proj.context(), vol), rb(proj.context(), bone), rs(proj.context(), slope);

    ctl_amd::SpectralEffectsExtension spectral(proj);
    spectral.setSpectralInformation(3, { 1.0, 1.1, 0.9, 1.2,   2.0, 2.0, 1.8, 2.1,   0.5, 0.4, 0.6, 0.5 });
    spectral.setDetectorResponse({ 0.8, 1.0, 0.95 });
    const ctl_amd::ProjectionData e = spectral.projectComposite({ &rv, &rb }, { { 0.4f, 0.25f, 0.18f }, { 1.3f, 0.5f, 0.3f } });

    ctl_amd::DynamicProjectorExtension dynamic(proj);
    const ctl_amd::ProjectionData d = dynamic.project(rs, rv, { 0.0, 10.0, 25.0, 40.0 });

    std::FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 3;
    std::fwrite(vol.data.data(), sizeof(float), vol.data.size(), f);
    std::fwrite(e.data.data(), sizeof(float), e.data.size(), f);
    std::fwrite(d.data.data(), sizeof(float), d.data.size(), f);
    std::fclose(f);
    std::printf("ok\n");
    return 0;
}
