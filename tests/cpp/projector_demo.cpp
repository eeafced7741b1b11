// Drives the header-only C++ mirror (include/ctlproj.hpp) the way CTL code drives
// RayCasterProjector: configure from encoded geometry, project a volume, count progress signals.
// Writes the projections as raw float32 to argv[1]; prints "views=<n> progress=<n>".
#include "ctlproj.hpp"

#include <cstdio>
#include <vector>

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s out.f32\n", argv[0]);
        return 2;
    }
    // 64^3 x 1 mm phantom (seed 1000, 10 ellipsoids), flat panel 256x128 @ 1 mm, C-arm 1200 mm,
    // 36 views over 360 degrees, ray step 0.1 voxel
    ctl_amd::VoxelVolume vol;
    vol.dims = { { 64, 64, 64 } };
    vol.data.resize(64 * 64 * 64);
    ctl_amd::check(ctlproj_phantom_ellipsoids(vol.dims.data(), vol.voxelSize.data(), 10, 1000, 4.0f,
                                              24.0f, 0.005f, 0.03f, 0.02f, vol.data.data()));
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
    t.p[2] = 350.0 * 3.14159265358979323846 / 180.0;
    std::vector<double> pmats(36 * 12);
    ctl_amd::check(ctlproj_encode_trajectory(&g, &det, &t, 36, pmats.data(), pmats.size()));

    ctl_amd::RayCasterProjector proj;
    proj.settings().raySampling = 0.1f;
    int progress = 0;
    proj.projectionFinished = [&](int) { ++progress; };
    proj.configure(pmats, { 256, 128, 1 });
    ctl_amd::ProjectionData d = proj.project(vol);

    // a volume without data gives 0 views, as in the reference
    ctl_amd::VoxelVolume empty;
    const bool empty_ok = proj.project(empty).nbViews == 0;

    std::FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 3;
    std::fwrite(d.data.data(), sizeof(float), d.data.size(), f);
    std::fclose(f);
    std::printf("views=%u progress=%d empty_ok=%d\n", d.nbViews, progress, int(empty_ok));
    return 0;
}
