// Drives the C++ mirrors of VolumeSlicer, VolumeResampler and RadonTransform3D (include/ctlproj.hpp)
// the way CTL code drives the OCL classes. Writes, as raw float32 to argv[1]: the phantom
// (40 x 36 x 32, voxel 0.9 x 1.1 x 0.8 mm, offset (1.5, -2, 0.7)), one slice at azimuth 0.4,
// polar 1.1, distance 2.5 (default size and resolution), 7 resampled points, and a 3 x 2 x 4
// sampleTransform; prints "slice=<w>x<h> radon=<side>".
#include "ctlproj.hpp"

#include <cstdio>
#include <vector>

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s out.f32\n", argv[0]);
        return 2;
    }
    ctl_amd::VoxelVolume vol;
    vol.dims = { { 40, 36, 32 } };
    vol.voxelSize = { { 0.9f, 1.1f, 0.8f } };
    vol.offset = { { 1.5f, -2.0f, 0.7f } };
    vol.data.resize(size_t(40) * 36 * 32);
    ctl_amd::check(ctlproj_phantom_ellipsoids(vol.dims.data(), vol.voxelSize.data(), 8, 77, 2.0f, 12.0f, 0.2f,
                                              1.0f, 0.1f, vol.data.data()));
    ctl_amd::Context ctx;

    ctl_amd::VolumeSlicer slicer(ctx, vol);
    const ctl_amd::Chunk2D s = slicer.slice(0.4, 1.1, 2.5);

    ctl_amd::VolumeResampler resampler(ctx, vol);
    const std::vector<float> pts = { 0.0f, 0.0f, 0.0f,   5.2f, -3.1f, 4.0f,   -17.0f, 16.0f, -11.0f,
                                     1.5f, -2.0f, 0.7f,  20.0f, 0.0f, 0.0f,   -9.3f, 7.7f, 2.2f,
                                     3.3f, -18.0f, 11.9f };
    const std::vector<float> r = resampler.sample(pts);

    ctl_amd::RadonTransform3D radon(ctx, vol);
    const std::vector<float> azi = { 0.0f, 1.0f, 2.0f }, pol = { 0.5f, 1.5f }, dist = { -6.0f, -1.0f, 0.5f, 7.0f };
    const std::vector<float> t = radon.sampleTransform(azi, pol, dist);

    std::FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 3;
    std::fwrite(vol.data.data(), sizeof(float), vol.data.size(), f);
    std::fwrite(s.data.data(), sizeof(float), s.data.size(), f);
    std::fwrite(r.data(), sizeof(float), r.size(), f);
    std::fwrite(t.data(), sizeof(float), t.size(), f);
    std::fclose(f);
    std::printf("slice=%ux%u radon=%u\n", s.width, s.height, radon.sliceSide());
    return 0;
}
