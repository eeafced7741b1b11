// FP64 acquisition geometry and projection-matrix encoder (host side of configure()).
#include "geometry.h"

#include <cmath>
#include <limits>

namespace ctlgeo {

namespace {
constexpr double kPi = 3.14159265358979323846;
constexpr double kPi2 = 1.57079632679489661923;
} // namespace

Mat3 Mat3::eye() { return Mat3{ { 1, 0, 0, 0, 1, 0, 0, 0, 1 } }; }

Mat3 Mat3::transposed() const
{
    Mat3 t;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t(r, c) = (*this)(c, r);
    return t;
}

Mat3 operator*(const Mat3& a, const Mat3& b)
{
    Mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a(i, k) * b(k, j);
            r(i, j) = s;
        }
    return r;
}

Vec3 operator*(const Mat3& a, const Vec3& x)
{
    Vec3 r;
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += a(i, k) * x[k];
        r[i] = s;
    }
    return r;
}

Vec3 operator+(const Vec3& a, const Vec3& b) { return Vec3{ { a[0] + b[0], a[1] + b[1], a[2] + b[2] } }; }
Vec3 operator-(const Vec3& a, const Vec3& b) { return Vec3{ { a[0] - b[0], a[1] - b[1], a[2] - b[2] } }; }

Mat3 rotation(double angle, Axis axis)
{
    const double s = std::sin(angle), c = std::cos(angle);
    switch (axis) {
    case X: return Mat3{ { 1, 0, 0, 0, c, -s, 0, s, c } };
    case Y: return Mat3{ { c, 0, s, 0, 1, 0, -s, 0, c } };
    default: return Mat3{ { c, -s, 0, s, c, 0, 0, 0, 1 } };
    }
}

// ---------------------------------------------------------------------------------------------
// detectors

double Detector::curvatureRadius() const
{
    if (std::fabs(angulation) <= 1e-12) return std::numeric_limits<double>::max();
    const double d = spacing + moduleWidth() * std::cos(angulation / 2.0);
    return d / std::sqrt(2.0 * (1.0 - std::cos(angulation)));
}

double Detector::fanAngle() const
{
    return double(modules - 1) * angulation + 2.0 * std::atan(0.5 * moduleWidth() / curvatureRadius());
}

std::vector<Location> Detector::moduleLocations() const
{
    if (!cylindrical) return std::vector<Location>(1); // flatpaneldetector.cpp:47-50
    // cylindricaldetector.cpp:154-205: rotate the point (0,0,r) about y by each module's
    // angulation, then translate by -r along z; the rotation is stored in passive form.
    std::vector<Location> locs(modules);
    const double r = curvatureRadius();
    const Vec3 pt{ { 0.0, 0.0, r } };
    for (uint32_t m = 0; m < modules; ++m) {
        const double ang = (double(m) - double(modules) * 0.5 + 0.5) * angulation;
        const Mat3 rot = rotation(ang, Y);
        Vec3 p = rot * pt;
        p[2] -= r;
        locs[m].position = p;
        locs[m].rotation = rot.transposed();
    }
    return locs;
}

// ---------------------------------------------------------------------------------------------
// gantries (no displacements: the displacement terms of abstractgantry.h:159-245 are identity)

namespace {
Mat3 tubularRotation(const Gantry& g) // tubulargantry.cpp:200-207
{
    return rotation(g.tilt, X) * rotation(g.rotation_angle, Z);
}
Mat3 tubularSourceRotation(const Gantry& g) // tubulargantry.cpp:146-150
{
    return tubularRotation(g) * rotation(kPi2, Z) * rotation(-kPi2, X);
}
Mat3 sourceRotation(const Gantry& g)
{
    return g.type == Gantry::CArm ? g.carm_location.rotation : tubularSourceRotation(g);
}
} // namespace

Vec3 Gantry::sourcePosition() const
{
    if (type == CArm) return carm_location.position; // carmgantry.cpp:34
    return tubularRotation(*this) * Vec3{ { sid, 0.0, -pitch_position } }; // tubulargantry.cpp:131-135
}

Mat3 Gantry::detectorRotation() const
{
    if (type == CArm) return carm_location.rotation.transposed(); // carmgantry.cpp:68-71
    return tubularSourceRotation(*this).transposed();              // tubulargantry.cpp:187-193
}

Vec3 Gantry::detectorPosition() const
{
    if (type == CArm) // carmgantry.cpp:48-53: position - R * (0, 0, -span)
        return carm_location.position - carm_location.rotation * Vec3{ { 0.0, 0.0, -carm_span } };
    return tubularRotation(*this) * Vec3{ { -(sdd - sid), 0.0, -pitch_position } }; // :166-172
}

// ---------------------------------------------------------------------------------------------
// encoder

namespace {
double det3(const Mat3& m)
{
    return m(0, 0) * (m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2)) -
           m(0, 1) * (m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2)) +
           m(0, 2) * (m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1));
}

// GeometryEncoder::intrinsicParameterMatrix (geometryencoder.cpp:311-360)
Mat3 intrinsic(const Vec3& dev, uint32_t nx, uint32_t ny, double pw, double ph, double skew)
{
    const double f = std::fabs(dev[2]);
    const double fx = f / pw, fy = f / ph;
    double ppx = dev[0] / pw;
    const double ppy = dev[1] / ph;
    ppx += ppy * skew / fy;
    const double mx = 0.5 * double(nx - 1), my = 0.5 * double(ny - 1);
    return Mat3{ { fx, skew, mx - ppx, 0.0, fy, my - ppy, 0.0, 0.0, 1.0 } };
}
} // namespace

void normalize(PMat& P)
{
    Mat3 M{ { P.p[0], P.p[1], P.p[2], P.p[4], P.p[5], P.p[6], P.p[8], P.p[9], P.p[10] } };
    double denom = std::sqrt(P.p[8] * P.p[8] + P.p[9] * P.p[9] + P.p[10] * P.p[10]);
    denom = std::copysign(denom, det3(M));
    for (double& x : P.p) x /= denom;
}

void encodeSingleView(const Gantry& g, const Detector& d, const Vec3& focal_spot, PMat* out)
{
    const Vec3 src = g.sourcePosition() + sourceRotation(g) * focal_spot; // geometryencoder.cpp:289-293
    const Vec3 det_pos = g.detectorPosition();
    const Mat3 det_rot = g.detectorRotation();
    const Mat3 det_rot_t = det_rot.transposed();
    const auto locs = d.moduleLocations();
    for (size_t m = 0; m < locs.size(); ++m) {
        const Vec3 mod_pos = det_pos + det_rot_t * locs[m].position;
        const Mat3 total_rot = locs[m].rotation * det_rot;
        const Vec3 dev = total_rot * (mod_pos - src);
        const Mat3 K = intrinsic(dev, d.channels, d.rows, d.pixel_w, d.pixel_h, d.skew);
        // ProjectionMatrix::compose(K, R, c) = [K R | -K R c] (projectionmatrix.cpp:43-48)
        const Mat3 KR = K * total_rot;
        const Vec3 t = KR * src;
        PMat P;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) P.p[4 * r + c] = KR(r, c);
            P.p[4 * r + 3] = -t[r];
        }
        normalize(P);
        out[m] = P;
    }
}

// ---------------------------------------------------------------------------------------------
// trajectories

namespace {
// fixed rotation shared by the C-arm protocols: Rz(pi/2 + start) * Rx(-pi/2)
Mat3 carmFixed(double start_angle) { return rotation(kPi2 + start_angle, Z) * rotation(-kPi2, X); }

double shortScanFan(const Detector& d, const Gantry& g) // trajectories.cpp:211-236
{
    double width;
    if (d.cylindrical) width = 2.0 * d.curvatureRadius() * std::sin(d.fanAngle() / 2.0);
    else width = d.moduleWidth();
    return 2.0 * std::tan(0.5 * width / g.carm_span);
}
} // namespace

bool prepareView(const Trajectory& t, uint32_t v, uint32_t n, const Detector& d, Gantry& g)
{
    const double* p = t.p;
    switch (t.type) {
    case Trajectory::ShortScan: { // trajectories.cpp:177-205
        if (g.type != Gantry::CArm) return false;
        const double span = p[2] >= 0.0 ? p[2] : kPi + shortScanFan(d, g);
        const double inc = n > 1 ? span / double(n - 1) : 0.0;
        const Mat3 rot = rotation(v * inc, Z) * carmFixed(p[1]);
        g.carm_location.rotation = rot;
        g.carm_location.position = rot * Vec3{ { 0.0, 0.0, -p[0] } };
        return true;
    }
    case Trajectory::Wobble: { // trajectories.cpp:71-96
        if (g.type != Gantry::CArm) return false;
        const double inc = n > 1 ? p[0] / double(n - 1) : 0.0;
        const double phase = std::sin(double(v) / double(n) * 2.0 * kPi * p[4]);
        const Mat3 rot = rotation(v * inc, Z) * carmFixed(p[2]) * rotation(phase * p[3], X);
        g.carm_location.rotation = rot;
        g.carm_location.position = rot * Vec3{ { 0.0, 0.0, -p[1] } };
        return true;
    }
    case Trajectory::CirclePlusLine: { // trajectories.cpp:125-175
        if (g.type != Gantry::CArm) return false;
        const uint32_t n_line = uint32_t(std::floor(n * p[3]));
        const uint32_t n_circle = n - n_line;
        const Vec3 init{ { 0.0, 0.0, -p[1] } };
        const Mat3 fixed = carmFixed(p[4]);
        if (v < n_circle) {
            const double inc = n_circle > 1 ? p[0] / double(n_circle - 1) : 0.0;
            const Mat3 rot = rotation(v * inc, Z) * fixed;
            g.carm_location.rotation = rot;
            g.carm_location.position = rot * init;
        } else {
            const uint32_t lv = v - n_circle;
            const double inc = n_line > 1 ? p[2] / double(n_line - 1) : 0.0;
            const Mat3 rot = rotation(p[0] / 2.0, Z) * fixed;
            Vec3 pos = rot * init;
            pos[2] += (lv - n_line / 2.0) * inc;
            g.carm_location.rotation = rot;
            g.carm_location.position = pos;
        }
        return true;
    }
    case Trajectory::Helical: // trajectories.cpp:33-50
        if (g.type != Gantry::Tubular) return false;
        g.pitch_position = v * p[1] + p[2];
        g.rotation_angle = v * p[0] + p[3];
        return true;
    case Trajectory::Axial: { // trajectories.cpp:239-256 (the increment is a float there)
        if (g.type != Gantry::Tubular) return false;
        const float inc = n > 0 ? float(2.0 * kPi / n) : 0.0f;
        g.rotation_angle = double(float(v) * inc) + p[0];
        return true;
    }
    }
    return false;
}

} // namespace ctlgeo
