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

void encodeModules(const Vec3& src, const Location* modules, size_t n, uint32_t channels, uint32_t rows,
                   double pw, double ph, double skew, PMat* out)
{
    for (size_t m = 0; m < n; ++m) {
        const Vec3 dev = modules[m].rotation * (modules[m].position - src);
        const Mat3 K = intrinsic(dev, channels, rows, pw, ph, skew);
        // ProjectionMatrix::compose(K, R, c) = [K R | -K R c] (projectionmatrix.cpp:43-48)
        const Mat3 KR = K * modules[m].rotation;
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

namespace {
// M = R Q with R upper triangular, positive diagonal, Q orthogonal: Gram-Schmidt on the rows of M
// from the last one up. This is the unique RQ decomposition (mat::RQdecomposition with `unique`,
// matrix_algorithm.cpp:67-87); it needs a regular M.
bool rq(const Mat3& M, Mat3& R, Mat3& Q)
{
    R = Mat3{ { 0, 0, 0, 0, 0, 0, 0, 0, 0 } };
    for (int i = 2; i >= 0; --i) {
        double v[3] = { M(i, 0), M(i, 1), M(i, 2) };
        for (int j = i + 1; j < 3; ++j) {
            const double rij = v[0] * Q(j, 0) + v[1] * Q(j, 1) + v[2] * Q(j, 2);
            R(i, j) = rij;
            for (int k = 0; k < 3; ++k) v[k] -= rij * Q(j, k);
        }
        const double nrm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (!(nrm > 0.0)) return false;
        R(i, i) = nrm;
        for (int k = 0; k < 3; ++k) Q(i, k) = v[k] / nrm;
    }
    return true;
}

Vec3 cross(const Vec3& a, const Vec3& b)
{
    return Vec3{ { a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0] } };
}
} // namespace

bool decodeView(const PMat* P, uint32_t n, uint32_t channels, uint32_t rows, int reference, double a, double b,
                DecodedView* out)
{
    // source position from module 0 (ProjectionMatrix::sourcePosition, projectionmatrix.cpp:264-274)
    const double* p = P[0].p;
    auto det = [&](int c0, int c1, int c2) {
        const Mat3 m{ { p[c0], p[c1], p[c2], p[4 + c0], p[4 + c1], p[4 + c2], p[8 + c0], p[8 + c1], p[8 + c2] } };
        return det3(m);
    };
    const double dm = det(0, 1, 2);
    if (dm == 0.0) return false;
    out->source = Vec3{ { det(1, 2, 3) / -dm, -det(0, 2, 3) / -dm, det(0, 1, 3) / -dm } };
    // intrinsics of module 0 (intrinsicMatK: R of the unique RQ scaled to R22 = 1): focal lengths
    // in pixels and skew (geometrydecoder.cpp computePixelSize)
    Mat3 R, Q;
    const Mat3 M0{ { p[0], p[1], p[2], p[4], p[5], p[6], p[8], p[9], p[10] } };
    if (!rq(M0, R, Q)) return false;
    const double fx = R(0, 0) / R(2, 2), fy = R(1, 1) / R(2, 2);
    out->skew = R(0, 1) / R(2, 2);
    switch (reference) {
    case DecodePixelSize: out->pixel_w = a; out->pixel_h = b; break;
    case DecodePixelWidth: out->pixel_w = a; out->pixel_h = a * fx / fy; break;
    case DecodePixelHeight: out->pixel_w = a * fy / fx; out->pixel_h = a; break;
    case DecodeSourceDetectorDistance: out->pixel_w = a / fx; out->pixel_h = a / fy; break;
    default: return false;
    }
    // module locations (computeModuleLocations): centre pixel direction M^-1 [x, y, 1] scaled by
    // R00 (directionSourceToPixel, NormalizeByX) and the pixel width, from the source; rotation Q
    out->modules.resize(n);
    const double cx = 0.5 * double(channels - 1), cy = 0.5 * double(rows - 1);
    Vec3 centroid{ { 0, 0, 0 } };
    for (uint32_t m = 0; m < n; ++m) {
        const double* q = P[m].p;
        const Mat3 Mm{ { q[0], q[1], q[2], q[4], q[5], q[6], q[8], q[9], q[10] } };
        if (!rq(Mm, R, Q)) return false;
        const double sgn = det3(Mm) < 0.0 ? -1.0 : 1.0; // sign(det M) M^-1 [x, y, 1]
        const double dz = 1.0 / R(2, 2);
        const double dy = (cy - dz * R(1, 2)) / R(1, 1);
        const double dx = (cx - dy * R(0, 1) - dz * R(0, 2)) / R(0, 0);
        const Vec3 d = Q.transposed() * Vec3{ { dx, dy, dz } };
        const double s = sgn * R(0, 0) * out->pixel_w;
        out->modules[m].position = out->source + Vec3{ { d[0] * s, d[1] * s, d[2] * s } };
        out->modules[m].rotation = Q;
        centroid = centroid + out->modules[m].position;
    }
    // source rotation (computeSourceRotation): columns r1, r2, s with s the unit vector from the
    // source to the module centroid, r1 = s x e (e the axis least aligned with s), r2 = s x r1
    for (int k = 0; k < 3; ++k) centroid[k] /= double(n);
    Vec3 s = centroid - out->source;
    const double ns = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    if (!(ns > 0.0)) return false;
    for (int k = 0; k < 3; ++k) s[k] /= ns;
    int axis = std::fabs(s[0]) < std::fabs(s[1]) ? 0 : 1;
    axis = std::fabs(s[axis]) < std::fabs(s[2]) ? axis : 2;
    Vec3 e{ { 0, 0, 0 } };
    e[axis] = 1.0;
    const Vec3 r1 = cross(s, e), r2 = cross(s, r1);
    out->source_rotation = Mat3{ { r1[0], r2[0], s[0], r1[1], r2[1], s[1], r1[2], r2[2], s[2] } };
    return true;
}

void encodeSingleView(const Gantry& g, const Detector& d, const Vec3& focal_spot, PMat* out)
{
    const Vec3 src = g.sourcePosition() + sourceRotation(g) * focal_spot; // geometryencoder.cpp:289-293
    const Vec3 det_pos = g.detectorPosition();
    const Mat3 det_rot = g.detectorRotation();
    const Mat3 det_rot_t = det_rot.transposed();
    auto locs = d.moduleLocations();
    for (auto& l : locs) { // module locations in world coordinates
        l.position = det_pos + det_rot_t * l.position;
        l.rotation = l.rotation * det_rot;
    }
    encodeModules(src, locs.data(), locs.size(), d.channels, d.rows, d.pixel_w, d.pixel_h, d.skew, out);
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
