// FP64 acquisition geometry: gantries, detectors, trajectories and the projection-matrix encoder.
//
// The conventions follow CTL's components (carmgantry.cpp:48-71, tubulargantry.cpp:131-207,
// abstractgantry.h:159-245, flatpaneldetector.cpp:47-50, cylindricaldetector.cpp:127-205) and
// GeometryEncoder (geometryencoder.cpp:83-135, 289-360), so that encoded matrices agree with the
// reference's to rounding.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace ctlgeo {

struct Vec3 {
    double v[3];
    double operator[](int i) const { return v[i]; }
    double& operator[](int i) { return v[i]; }
};

struct Mat3 {
    double m[9]; // row-major
    double operator()(int r, int c) const { return m[3 * r + c]; }
    double& operator()(int r, int c) { return m[3 * r + c]; }
    static Mat3 eye();
    Mat3 transposed() const;
};

Mat3 operator*(const Mat3& a, const Mat3& b);
Vec3 operator*(const Mat3& a, const Vec3& x);
Vec3 operator+(const Vec3& a, const Vec3& b);
Vec3 operator-(const Vec3& a, const Vec3& b);

enum Axis { X = 0, Y = 1, Z = 2 };

// Rotation about a coordinate axis (mat::rotationMatrix, matrix_utils.tpp:21-41).
Mat3 rotation(double angle, Axis axis);

struct Location { // mat::Location (matrix_utils.h:23-33)
    Vec3 position{ { 0.0, 0.0, 0.0 } };
    Mat3 rotation = Mat3::eye();
};

struct Detector {
    uint32_t channels = 1, rows = 1;     // pixels per module (width, height)
    double pixel_w = 1.0, pixel_h = 1.0; // mm
    uint32_t modules = 1;
    double angulation = 0.0; // rad between adjacent modules; 0 with one module = flat panel
    double spacing = 0.0;    // mm between adjacent modules
    double skew = 0.0;       // AbstractDetector::skewCoefficient
    bool cylindrical = false;

    double moduleWidth() const { return channels * pixel_w; } // cylindricaldetector.cpp:280-283
    double curvatureRadius() const;                          // cylindricaldetector.cpp:127-136
    double fanAngle() const;                                 // cylindricaldetector.cpp:270-274
    std::vector<Location> moduleLocations() const;
};

struct Gantry {
    enum Type { CArm = 0, Tubular = 1 } type = CArm;
    // C-arm (carmgantry.cpp): span and the location of the source
    double carm_span = 1000.0;
    Location carm_location;
    // tubular (tubulargantry.cpp): distances, rotation, pitch position and tilt
    double sdd = 1000.0, sid = 500.0, rotation_angle = 0.0, pitch_position = 0.0, tilt = 0.0;

    Vec3 sourcePosition() const;
    Mat3 detectorRotation() const; // world -> detector CT coordinates (passive form)
    Vec3 detectorPosition() const;
};

// 3x4 FP64 projection matrix, row-major.
struct PMat {
    double p[12];
};

// GeometryEncoder::encodeSingleViewGeometry (geometryencoder.cpp:104-135): one P per module.
void encodeSingleView(const Gantry& g, const Detector& d, const Vec3& focal_spot, PMat* out);

// Trajectories (prepare protocols of trajectories.cpp). Each sets the gantry for view `v` of `n`.
struct Trajectory {
    enum Type {
        ShortScan = 0,     // (source_to_iso, start_angle, angle_span)  trajectories.cpp:177-205
        Wobble = 1,        // (angle_span, source_to_iso, start_angle, wobble_angle, wobble_freq) :71-96
        CirclePlusLine = 2,// (angle_span, source_to_iso, line_length, fraction_line, start_angle) :125-175
        Helical = 3,       // (angle_increment, pitch_increment, start_pitch, start_angle) :33-50
        Axial = 4          // (start_angle) :239-256
    } type = ShortScan;
    double p[6] = { 0, 0, 0, 0, 0, 0 };
};

// Returns false if the trajectory does not apply to the gantry type.
bool prepareView(const Trajectory& t, uint32_t view, uint32_t n_views, const Detector& d, Gantry& g);

// Normalisation of a P-matrix (ProjectionMatrix::normalized, projectionmatrix.cpp:143-150).
void normalize(PMat& P);

// Encoding of generic components (GeometryEncoder::encodeSingleViewGeometry for a GenericGantry /
// GenericDetector): source position and per-module world locations (position of the module
// centre, rotation world -> module), pixels per module and pixel size.
void encodeModules(const Vec3& src, const Location* modules, size_t n, uint32_t channels, uint32_t rows,
                   double pw, double ph, double skew, PMat* out);

// GeometryDecoder (geometrydecoder.cpp): physical geometry of one view from its P-matrices.
enum DecodeReference {
    DecodePixelSize = 0,              // pixel width and height given
    DecodePixelWidth = 1,             // PhysicalDimension::PixelWidth
    DecodePixelHeight = 2,            // PhysicalDimension::PixelHeight
    DecodeSourceDetectorDistance = 3, // PhysicalDimension::SourceDetectorDistance
};
struct DecodedView {
    Vec3 source;
    Mat3 source_rotation; // computeSourceRotation (columns r1, r2, source->detector direction)
    double pixel_w = 0, pixel_h = 0, skew = 0;
    std::vector<Location> modules; // module centres and rotations (world -> module)
};
// reference DecodePixelSize: a, b = pixel width, height; otherwise a = the reference length in mm.
bool decodeView(const PMat* P, uint32_t n_modules, uint32_t channels, uint32_t rows, int reference,
                double a, double b, DecodedView* out);

} // namespace ctlgeo
