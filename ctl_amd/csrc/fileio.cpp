// DEN and NRRD files for volumes, projections and P-matrices (io/den, io/nrrd of the reference).
//
// DEN (io/den/den_header.tpp, denfileio.h): a 6-byte header of three little-endian uint16 values
// rows, cols, count, then count matrices of rows x cols elements, cols fastest. The element type
// is not stored; it follows from the data size (1, 2, 4 or 8 bytes per element: uchar, ushort,
// float, double). Volumes are cols = X, rows = Y, count = Z; projections cols = channels,
// rows = detector rows, count = modules x views; P-matrices 3 x 4 per (view, module).
//
// NRRD (io/nrrd/nrrdfileio.tpp): the text header "NRRD000<v>", "field: value" lines, "key:=value"
// pairs and "#" comments up to an empty line, then raw little-endian data. Written: type,
// dimension, sizes, labels, encoding raw, endian little, optional spacings (+ units mm) and space
// origin, then the caller's key:=value pairs.
#include "ctlproj.h"
#include "errors.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

using namespace ctlrt;

namespace {

bool littleEndianHost()
{
    const uint16_t one = 1;
    return *reinterpret_cast<const uint8_t*>(&one) == 1;
}

ctlproj_status ioFail(const std::string& msg) { return fail(CTLPROJ_ERR_IO, msg); }

int64_t fileSize(std::ifstream& f)
{
    const auto here = f.tellg();
    f.seekg(0, std::ios::end);
    const auto end = f.tellg();
    f.seekg(here);
    return int64_t(end);
}

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

// NRRD element type names (nrrd specification, "type:" field) -> bytes per element, code
int nrrdType(const std::string& t, int* code)
{
    static const struct { const char* name; int bytes; int code; } types[] = {
        { "uchar", 1, CTLPROJ_ELEM_U8 }, { "unsigned char", 1, CTLPROJ_ELEM_U8 },
        { "uint8", 1, CTLPROJ_ELEM_U8 }, { "uint8_t", 1, CTLPROJ_ELEM_U8 },
        { "ushort", 2, CTLPROJ_ELEM_U16 }, { "unsigned short", 2, CTLPROJ_ELEM_U16 },
        { "unsigned short int", 2, CTLPROJ_ELEM_U16 }, { "uint16", 2, CTLPROJ_ELEM_U16 },
        { "uint16_t", 2, CTLPROJ_ELEM_U16 }, { "float", 4, CTLPROJ_ELEM_F32 },
        { "double", 8, CTLPROJ_ELEM_F64 },
    };
    for (const auto& e : types)
        if (t == e.name) {
            *code = e.code;
            return e.bytes;
        }
    return 0;
}

const char* nrrdTypeName(int elem)
{
    switch (elem) {
    case CTLPROJ_ELEM_U8: return "uchar";
    case CTLPROJ_ELEM_U16: return "ushort";
    case CTLPROJ_ELEM_F32: return "float";
    case CTLPROJ_ELEM_F64: return "double";
    default: return nullptr;
    }
}

bool validElem(int elem)
{
    return elem == CTLPROJ_ELEM_U8 || elem == CTLPROJ_ELEM_U16 || elem == CTLPROJ_ELEM_F32 ||
           elem == CTLPROJ_ELEM_F64;
}

struct NrrdHeader {
    ctlproj_nrrd_header info{};
    std::string keyvals; // "key:=value\n" lines in file order
};

ctlproj_status parseNrrd(const char* path, NrrdHeader& h)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return ioFail(std::string("cannot open file: ") + path);
    std::string line;
    if (!std::getline(f, line)) return ioFail("empty file");
    if (line.size() != 8 || line.compare(0, 7, "NRRD000") != 0 || !std::isdigit(static_cast<unsigned char>(line[7])))
        return ioFail("no valid nrrd file - the magic first line is missing");
    h.info.version = uint32_t(line[7] - '0');
    bool have_type = false, have_dim = false, have_sizes = false;
    bool ended = false;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) {
            ended = true;
            break;
        }
        if (line[0] == '#') continue; // comment
        const size_t kv = line.find(":=");
        if (kv != std::string::npos) {
            h.keyvals += line.substr(0, kv) + ":=" + line.substr(kv + 2) + "\n";
            continue;
        }
        const size_t c = line.find(": ");
        if (c == std::string::npos) return ioFail("invalid header entry: " + line);
        const std::string field = trim(line.substr(0, c)), desc = trim(line.substr(c + 2));
        std::istringstream is(desc);
        if (field == "type") {
            int code = 0;
            h.info.elem_bytes = uint32_t(nrrdType(desc, &code));
            h.info.elem = code;
            if (!h.info.elem_bytes) return ioFail("unsupported data type: " + desc);
            have_type = true;
        } else if (field == "dimension") {
            if (!(is >> h.info.ndim) || h.info.ndim < 1 || h.info.ndim > 4)
                return ioFail("unsupported dimension: " + desc);
            have_dim = true;
        } else if (field == "sizes") {
            uint32_t n = 0;
            uint64_t v;
            while (n < 4 && (is >> v)) h.info.sizes[n++] = v;
            if (n == 0) return ioFail("invalid sizes: " + desc);
            h.info.nsizes = n;
            have_sizes = true;
        } else if (field == "encoding") {
            if (desc != "raw") return ioFail("unsupported data encoding: " + desc);
        } else if (field == "endian") {
            if (desc != "little") return ioFail("unsupported endianness: " + desc);
        } else if (field == "spacings") {
            for (int k = 0; k < 3 && (is >> h.info.spacing[k]); ++k) h.info.has_spacing = 1;
        } else if (field == "space origin") {
            double o[3];
            if (std::sscanf(desc.c_str(), " ( %lf , %lf , %lf )", &o[0], &o[1], &o[2]) == 3) {
                for (int k = 0; k < 3; ++k) h.info.origin[k] = o[k];
                h.info.has_origin = 1;
            }
        }
        // other fields (labels, units, space, kinds, ...) carry no data layout
    }
    if (!ended) return ioFail("nrrd header does not end with an empty line");
    if (!have_type || !have_dim || !have_sizes || h.info.nsizes != h.info.ndim)
        return ioFail("nrrd header lacks type, dimension or sizes");
    if (h.info.elem_bytes > 1 && !littleEndianHost()) return ioFail("big-endian hosts are not supported");
    h.info.header_bytes = uint64_t(f.tellg());
    uint64_t n = 1;
    for (uint32_t k = 0; k < h.info.ndim; ++k) n *= h.info.sizes[k];
    h.info.count = n;
    const int64_t total = fileSize(f);
    if (total < 0 || uint64_t(total) - h.info.header_bytes < n * h.info.elem_bytes)
        h.info.data_complete = 0;
    else
        h.info.data_complete = 1;
    return CTLPROJ_OK;
}

} // namespace

extern "C" {

// ---- DEN --------------------------------------------------------------------------------------

ctlproj_status ctlproj_den_info(const char* path, uint32_t* rows, uint32_t* cols, uint32_t* count,
                                uint32_t* elem_bytes)
{
    if (!path || !rows || !cols || !count || !elem_bytes) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return ioFail(std::string("cannot open file: ") + path);
    uint8_t b[6];
    if (!f.read(reinterpret_cast<char*>(b), 6)) return ioFail("file shorter than a DEN header");
    *rows = uint32_t(b[0]) | uint32_t(b[1]) << 8;
    *cols = uint32_t(b[2]) | uint32_t(b[3]) << 8;
    *count = uint32_t(b[4]) | uint32_t(b[5]) << 8;
    const uint64_t n = uint64_t(*rows) * *cols * *count;
    const int64_t total = fileSize(f);
    if (n == 0 || total < 6) return ioFail("DEN header with a zero size");
    const uint64_t data = uint64_t(total) - 6;
    // den_utils.tpp: the element type follows from bytes per element
    if (data % n != 0 || (data / n != 1 && data / n != 2 && data / n != 4 && data / n != 8))
        return ioFail("DEN data size does not fit the header");
    *elem_bytes = uint32_t(data / n);
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_den_read(const char* path, void* out, size_t out_bytes)
{
    uint32_t r, c, n, eb;
    ctlproj_status st = ctlproj_den_info(path, &r, &c, &n, &eb);
    if (st != CTLPROJ_OK) return st;
    const uint64_t bytes = uint64_t(r) * c * n * eb;
    if (!out || out_bytes != bytes)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output holds " + std::to_string(out_bytes) +
                                                     " bytes, the file " + std::to_string(bytes));
    if (eb > 1 && !littleEndianHost()) return ioFail("big-endian hosts are not supported");
    std::ifstream f(path, std::ios::binary);
    f.seekg(6);
    if (!f.read(static_cast<char*>(out), std::streamsize(bytes))) return ioFail("short read");
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_den_write(const char* path, const void* data, uint32_t elem_bytes, uint32_t rows,
                                 uint32_t cols, uint32_t count)
{
    if (!path || !data) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "element size must be 1, 2, 4 or 8 bytes");
    // den_header.h isOutOfBounds: every size fits an unsigned short
    if (rows == 0 || cols == 0 || count == 0 || rows > 0xffff || cols > 0xffff || count > 0xffff)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "DEN sizes must be 1..65535");
    if (elem_bytes > 1 && !littleEndianHost()) return ioFail("big-endian hosts are not supported");
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) return ioFail(std::string("cannot create file: ") + path);
    const uint8_t b[6] = { uint8_t(rows), uint8_t(rows >> 8), uint8_t(cols), uint8_t(cols >> 8),
                           uint8_t(count), uint8_t(count >> 8) };
    f.write(reinterpret_cast<const char*>(b), 6);
    f.write(static_cast<const char*>(data), std::streamsize(uint64_t(rows) * cols * count * elem_bytes));
    return f ? CTLPROJ_OK : ioFail("write failed");
}

// ---- NRRD -------------------------------------------------------------------------------------

ctlproj_status ctlproj_nrrd_read_header(const char* path, ctlproj_nrrd_header* info, char* keyvals, size_t keyvals_cap)
{
    if (!path || !info) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    NrrdHeader h;
    ctlproj_status st = parseNrrd(path, h);
    if (st != CTLPROJ_OK) return st;
    *info = h.info;
    if (keyvals && keyvals_cap) {
        const size_t n = std::min(keyvals_cap - 1, h.keyvals.size());
        std::memcpy(keyvals, h.keyvals.data(), n);
        keyvals[n] = '\0';
    }
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_nrrd_read(const char* path, void* out, size_t out_bytes)
{
    if (!path || !out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    NrrdHeader h;
    ctlproj_status st = parseNrrd(path, h);
    if (st != CTLPROJ_OK) return st;
    const uint64_t bytes = h.info.count * h.info.elem_bytes;
    if (out_bytes != bytes)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output holds " + std::to_string(out_bytes) +
                                                     " bytes, the file " + std::to_string(bytes));
    if (!h.info.data_complete) return ioFail("raw data size of file does not fit to dimensions in nrrd header");
    std::ifstream f(path, std::ios::binary);
    f.seekg(std::streamoff(h.info.header_bytes));
    if (!f.read(static_cast<char*>(out), std::streamsize(bytes))) return ioFail("short read");
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_nrrd_write(const char* path, const void* data, int elem, uint32_t ndim,
                                  const uint64_t* sizes, const char* labels, const double* spacing,
                                  const double* origin, const char* keyvals)
{
    if (!path || !data || !sizes) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!validElem(elem)) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "unknown element type");
    if (ndim < 1 || ndim > 4) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "dimension must be 1..4");
    uint64_t n = 1;
    for (uint32_t k = 0; k < ndim; ++k) {
        if (sizes[k] == 0) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "zero size");
        n *= sizes[k];
    }
    if (elem > 1 && !littleEndianHost()) return ioFail("big-endian hosts are not supported");
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) return ioFail(std::string("cannot create file: ") + path);
    f << "NRRD0004\n";
    f << "type: " << nrrdTypeName(elem) << '\n';
    f << "dimension: " << ndim << '\n';
    f << "sizes:";
    for (uint32_t k = 0; k < ndim; ++k) f << ' ' << sizes[k];
    f << '\n';
    if (labels && *labels) f << "labels: " << labels << '\n';
    f << "encoding: raw\n";
    f << "endian: little\n";
    char buf[256];
    if (spacing) {
        std::snprintf(buf, sizeof buf, "spacings: %.9g %.9g %.9g\n", spacing[0], spacing[1], spacing[2]);
        f << buf << "units:";
        for (uint32_t k = 0; k < ndim; ++k) f << " \"mm\"";
        f << '\n';
    }
    if (origin) {
        std::snprintf(buf, sizeof buf, "space: scanner-xyz\nspace origin: (%.9g,%.9g,%.9g)\n", origin[0],
                      origin[1], origin[2]);
        f << buf;
    }
    if (keyvals && *keyvals) {
        std::string kv(keyvals);
        if (kv.back() != '\n') kv += '\n';
        f << kv;
    }
    f << '\n';
    f.write(static_cast<const char*>(data), std::streamsize(n * uint64_t(elem)));
    return f ? CTLPROJ_OK : ioFail("write failed");
}

} // extern "C"
