// sf_scene.cpp -- scene input on the host: the bgeo V5 point reader behind
// ReadBgeoFileForGPU (seepage_flows/include/kiri_utils.h:252-308) and the
// sand shape builders CudaVolumeEmitter::BuildSeepageflowShapeVolume /
// BuildSeepageflowShapeMultiVolume (src/kiri_pbs_cuda/emitter/
// cuda_volume_emitter.cpp:109-175).  Nothing here touches the GPU.
//
// The arithmetic is the reference's FP32 host arithmetic, evaluated in the
// order it is written and without contraction into FMAs.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/sf_seepage.h"

#pragma clang fp contract(off)

namespace sf {
int set_error(int code, const std::string& msg);
}
using sf::set_error;

namespace {

// Whole file, inflated when it is gzip (gzread passes other files through).
int slurp(const char* path, std::vector<unsigned char>& out) {
    gzFile f = gzopen(path, "rb");
    if (!f) return set_error(SF_E_IO, std::string("cannot open ") + path);
    unsigned char buf[1 << 16];
    for (;;) {
        const int got = gzread(f, buf, sizeof(buf));
        if (got < 0) {
            int zerr = 0;
            const std::string msg = gzerror(f, &zerr);
            gzclose(f);
            return set_error(SF_E_IO, std::string("read error in ") + path + ": " + msg);
        }
        if (got == 0) break;
        out.insert(out.end(), buf, buf + got);
    }
    gzclose(f);
    return SF_OK;
}

struct BeIn {
    const unsigned char* p;
    size_t n, at = 0;
    bool ok = true;
    bool need(size_t k) {
        if (!ok || n - at < k) ok = false;
        return ok;
    }
    uint32_t u32() {
        if (!need(4)) return 0;
        const uint32_t v = (uint32_t)p[at] << 24 | (uint32_t)p[at + 1] << 16 | (uint32_t)p[at + 2] << 8 | p[at + 3];
        at += 4;
        return v;
    }
    int32_t i32() { return (int32_t)u32(); }
    uint16_t u16() {
        if (!need(2)) return 0;
        const uint16_t v = (uint16_t)(p[at] << 8 | p[at + 1]);
        at += 2;
        return v;
    }
    std::string str(size_t len) {
        if (!need(len)) return std::string();
        std::string s((const char*)p + at, len);
        at += len;
        return s;
    }
    void skip(size_t k) {
        if (need(k)) at += k;
    }
};

float word_as_float(uint32_t w) {
    float f;
    std::memcpy(&f, &w, 4);
    return f;
}

// One point attribute: Houdini type (0 float, 1 int, 4 indexed string,
// 5 vector), component count and first word within a point record.
struct PointAttr {
    std::string name;
    int32_t type;
    uint16_t size;
    size_t word;
};

// Partio's readBGEO (extlib/include/partio/io/BGEO.cpp:196-309) with its
// point-attribute table (getAttributes, :85-150): each point is the
// homogeneous position (4 words) followed by every point attribute's words.
// Returns the point record size in words and the attribute table; `points`
// is the byte offset of the first record.
int parse_header(const std::vector<unsigned char>& file, const char* path, size_t* n_points, size_t* record_words,
                 std::vector<PointAttr>* attrs, size_t* points) {
    BeIn in{file.data(), file.size()};
    const std::string magic = in.str(4);
    if (!in.ok) return set_error(SF_E_IO, std::string(path) + ": not a bgeo file (too short)");
    if (magic != "Bgeo") {
        if (magic == std::string("\x7f" "NSJ", 4))
            return set_error(SF_E_IO, std::string(path) + ": new (JSON) bgeo format; only the classic V5 format "
                                                          "is read (write .bhclassic from Houdini)");
        return set_error(SF_E_IO, std::string(path) + ": bad bgeo magic");
    }
    in.skip(1);  // version character ('V')
    const int32_t version = in.i32();
    const int32_t np = in.i32();
    in.i32();  // nPrims
    in.i32();  // nPointGroups
    in.i32();  // nPrimGroups
    const int32_t n_point_attr = in.i32();
    in.i32();  // nVertexAttrib
    in.i32();  // nPrimAttrib
    in.i32();  // nAttrib (detail)
    if (!in.ok) return set_error(SF_E_IO, std::string(path) + ": truncated bgeo header");
    if (version != 5) return set_error(SF_E_IO, std::string(path) + ": bgeo version must be 5");
    if (np < 0 || n_point_attr < 0) return set_error(SF_E_IO, std::string(path) + ": negative bgeo counts");
    size_t words = 4;
    attrs->clear();
    for (int32_t a = 0; a < n_point_attr; ++a) {
        const std::string name = in.str(in.u16());
        const uint16_t size = in.u16();
        const int32_t type = in.i32();
        if (!in.ok) return set_error(SF_E_IO, std::string(path) + ": truncated bgeo attribute table");
        if (type == 0 || type == 1 || type == 5) {  // float, int, vector: `size` default words follow
            in.skip(4 * (size_t)size);
        } else if (type == 4) {                     // indexed string: the string table follows
            const int32_t strings = in.i32();
            if (strings < 0) return set_error(SF_E_IO, std::string(path) + ": negative string-table size");
            for (int32_t s = 0; s < strings && in.ok; ++s) in.skip(in.u16());
        } else if (type == 2) {
            return set_error(SF_E_IO, std::string(path) + ": attribute '" + name + "' has type string, unsupported");
        } else {
            return set_error(SF_E_IO, std::string(path) + ": attribute '" + name + "' has unknown type " +
                                          std::to_string(type));
        }
        if (!in.ok) return set_error(SF_E_IO, std::string(path) + ": truncated bgeo attribute table");
        attrs->push_back(PointAttr{name, type, size, words});
        words += size;
    }
    if ((file.size() - in.at) / 4 / words < (size_t)np)
        return set_error(SF_E_IO, std::string(path) + ": truncated bgeo point data");
    *n_points = (size_t)np;
    *record_words = words;
    *points = in.at;
    return SF_OK;
}

// first attribute of that name: the one a lookup by name finds
const PointAttr* find_attr(const std::vector<PointAttr>& attrs, const char* name) {
    for (const PointAttr& a : attrs)
        if (a.name == name && a.size >= 1) return &a;
    return nullptr;
}

int read_bgeo_file(const char* path, std::vector<unsigned char>& file, size_t* np, size_t* words,
                   std::vector<PointAttr>* attrs, size_t* at) {
    if (int rc = slurp(path, file)) return rc;
    return parse_header(file, path, np, words, attrs, at);
}

uint32_t word_at(const unsigned char* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

int check_volume(const sf_volume_data* d, const float* shape4, size_t n, bool multi) {
    if (!d) return set_error(SF_E_INVALID, "null volume data");
    if (n > 0 && !shape4) return set_error(SF_E_INVALID, "null shape");
    if (d->size > d->capacity || n > d->capacity - d->size)
        return set_error(SF_E_CAPACITY, "volume data capacity too small for the shape");
    if (n > 0 && (!d->pos3 || !d->col3 || !d->label || !d->mass || !d->radius))
        return set_error(SF_E_INVALID, "null volume arrays");
    if (multi && n > 0 && (!d->cda0asat3 || !d->amcamcp2))
        return set_error(SF_E_INVALID, "the multi-material volume needs cda0asat3 and amcamcp2");
    return SF_OK;
}

// Shared body of both builders: appends the shape as sand (label 1) with
// mass = density * ((4/3) * pi * r^3), tracks the smallest radius and the
// lowest y of this shape, and with offset_y lowers every particle held so far
// so that this shape's lowest point sits 5 smallest radii above the floor.
void append_shape(sf_volume_data* d, const float* shape4, size_t n, sf_float3 color, float density, bool multi,
                  sf_float3 cda0asat, float amc, float amcp, int offset_y, float world_lowest_y) {
    const float huge = (float)std::numeric_limits<size_t>::max();  // Huge<size_t>() stored in a float
    const float kPi = 3.141592653589793238462643383279502884197f;   // KIRI_PI (kiri_pbs_pch.cuh:54)
    float min_y = huge;
    for (size_t i = 0; i < n; ++i) {
        const float* s = shape4 + 4 * i;
        const float radius = s[3];
        const size_t k = d->size + i;
        d->pos3[3 * k] = s[0];
        d->pos3[3 * k + 1] = s[1];
        d->pos3[3 * k + 2] = s[2];
        d->col3[3 * k] = color.x;
        d->col3[3 * k + 1] = color.y;
        d->col3[3 * k + 2] = color.z;
        d->label[k] = 1;
        d->radius[k] = radius;
        d->mass[k] = density * ((4.f / 3.f) * kPi * powf(radius, 3.f));
        d->sand_min_radius = std::min(radius, d->sand_min_radius);
        if (multi) {
            d->cda0asat3[3 * k] = cda0asat.x;
            d->cda0asat3[3 * k + 1] = cda0asat.y;
            d->cda0asat3[3 * k + 2] = cda0asat.z;
            d->amcamcp2[2 * k] = amc;
            d->amcamcp2[2 * k + 1] = amcp;
        }
        min_y = std::min(min_y, s[1]);
    }
    d->size += n;
    if (offset_y && n > 0) {
        const float off = min_y - (world_lowest_y + d->sand_min_radius * 5.f);
        for (size_t k = 0; k < d->size; ++k) {
            d->pos3[3 * k] = d->pos3[3 * k] - 0.f;
            d->pos3[3 * k + 1] = d->pos3[3 * k + 1] - off;
            d->pos3[3 * k + 2] = d->pos3[3 * k + 2] - 0.f;
        }
    }
}

}  // namespace

extern "C" {

int sf_read_bgeo(const char* path, int flip_yz, float* shape4, size_t capacity, size_t* n_out) {
    if (!path || !n_out) return set_error(SF_E_INVALID, "null path or n_out");
    std::vector<unsigned char> file;
    std::vector<PointAttr> attrs;
    size_t np = 0, words = 0, at = 0;
    if (int rc = read_bgeo_file(path, file, &np, &words, &attrs, &at)) return rc;
    const PointAttr* ps_attr = find_attr(attrs, "pscale");
    const long pscale = ps_attr ? (long)ps_attr->word : -1;
    *n_out = np;
    if (!shape4) return SF_OK;
    if (capacity < np) return set_error(SF_E_CAPACITY, "shape capacity too small for the file's points");
    for (size_t i = 0; i < np; ++i) {
        const unsigned char* rec = file.data() + at + 4 * words * i;
        const float x = word_as_float(word_at(rec)), y = word_as_float(word_at(rec + 4)),
                    z = word_as_float(word_at(rec + 8));
        const float ps = pscale >= 0 ? word_as_float(word_at(rec + 4 * pscale)) : 0.01f;
        float* o = shape4 + 4 * i;
        o[0] = x;
        o[1] = flip_yz ? z : y;
        o[2] = flip_yz ? y : z;
        o[3] = ps;
    }
    return SF_OK;
}

// Partio2VTK (kiri_utils.h:54-250): legacy VTK 4.1, binary (big-endian), an
// unstructured grid with one vertex cell per particle (cell point = the "id"
// attribute, else the index), then every other attribute as FIELD data:
// float and vector as float, int as int.
int sf_bgeo_to_vtk(const char* bgeo_path, const char* vtk_path) {
    if (!bgeo_path || !vtk_path) return set_error(SF_E_INVALID, "null path");
    std::vector<unsigned char> file;
    std::vector<PointAttr> attrs;
    size_t np = 0, words = 0, at = 0;
    if (int rc = read_bgeo_file(bgeo_path, file, &np, &words, &attrs, &at)) return rc;
    if (np == 0) return SF_OK;  // nothing is written for an empty file
    if (np > 0x7fffffffu) return set_error(SF_E_INVALID, "too many particles for a VTK file");
    const unsigned char* rec0 = file.data() + at;
    const size_t stride = 4 * words;
    std::string out;
    out.reserve(np * (12 + 8 + 4) + 256);
    auto text = [&](const std::string& t) { out += t; };
    auto word = [&](uint32_t w) {  // big-endian, as the file stores it
        out.push_back((char)(w >> 24));
        out.push_back((char)(w >> 16));
        out.push_back((char)(w >> 8));
        out.push_back((char)w);
    };
    const std::string n = std::to_string(np);
    text("# vtk DataFile Version 4.1\nSeepage Flow Framework Particles Data\nBINARY\nDATASET UNSTRUCTURED_GRID\n");
    text("POINTS " + n + " float\n");
    for (size_t i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) word(word_at(rec0 + stride * i + 4 * c));
    text("\n");
    const PointAttr* id = find_attr(attrs, "id");
    text("CELLS " + n + " " + std::to_string(2 * np) + "\n");
    for (size_t i = 0; i < np; ++i) {
        word(1u);
        word(id ? word_at(rec0 + stride * i + 4 * id->word) : (uint32_t)i);
    }
    text("\n");
    text("CELL_TYPES " + n + "\n");
    for (size_t i = 0; i < np; ++i) word(1u);  // VTK_VERTEX
    text("\n");
    std::vector<const PointAttr*> fields;
    for (const PointAttr& a : attrs)
        if (&a != id && (a.type == 0 || a.type == 1 || a.type == 5)) fields.push_back(&a);
    text("POINT_DATA " + n + "\n");
    text("FIELD FieldData " + std::to_string(fields.size()) + "\n");
    for (const PointAttr* a : fields) {
        std::string name;
        bool space = false;
        for (char ch : a->name) {  // runs of whitespace become one '_'
            const bool ws = ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\f' || ch == '\v';
            if (ws && !space) name += '_';
            if (!ws) name += ch;
            space = ws;
        }
        const int count = a->type == 5 ? 3 : (int)a->size;
        text(name + " " + std::to_string(a->size) + " " + n + (a->type == 1 ? " int\n" : " float\n"));
        for (size_t i = 0; i < np; ++i)
            for (int c = 0; c < count && (size_t)c < a->size; ++c) word(word_at(rec0 + stride * i + 4 * (a->word + c)));
        text("\n");
    }
    FILE* f = std::fopen(vtk_path, "wb");
    if (!f) return set_error(SF_E_IO, std::string("cannot open ") + vtk_path + " for writing");
    const size_t wrote = std::fwrite(out.data(), 1, out.size(), f);
    const int closed = std::fclose(f);
    if (wrote != out.size() || closed != 0) return set_error(SF_E_IO, std::string("write error on ") + vtk_path);
    return SF_OK;
}

int sf_build_seepage_shape_volume(sf_volume_data* d, const float* shape4, size_t n, sf_float3 color,
                                  float sand_density, int offset_y, float world_lowest_y) {
    if (int rc = check_volume(d, shape4, n, false)) return rc;
    d->sand_min_radius = (float)std::numeric_limits<size_t>::max();
    append_shape(d, shape4, n, color, sand_density, false, sf_float3{0.f, 0.f, 0.f}, 0.f, 0.f, offset_y,
                 world_lowest_y);
    return SF_OK;
}

int sf_build_seepage_shape_multi_volume(sf_volume_data* d, const float* shape4, size_t n, sf_float3 color,
                                        float sand_density, sf_float3 cda0asat, float amc, float amcp, int offset_y,
                                        float world_lowest_y) {
    if (int rc = check_volume(d, shape4, n, true)) return rc;
    if (d->size == 0) d->sand_min_radius = (float)std::numeric_limits<size_t>::max();
    append_shape(d, shape4, n, color, sand_density, true, cda0asat, amc, amcp, offset_y, world_lowest_y);
    return SF_OK;
}

}  // extern "C"
