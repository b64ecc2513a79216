// sf_host.cpp -- host-only part of the C-ABI: error text, parameter
// validation, the world-boundary particle builder and the bgeo V5 writer.
// Nothing here touches the GPU, so all of it is testable on a CPU-only host.
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sf_seepage.h"

namespace sf {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

static bool finite3(sf_float3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

int validate(const sf_params* p, const sf_boundary_params* b) {
    if (!p || !b) return set_error(SF_E_INVALID, "null parameter struct");
    if (!(b->kernel_radius > 0.f) || !std::isfinite(b->kernel_radius))
        return set_error(SF_E_INVALID, "kernel_radius must be finite and > 0");
    if (!finite3(b->lowest_point) || !finite3(b->highest_point))
        return set_error(SF_E_INVALID, "box corners must be finite");
    sf_int3 g;
    if (sf_grid_size(b->lowest_point, b->highest_point, b->kernel_radius, &g) != SF_OK) return SF_E_INVALID;
    if (b->grid_size.x < 1 || b->grid_size.y < 1 || b->grid_size.z < 1)
        return set_error(SF_E_INVALID, "grid_size must be >= 1 in every axis");
    if (g.x != b->grid_size.x || g.y != b->grid_size.y || g.z != b->grid_size.z)
        return set_error(SF_E_INVALID, "grid_size differs from make_int3((highest - lowest) / kernel_radius)");
    double cells = (double)b->grid_size.x * b->grid_size.y * b->grid_size.z;
    if (cells > 2147483000.0) return set_error(SF_E_INVALID, "grid has more than 2^31 cells");
    static_assert(offsetof(sf_params, sf_amc_p) == offsetof(sf_params, sph_mass) + 24 * sizeof(float),
                  "sph_mass .. sf_amc_p must be 25 consecutive floats");
    {  // the 25 scalar fields sph_mass .. sf_amc_p are consecutive floats
        const float* f = &p->sph_mass;
        for (int q = 0; q < 25; ++q)
            if (!std::isfinite(f[q])) return set_error(SF_E_INVALID, "sf_params has a non-finite field");
        if (!finite3(p->gravity) || !finite3(p->sf_dry_sand_color) || !finite3(p->sf_wet_sand_color))
            return set_error(SF_E_INVALID, "sf_params has a non-finite vector field");
        if (!(p->dt > 0.f) || !std::isfinite(p->dt)) return set_error(SF_E_INVALID, "dt must be finite and > 0");
    }
    if (!(p->sph_density > 0.f)) return set_error(SF_E_INVALID, "sph_density must be > 0");
    if (!(p->dem_density > 0.f)) return set_error(SF_E_INVALID, "dem_density must be > 0");
    if (!(p->sph_mass > 0.f)) return set_error(SF_E_INVALID, "sph_mass must be > 0");
    if (!(p->dem_particle_radius > 0.f) || !(p->sph_particle_radius > 0.f))
        return set_error(SF_E_INVALID, "particle radii must be > 0");
    if (p->solver_type != SF_SPH_SOLVER && p->solver_type != SF_WCSPH_SOLVER)
        return set_error(SF_E_INVALID, "solver_type must be SF_SPH_SOLVER or SF_WCSPH_SOLVER");
    if (p->sf_type != SF_TYPE_SF && p->sf_type != SF_TYPE_MULTI_SF)
        return set_error(SF_E_INVALID, "sf_type must be SF_TYPE_SF or SF_TYPE_MULTI_SF");
    if (!(p->sf_cmc_p > 0.f) || !(p->sf_amc_p > 0.f))
        return set_error(SF_E_INVALID, "Bezier control abscissae sf_cmc_p / sf_amc_p must be > 0");
    return SF_OK;
}

// ---- big-endian writer ------------------------------------------------------
struct BeOut {
    std::vector<unsigned char> buf;
    void u8(unsigned char v) { buf.push_back(v); }
    void u16(uint16_t v) {
        u8((unsigned char)(v >> 8));
        u8((unsigned char)v);
    }
    void u32(uint32_t v) {
        u8((unsigned char)(v >> 24));
        u8((unsigned char)(v >> 16));
        u8((unsigned char)(v >> 8));
        u8((unsigned char)v);
    }
    void i32(int32_t v) { u32((uint32_t)v); }
    void f32(float v) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        u32(u);
    }
    void str(const char* s) {
        uint16_t n = (uint16_t)std::strlen(s);
        u16(n);
        for (uint16_t i = 0; i < n; ++i) u8((unsigned char)s[i]);
    }
};

}  // namespace sf

using namespace sf;

extern "C" {

int sf_abi_version(void) { return SF_ABI_VERSION; }

const char* sf_last_error(void) { return sf::g_last_error.c_str(); }

int sf_validate_params(const sf_params* p, const sf_boundary_params* b) { return sf::validate(p, b); }

int sf_grid_size(sf_float3 lo, sf_float3 hi, float h, sf_int3* out) {
    if (!out) return set_error(SF_E_INVALID, "null output");
    if (!(h > 0.f)) return set_error(SF_E_INVALID, "kernel radius must be > 0");
    const float e[3] = {(hi.x - lo.x) / h, (hi.y - lo.y) / h, (hi.z - lo.z) / h};
    // the truncation below is defined only for values an int holds
    for (float v : e)
        if (!(v >= 0.f && v < 1073741824.f))
            return set_error(SF_E_INVALID, "box extent / kernel radius must be in [0, 2^30) per axis");
    out->x = (int)e[0];
    out->y = (int)e[1];
    out->z = (int)e[2];
    return SF_OK;
}

// CudaBoundaryEmitter::BuildWorldBoundary (cuda_boundary_emitter.cpp:12-80):
// the loop counters are size_t compared against float plane extents, and
// positions are lowest + (float)i * spacing in FP32.
int sf_build_world_boundary(sf_float3 lo, sf_float3 hi, float r, size_t capacity, float* pos3, uint64_t* label,
                            size_t* n_out) {
    if (!n_out) return set_error(SF_E_INVALID, "null n_out");
    if (!(r > 0.f)) return set_error(SF_E_INVALID, "particle radius must be > 0");
    const float sp = r * 2.f;
    const float sx = (hi.x - lo.x) / sp, sy = (hi.y - lo.y) / sp, sz = (hi.z - lo.z) / sp;
    if (!(sx >= 0.f && sy >= 0.f && sz >= 0.f) || sx > 1e7f || sy > 1e7f || sz > 1e7f)
        return set_error(SF_E_INVALID, "box extents must be >= 0 and at most 1e7 spacings");
    // at most (e + 2)^2 points per plane: keep the walk (and a count-only
    // call, which runs it too) bounded
    if (2.0 * ((sx + 2.0) * (sz + 2.0) + (sx + 2.0) * (sy + 2.0) + (sy + 2.0) * (sz + 2.0)) > 2e9)
        return set_error(SF_E_INVALID, "the box needs more than 2e9 boundary particles");
    struct Plane {
        int a, b;         // axes iterated (0=x,1=y,2=z)
        int fixed_axis;   // axis held constant
        bool at_high;     // constant value lowest or highest
    };
    // bottom(ZX,lo.y) top(ZX,hi.y) back(XY,lo.z) front(XY,hi.z) left(YZ,lo.x) right(YZ,hi.x)
    const Plane planes[6] = {{0, 2, 1, false}, {0, 2, 1, true}, {0, 1, 2, false},
                             {0, 1, 2, true},  {1, 2, 0, false}, {1, 2, 0, true}};
    const float ext[3] = {sx, sy, sz};
    const float lov[3] = {lo.x, lo.y, lo.z}, hiv[3] = {hi.x, hi.y, hi.z};
    size_t k = 0;
    for (const Plane& pl : planes) {
        for (size_t i = 0; (float)i <= ext[pl.a]; ++i) {
            for (size_t j = 0; (float)j <= ext[pl.b]; ++j) {
                if (pos3) {
                    if (k >= capacity) return set_error(SF_E_CAPACITY, "boundary output capacity too small");
                    float v[3];
                    v[pl.a] = lov[pl.a] + (float)i * sp;
                    v[pl.b] = lov[pl.b] + (float)j * sp;
                    v[pl.fixed_axis] = pl.at_high ? hiv[pl.fixed_axis] : lov[pl.fixed_axis];
                    pos3[3 * k] = v[0];
                    pos3[3 * k + 1] = v[1];
                    pos3[3 * k + 2] = v[2];
                    if (label) label[k] = 0;
                }
                ++k;
            }
        }
    }
    *n_out = k;
    return SF_OK;
}

// Uncompressed Houdini bgeo V5, attribute layout of ExportBgeoFileCUDA
// (kiri_utils.h:390-394) as Partio::writeBGEO lays it out (BGEO.cpp:311-476).
int sf_write_bgeo(const char* path, size_t n, const float* pos3, const float* col3, const float* radius,
                  const uint64_t* label) {
    if (!path) return set_error(SF_E_INVALID, "null path");
    if (n > 0 && (!pos3 || !col3 || !radius || !label)) return set_error(SF_E_INVALID, "null particle arrays");
    if (n > 0x7fffffffu) return set_error(SF_E_INVALID, "too many particles for a bgeo file");
    BeOut o;
    o.buf.reserve(64 + n * 36);
    o.u32(0x4267656fu);  // 'Bgeo'
    o.u8('V');
    o.i32(5);            // version
    o.i32((int32_t)n);   // nPoints
    o.i32(0);            // nPrims
    o.i32(0);            // nPointGroups
    o.i32(0);            // nPrimGroups
    o.i32(3);            // nPointAttrib (Cd, pscale, label; position is implicit)
    o.i32(0);            // nVertexAttrib
    o.i32(0);            // nPrimAttrib
    o.i32(0);            // nAttrib (fixed)
    struct Attr {
        const char* name;
        uint16_t count;
        int32_t type;
    };
    const Attr attrs[3] = {{"Cd", 3, 0}, {"pscale", 1, 0}, {"label", 1, 1}};
    for (const Attr& a : attrs) {
        o.str(a.name);
        o.u16(a.count);
        o.i32(a.type);
        for (int i = 0; i < a.count; ++i) o.i32(0);
    }
    for (size_t i = 0; i < n; ++i) {
        o.f32(pos3[3 * i]);
        o.f32(pos3[3 * i + 1]);
        o.f32(pos3[3 * i + 2]);
        o.f32(1.f);
        o.f32(col3[3 * i]);
        o.f32(col3[3 * i + 1]);
        o.f32(col3[3 * i + 2]);
        o.f32(radius[i]);
        o.i32((int32_t)label[i]);
    }
    o.u8(0x00);
    o.u8(0xff);
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return set_error(SF_E_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    size_t w = std::fwrite(o.buf.data(), 1, o.buf.size(), fp);
    int rc = std::fclose(fp);
    if (w != o.buf.size() || rc != 0) return set_error(SF_E_IO, std::string("short write to ") + path);
    return SF_OK;
}

}  // extern "C"

namespace sf {
// mkdir -p (ExportBgeoFileCUDA creates the export folder, kiri_utils.h:378-388)
int make_dirs(const std::string& path) {
    if (path.empty()) return SF_OK;
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty() && cur != "/") {
                struct stat st;
                if (stat(cur.c_str(), &st) != 0) {
                    if (mkdir(cur.c_str(), 0775) != 0 && errno != EEXIST)
                        return set_error(SF_E_IO, "cannot create folder " + cur + ": " + std::strerror(errno));
                } else if (!S_ISDIR(st.st_mode)) {
                    return set_error(SF_E_IO, cur + " exists and is not a folder");
                }
            }
        }
        if (i < path.size()) cur.push_back(path[i]);
    }
    return SF_OK;
}
}  // namespace sf
