// seepage_main.cpp -- a main.cpp-style driver (seepage_flows/src/seepageflow/main.cpp:25-297)
// written against the KIRI drop-in layer: it sets the global parameter
// structs, builds world-boundary particles and a particle block, creates the
// system and runs the adaptive sub-step frame loop with a bgeo per frame.
//
// With --shape, it is the reference's own scene set-up (main.cpp:25-221):
// sand objects read from <resources>bgeo/<FOLDER>/<NAME>.bgeo (the reference
// loads bunny/bunny), built with BuildSeepageflowShapeMultiVolume onto the
// floor of a 1.5 x 1.5 x 2.5 m box, MULTI_SF, and water from the circle
// emitter, which main.cpp's Update() switches on at the first frame.
// Without it, the driver builds the small water-block + sand-pile scene
// (16,384 water + 8,192 sand, 1 m box) from the same lattice / SplitMix64
// jitter rules as sph_seepage_flows_amd/scenes.py.  Either way its frames can
// be compared bit for bit with the Python host.
//
//   seepage_main [--frames N] [--out FOLDER] [--render-interval S] [--fixed] [--export-initial] [--emit]
//                [--resources DIR] [--shape FOLDER/NAME[:CD]]... [--flip-yz] [--no-emit] [--device-frame] [--solver-api]
//
// --emit switches on main.cpp's inflow emitter (main.cpp:85-95: a circle of
// radius 0.17 at the box centre + (size/10, size/5, size/2.3), velocity
// (0, 0, -5)) in the small scene, with room for 20,000 more particles.
// --shape adds one sand object (CD: its drag coefficient, 0.5 by default as
// main.cpp:137-138); --no-emit keeps the shape scene's emitter off.
// --device-frame runs each frame with CudaSFSystem::UpdateFrame (the same
// loop inside the library) instead of calling UpdateSystem per sub-step.
// --solver-api runs each sub-step as CudaSFSystem::OnUpdateSolver does
// (cuda_sf_system.cu:66-84): searcher->BuildGNSearcher(particles), then
// solver->UpdateSolver(particles, boundaries, cellStart, bCellStart, ...)
// (no emitter).
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kiri_seepage.hpp"

using namespace KIRI;

namespace {

bool gAdaptive = true;  // adaptive_sub_timestep of main.cpp:198

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float uniform01(uint64_t seed, uint64_t stream, uint64_t idx) {
    const uint64_t key = (seed * 0xD1B54A32D192ED03ull) ^ (stream << 40) ^ idx;
    return (float)(splitmix64(key) >> 40) * (float)(1.0 / (1 << 24));
}
// (i, j, k) lattice, i outermost, truncated to n, optional per-component jitter
void lattice(Vec_Float3& out, float3 o, float spacing, int nx, int ny, int nz, size_t n, float jitter, uint64_t seed,
             uint64_t stream) {
    size_t k = 0;
    for (int i = 0; i < nx && k < n; ++i)
        for (int j = 0; j < ny && k < n; ++j)
            for (int l = 0; l < nz && k < n; ++l, ++k) {
                float3 p = make_float3(o.x + (float)i * spacing, o.y + (float)j * spacing, o.z + (float)l * spacing);
                if (jitter != 0.f) {
                    p.x = p.x + (uniform01(seed, stream, 3 * k) * 2.f - 1.f) * jitter;
                    p.y = p.y + (uniform01(seed, stream, 3 * k + 1) * 2.f - 1.f) * jitter;
                    p.z = p.z + (uniform01(seed, stream, 3 * k + 2) * 2.f - 1.f) * jitter;
                }
                out.push_back(p);
            }
}

CudaSFSystemPtr SetupParams(bool emit) {
    const float3 lo = make_float3(0.f, 0.f, 0.f), hi = make_float3(1.f, 1.f, 1.f);
    CUDA_SEEPAGEFLOW_APP_PARAMS.max_num = emit ? 24576 + 20000 : 24576;
    auto& P = CUDA_SEEPAGEFLOW_PARAMS;
    const float r = 0.01f, rs = 0.01f, rho0 = 1000.f, cs = 100.f;
    P.sph_density = rho0;
    P.sph_particle_radius = r;
    P.sph_kernel_radius = 4.f * r;
    const float diam = 2.f * r, h = P.sph_kernel_radius;
    P.sph_mass = 0.8f * diam * diam * diam * rho0;                 // main.cpp:44-45
    P.sph_stiff = 0.001f * cs * cs * rho0 / 7.f;                    // main.cpp:53
    P.sph_visc = 0.05f;
    P.sph_nu = (P.sph_visc + P.sph_visc) * h * cs;                  // main.cpp:49
    P.sph_bnu = 0.2f * h * cs;                                      // main.cpp:51
    P.dem_density = 2650.f;
    P.dem_particle_radius = rs;
    P.dem_mass = P.dem_density * ((float)(4.0 / 3.0) * 3.1415926f * (rs * rs * rs));  // main.cpp:61
    P.dem_kernel_radius = 4.f * rs;
    P.dem_young = 1e6f;
    P.dem_poisson = 0.3f;
    P.dem_tan_friction_angle = 0.5f;
    P.dem_damping = 0.4f;
    P.sf_c0 = 0.7f; P.sf_csat = 0.f; P.sf_cmc = 1.f; P.sf_cmc_p = 0.01f; P.sf_cd = 0.5f;   // main.cpp:68-77
    P.sf_a0 = 0.f; P.sf_asat = 0.8f; P.sf_amc = 1.5f; P.sf_amc_p = 0.5f;
    P.sf_dry_sand_color = sf_float3{0.88f, 0.79552f, 0.5984f};
    P.sf_wet_sand_color = sf_float3{0.38f, 0.29f, 0.14f};
    P.gravity = sf_float3{0.f, -9.81f, 0.f};
    P.dt = 1e-4f;
    P.solver_type = WCSPH_SOLVER;
    P.sf_type = SF;

    auto& B = CUDA_BOUNDARY_PARAMS;
    B.lowest_point = to_sf(lo);
    B.highest_point = to_sf(hi);
    B.world_size = sf_float3{hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
    B.world_center = sf_float3{(hi.x + lo.x) / 2.f, (hi.y + lo.y) / 2.f, (hi.z + lo.z) / 2.f};
    B.kernel_radius = P.sph_kernel_radius;
    B.grid_size = to_sf(GridSize(lo, hi, B.kernel_radius));       // main.cpp:104

    BoundaryData boundaryData;
    CudaBoundaryEmitter().BuildWorldBoundary(boundaryData, lo, hi, P.sph_particle_radius);

    Vec_Float3 pos, col;
    Vec_SizeT label;
    Vec_Float mass, radius;
    lattice(pos, make_float3(r, r, r), (float)(2 * 0.01), 15, 25, 44, 16384, 0.f, 0, 0);
    const size_t nw = pos.size();
    lattice(pos, make_float3((float)(0.5 + 0.01), rs, rs), (float)(2 * 0.01), 15, 13, 44, 8192, (float)(0.1 * 0.01), 0, 1);
    const float3 water_col = make_float3((float)(127.0 / 255.0), (float)(205.0 / 255.0), (float)(255.0 / 255.0));
    const float3 dry = make_float3(0.88f, 0.79552f, 0.5984f);
    for (size_t i = 0; i < pos.size(); ++i) {
        const bool w = i < nw;
        label.push_back(w ? 0 : 1);
        mass.push_back(w ? P.sph_mass : P.dem_mass);
        radius.push_back(w ? P.sph_particle_radius : P.dem_particle_radius);
        col.push_back(w ? water_col : dry);
    }
    auto particles = std::make_shared<CudaSFParticles>(CUDA_SEEPAGEFLOW_APP_PARAMS.max_num, pos, col, label, mass, radius);
    auto searcher = std::make_shared<CudaGNSearcher>(lo, hi, particles->MaxSize(), B.kernel_radius, SEEPAGE);
    auto boundary_particles = std::make_shared<CudaBoundaryParticles>(boundaryData.pos, boundaryData.label);
    auto solver = std::make_shared<CudaWCSphSFSolver>(particles->MaxSize(), 0.f, 3.f, 40.f);
    auto boundary_searcher = std::make_shared<CudaGNBoundarySearcher>(lo, hi, boundary_particles->MaxSize(), B.kernel_radius);
    // sph emitter (main.cpp:85-95, 106-111)
    const float3 size = make_float3(hi.x - lo.x, hi.y - lo.y, hi.z - lo.z);
    const float3 center = make_float3((hi.x + lo.x) / 2.f, (hi.y + lo.y) / 2.f, (hi.z + lo.z) / 2.f);
    auto& E = CUDA_SPH_EMITTER_PARAMS;
    E.enable = true;
    E.run = false;  // main() switches it on at the first frame (main.cpp:228-233)
    E.emit_pos = make_float3(center.x + size.x / 10.f, center.y + size.y / 5.f, center.z + size.z / 2.3f);
    E.emit_vel = make_float3(0.f, 0.f, -5.f);
    E.emit_col = make_float3(127.f / 255.f, 205.f / 255.f, 255.f / 255.f);
    E.emit_radius = 0.17f;
    E.emit_width = 0.22f;
    E.emit_height = 0.18f;
    E.emit_type = CIRCLE;
    CudaEmitterPtr emitter = std::make_shared<CudaEmitter>(E.emit_pos, E.emit_vel, E.enable);
    std::printf("Number of total particles = %zu, boundary particles = %zu\n", particles->Size(), boundary_particles->Size());
    return std::make_shared<CudaSFSystem>(particles, boundary_particles, solver, searcher, boundary_searcher, emitter,
                                          gAdaptive);
}

struct ShapeSpec {
    std::string folder, name;
    float cd;
};

// main.cpp:25-221 with the sand objects of `shapes`
CudaSFSystemPtr SetupShapeScene(const std::vector<ShapeSpec>& shapes, bool flip_yz) {
    // scene config (main.cpp:33-37)
    const float3 lo = make_float3(0.f, 0.f, 0.f), hi = make_float3(1.5f, 1.5f, 2.5f);
    const float3 size = make_float3(hi.x - lo.x, hi.y - lo.y, hi.z - lo.z);
    const float3 center = make_float3((hi.x + lo.x) / 2.f, (hi.y + lo.y) / 2.f, (hi.z + lo.z) / 2.f);
    CUDA_SEEPAGEFLOW_APP_PARAMS.max_num = 500000;

    // sph params (main.cpp:39-56)
    auto& P = CUDA_SEEPAGEFLOW_PARAMS;
    P.sph_density = 1000.f;
    P.sph_particle_radius = 0.01f;
    P.sph_kernel_radius = 4.f * P.sph_particle_radius;
    const float diam = 2.f * P.sph_particle_radius;
    P.sph_mass = 0.8f * diam * diam * diam * P.sph_density;
    const float visc = 0.05f, sound_speed = 100.f;
    const float nu = (visc + visc) * P.sph_kernel_radius * sound_speed;
    const float boundary_friction = 0.2f;
    const float bnu = boundary_friction * P.sph_kernel_radius * sound_speed;
    P.sph_stiff = 0.001f * sound_speed * sound_speed * P.sph_density / 7.f;
    P.sph_visc = visc;
    P.sph_nu = nu;
    P.sph_bnu = bnu;

    // dem and seepage params (main.cpp:58-83)
    P.dem_density = 2700.f;
    P.dem_particle_radius = 0.005f;
    P.dem_mass = P.dem_density * ((4.f / 3.f) * 3.1415926f * powf(P.dem_particle_radius, 3.f));
    P.dem_young = 1e5f;
    P.dem_poisson = 0.3f;
    P.dem_tan_friction_angle = 0.5f;
    P.dem_damping = 0.4f;
    P.sf_c0 = 0.7f;
    P.sf_cd = 0.5f;
    P.sf_csat = 0.f;
    P.sf_cmc = 1.f;
    P.sf_cmc_p = 0.01f;
    P.sf_a0 = 0.f;
    P.sf_asat = 0.8f;
    P.sf_amc = 1.5f;
    P.sf_amc_p = 0.5f;
    P.sf_dry_sand_color = sf_float3{0.88f, 0.79552f, 0.5984f};
    P.sf_wet_sand_color = sf_float3{0.38f, 0.29f, 0.14f};
    P.dt = 0.5f * P.dem_particle_radius / sqrtf(P.dem_young / P.dem_density);
    P.gravity = sf_float3{0.f, -9.8f, 0.f};

    // sph emitter (main.cpp:85-95)
    auto& E = CUDA_SPH_EMITTER_PARAMS;
    E.enable = true;
    E.run = false;
    E.emit_pos = make_float3(center.x + size.x / 10.f, center.y + size.y / 5.f, center.z + size.z / 2.3f);
    E.emit_vel = make_float3(0.f, 0.f, -5.f);
    E.emit_col = make_float3(127.f / 255.f, 205.f / 255.f, 255.f / 255.f);
    E.emit_radius = 0.17f;
    E.emit_width = 0.22f;
    E.emit_height = 0.18f;
    E.emit_type = CIRCLE;

    // scene data (main.cpp:97-104)
    auto& B = CUDA_BOUNDARY_PARAMS;
    B.lowest_point = to_sf(lo);
    B.highest_point = to_sf(hi);
    B.world_size = to_sf(size);
    B.world_center = to_sf(center);
    B.kernel_radius = P.sph_kernel_radius;
    B.grid_size = to_sf(GridSize(lo, hi, B.kernel_radius));

    CudaEmitterPtr emitter = std::make_shared<CudaEmitter>(E.emit_pos, E.emit_vel, E.enable);
    BoundaryData boundaryData;
    CudaBoundaryEmitter().BuildWorldBoundary(boundaryData, lo, hi, P.sph_particle_radius);

    // shape sampling (main.cpp:117-166)
    P.sf_type = MULTI_SF;
    SeepageflowMultiVolumeData multiVolumeData;
    CudaVolumeEmitter volumeEmitter;
    for (size_t i = 0; i < shapes.size(); ++i) {
        const float3 cda0asat = make_float3(shapes[i].cd, P.sf_a0, P.sf_asat);
        const float2 amcamcp = make_float2(P.sf_amc, P.sf_amc_p);
        Vec_Float4 sandShape = ReadBgeoFileForGPU(shapes[i].folder, shapes[i].name, flip_yz);
        volumeEmitter.BuildSeepageflowShapeMultiVolume(multiVolumeData, sandShape, make_float3(0.88f, 0.79552f, 0.5984f),
                                                       P.dem_density, cda0asat, amcamcp, true, lo.y);
        std::printf("Object(%zu) Params: Cd A0 Asat Amc Amcp = %g, %g, %g, %g, %g\n", i + 1, cda0asat.x, cda0asat.y,
                    cda0asat.z, amcamcp.x, amcamcp.y);
    }

    // dt (main.cpp:168-172)
    P.dem_particle_radius = multiVolumeData.sandMinRadius;
    P.dt = 0.5f * P.dem_particle_radius / sqrtf(P.dem_young / P.dem_density);
    std::printf("Number of total particles = %zu\n", multiVolumeData.pos.size());

    auto particles = std::make_shared<CudaSFParticles>(CUDA_SEEPAGEFLOW_APP_PARAMS.max_num, multiVolumeData.pos,
                                                       multiVolumeData.col, multiVolumeData.label,
                                                       multiVolumeData.mass, multiVolumeData.radius,
                                                       multiVolumeData.cda0asat, multiVolumeData.amcamcp);
    auto searcher = std::make_shared<CudaGNSearcher>(lo, hi, particles->MaxSize(), B.kernel_radius, SEEPAGE_MULTI);
    auto boundary_particles = std::make_shared<CudaBoundaryParticles>(boundaryData.pos, boundaryData.label);
    std::printf("Number of Boundary Particles = %zu\n", boundary_particles->Size());
    auto solver = std::make_shared<CudaWCSphSFSolver>(particles->MaxSize());  // solver defaults (0, 3, 100)
    P.solver_type = WCSPH_SOLVER;
    auto boundary_searcher = std::make_shared<CudaGNBoundarySearcher>(lo, hi, boundary_particles->MaxSize(),
                                                                      B.kernel_radius);
    return std::make_shared<CudaSFSystem>(particles, boundary_particles, solver, searcher, boundary_searcher, emitter,
                                          gAdaptive);
}

bool parse_shape(const char* arg, ShapeSpec* out) {
    std::string a(arg);
    out->cd = 0.5f;
    const size_t colon = a.find(':');
    if (colon != std::string::npos) {
        out->cd = (float)std::atof(a.c_str() + colon + 1);
        a = a.substr(0, colon);
    }
    const size_t slash = a.rfind('/');
    if (slash == std::string::npos || slash == 0 || slash + 1 == a.size()) return false;
    out->folder = a.substr(0, slash);
    out->name = a.substr(slash + 1);
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    int frames = 2;
    bool export_initial = false, emit = false, no_emit = false, flip_yz = false, device_frame = false;
    bool solver_api = false;
    std::vector<ShapeSpec> shapes;
    float render_interval = 1.f / 30.f;
    std::string out = "export/bgeo/seepage_small_block";
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--frames") && i + 1 < argc) frames = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) out = argv[++i];
        else if (!std::strcmp(argv[i], "--render-interval") && i + 1 < argc) render_interval = (float)std::atof(argv[++i]);
        else if (!std::strcmp(argv[i], "--fixed")) gAdaptive = false;
        else if (!std::strcmp(argv[i], "--export-initial")) export_initial = true;
        else if (!std::strcmp(argv[i], "--emit")) emit = true;
        else if (!std::strcmp(argv[i], "--no-emit")) no_emit = true;
        else if (!std::strcmp(argv[i], "--flip-yz")) flip_yz = true;
        else if (!std::strcmp(argv[i], "--device-frame")) device_frame = true;
        else if (!std::strcmp(argv[i], "--solver-api")) solver_api = true;
        else if (!std::strcmp(argv[i], "--resources") && i + 1 < argc) DbPbrPath() = argv[++i];
        else if (!std::strcmp(argv[i], "--shape") && i + 1 < argc) {
            ShapeSpec sp;
            if (!parse_shape(argv[++i], &sp)) {
                std::fprintf(stderr, "--shape wants FOLDER/NAME[:CD], got %s\n", argv[i]);
                return 2;
            }
            shapes.push_back(sp);
        } else {
            std::fprintf(stderr, "unknown argument %s\n", argv[i]);
            return 2;
        }
    }
    if (!DbPbrPath().empty() && DbPbrPath().back() != '/') DbPbrPath() += '/';
    std::strncpy(CUDA_SEEPAGEFLOW_APP_PARAMS.bgeo_export_folder, out.c_str(),
                 sizeof(CUDA_SEEPAGEFLOW_APP_PARAMS.bgeo_export_folder) - 1);
    CUDA_SEEPAGEFLOW_APP_PARAMS.bgeo_export = true;
    CudaSFSystemPtr system;
    try {
        system = shapes.empty() ? SetupParams(emit) : SetupShapeScene(shapes, flip_yz);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "setup failed: %s\n", e.what());
        return 1;
    }
    if (export_initial) ExportBgeoFileCUDA(out, UInt2Str4Digit(0), system->GetSFParticles());
    // Update() at the first frame: the emitter runs with the app (main.cpp:228-233)
    if (CUDA_SPH_EMITTER_PARAMS.enable) CUDA_SPH_EMITTER_PARAMS.run = shapes.empty() ? emit : !no_emit;
    double total = 0.0;
    for (int frame = 1; frame <= frames; ++frame) {  // Update() main.cpp:223-283
        auto t0 = std::chrono::steady_clock::now();
        size_t sub = 0;
        if (device_frame) {
            sub = system->UpdateFrame(render_interval);
        } else if (system->GetAdaptiveSubTimeStep()) {
            float remaining_time = render_interval;
            while (remaining_time > 1e-6f) {
                if (solver_api) {
                    auto particles = system->GetSFParticles();
                    auto boundaries = system->GetBoundaries();
                    auto searcher = system->GetSearcher();
                    searcher->BuildGNSearcher(particles);
                    system->GetSolver()->UpdateSolver(particles, boundaries, searcher->GetCellStart(),
                                                      system->GetBoundarySearcher()->GetCellStart(), remaining_time,
                                                      CUDA_SEEPAGEFLOW_PARAMS, CUDA_BOUNDARY_PARAMS);
                } else {
                    system->UpdateSystem(remaining_time);
                }
                remaining_time -= remaining_time / static_cast<float>(system->GetNumOfSubTimeSteps());
                ++sub;
            }
        } else {
            const size_t k = system->GetNumOfSubTimeSteps();
            for (size_t i = 0; i < k; ++i, ++sub) system->UpdateSystem(render_interval);
        }
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        total += sec;
        std::printf("Simulation Frame=%d, sub-steps=%zu, Time Per Frame=%.4f s\n", frame, sub, sec);
        // the export call of main.cpp:264-275, with this loop's frame number
        if (CUDA_SEEPAGEFLOW_APP_PARAMS.bgeo_export)
        {
            auto particles = system->GetSFParticles();
            ExportBgeoFileCUDA(
                CUDA_SEEPAGEFLOW_APP_PARAMS.bgeo_export_folder,
                UInt2Str4Digit(frame),
                particles->GetPosPtr(),
                particles->GetColPtr(),
                particles->GetRadiusPtr(),
                particles->GetLabelPtr(),
                particles->Size());
        }
    }
    std::printf("Average Per Frame=%.4f s\n", total / frames);
    return 0;
}
