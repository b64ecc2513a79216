// kiri_seepage.hpp -- header-only C++ layer that re-creates the reference's
// seepage-flow class API (namespace KIRI) on top of the C-ABI in
// include/sf_seepage.h, so a main.cpp-style caller drops in:
//
//   reference                                         here
//   CudaSeepageflowParams / CUDA_SEEPAGEFLOW_PARAMS   = sf_params, inline global
//     (seepage_flows_cuda/include/kiri_pbs_cuda/data/cuda_seepageflow_params.h:25-84)
//   CudaBoundaryParams / CUDA_BOUNDARY_PARAMS         = sf_boundary_params, inline global
//     (.../data/cuda_boundary_params.h:19-29)
//   CudaBoundaryEmitter::BuildWorldBoundary           -> sf_build_world_boundary
//     (src/kiri_pbs_cuda/emitter/cuda_boundary_emitter.cpp:12-80)
//   CudaSFParticles(maxNum, pos, col, label, mass, radius[, cda0asat, amcamcp])
//     (.../particle/cuda_sf_particles.cuh:138-235)
//   CudaBoundaryParticles(pos, label)                 (.../particle/cuda_boundary_particles.cuh:21-28)
//   CudaGNSearcher / CudaGNBoundarySearcher           (.../searcher/cuda_neighbor_searcher.cuh:56-98)
//   CudaWCSphSFSolver(num, negativeScale, timeStepLimitScale, speedOfSound)
//     (.../solver/seepageflow/cuda_wcsph_sf_solver.cuh:19-68)
//   CudaSFSystem(...), UpdateSystem(renderInterval), GetNumOfSubTimeSteps()
//     (.../system/cuda_sf_system.cuh:15-50, cuda_base_system.cuh:17-44)
//   CudaGNBaseSearcher::BuildGNSearcher / GetCellStart / GetGridIdxArray
//     -> sf_build_neighbors, sf_particle_view (.../searcher/cuda_neighbor_searcher.cuh:29-50)
//   CudaWCSphSFSolver::UpdateSolver(particles, boundaries, cellStart, bCellStart,
//     renderInterval, params, bparams) -> sf_update_solver (.../solver/seepageflow/cuda_wcsph_sf_solver.cuh:36-43)
//   CudaSFParticles::Get*Ptr()                        -> sf_particle_view
//     (.../particle/cuda_sf_particles.cuh:170-190, cuda_particles.cuh:57): device arrays
//     in the reference's layout (float3 / float / size_t / float2), kept current by
//     the library after every step; writes into persistent fields are taken in at
//     the next step
//   ExportBgeoFileCUDA(folder, name, pos, col, radius, label, n)
//     -> sf_export_bgeo_device (seepage_flows/include/kiri_utils.h:364-448)
//
// Differences a caller sees: a particle object has device arrays only once a
// CudaSFSystem holds it (the reference copies to the device in the particle
// constructor; nothing in the reference's driver reads them before); the
// searcher and solver classes act on the system's handle, which owns the grid
// and the boundary particles.
//   CudaEmitter / CUDA_SPH_EMITTER_PARAMS             -> sf_set_emitter
//     (emitter/cuda_emitter.cpp:12-97, system/cuda_sf_system.cu:36-58,86-115)
//   ReadBgeoFileForGPU / ReadMultiBgeoFilesForGPU     -> sf_read_bgeo
//     (seepage_flows/include/kiri_utils.h:252-361); the root folder is
//     DB_PBR_PATH when the caller defines it (as the reference's build does),
//     else DbPbrPath(), settable at run time
//   Partio2VTK(vtk, bgeo)                             -> sf_bgeo_to_vtk (kiri_utils.h:54-250)
//   CudaVolumeEmitter::BuildSeepageflowShape(Multi)Volume
//     -> sf_build_seepage_shape_(multi_)volume (emitter/cuda_volume_emitter.cpp:109-175)
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sf_seepage.h"

namespace KIRI {

template <class T>
using SharedPtr = std::shared_ptr<T>;
using Vec_Float = std::vector<float>;
using Vec_Float2 = std::vector<float2>;
using Vec_Float3 = std::vector<float3>;
using Vec_Float4 = std::vector<float4>;
using Vec_SizeT = std::vector<size_t>;
using String = std::string;
using Vec_String = std::vector<std::string>;
static_assert(sizeof(float3) == 12 && sizeof(float2) == 8 && sizeof(float4) == 16 && sizeof(size_t) == 8,
              "the volume builders write the vectors' storage as packed float / uint64 arrays");

enum CudaSphSolverType { SPH_SOLVER = SF_SPH_SOLVER, WCSPH_SOLVER = SF_WCSPH_SOLVER, IISPH_SOLVER = 2 };
enum CudaSeepageflowType { SF = SF_TYPE_SF, MULTI_SF = SF_TYPE_MULTI_SF };
enum SearcherParticleType { SEEPAGE = 0, SEEPAGE_MULTI = 1 };

using CudaSeepageflowParams = sf_params;
using CudaBoundaryParams = sf_boundary_params;

struct CudaSeepageflowAppParams {
    size_t max_num = 100000;
    bool run = false;
    bool run_offline = false;
    int scene_data_idx = 0;
    char bgeo_export_folder[320] = "default";
    bool bgeo_export = false;
};

inline CudaSeepageflowParams CUDA_SEEPAGEFLOW_PARAMS{};
inline CudaBoundaryParams CUDA_BOUNDARY_PARAMS{};
inline CudaSeepageflowAppParams CUDA_SEEPAGEFLOW_APP_PARAMS{};

inline sf_float3 to_sf(float3 v) { return sf_float3{v.x, v.y, v.z}; }
inline sf_int3 to_sf(int3 v) { return sf_int3{v.x, v.y, v.z}; }

// The reference prints failures and carries on (kiri_pbs_pch.cuh:35-50); a
// handle that could not be created cannot carry on, so that case throws.
inline void sf_check(int rc, const char* what) {
    if (rc != SF_OK) {
        std::string msg = std::string(what) + ": " + sf_last_error();
        std::fprintf(stderr, "%s\n", msg.c_str());
        throw std::runtime_error(msg);
    }
}

// int3((highest - lowest) / kernelRadius) (main.cpp:104)
inline int3 GridSize(float3 lo, float3 hi, float h) {
    sf_int3 g{};
    sf_check(sf_grid_size(to_sf(lo), to_sf(hi), h, &g), "sf_grid_size");
    return make_int3(g.x, g.y, g.z);
}

struct BoundaryData {
    Vec_Float3 pos;
    Vec_SizeT label;
};

class CudaBoundaryEmitter {
  public:
    explicit CudaBoundaryEmitter(bool enable = true) : bEnable(enable) {}
    void BuildWorldBoundary(BoundaryData& data, const float3& lowest, const float3& highest, float particleRadius) {
        if (!bEnable) return;
        size_t n = 0;
        sf_check(sf_build_world_boundary(to_sf(lowest), to_sf(highest), particleRadius, 0, nullptr, nullptr, &n),
                 "sf_build_world_boundary");
        std::vector<float> p(3 * n);
        std::vector<uint64_t> l(n);
        sf_check(sf_build_world_boundary(to_sf(lowest), to_sf(highest), particleRadius, n, p.data(), l.data(), &n),
                 "sf_build_world_boundary");
        for (size_t i = 0; i < n; ++i) {
            data.pos.push_back(make_float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
            data.label.push_back((size_t)l[i]);
        }
    }
    void BuildBoundaryShapeVolume(BoundaryData& data, const std::vector<float4>& shape) {
        if (!bEnable) return;
        for (const float4& s : shape) {
            data.pos.push_back(make_float3(s.x, s.y, s.z));
            data.label.push_back(1);
        }
    }

  private:
    bool bEnable;
};

// Root folder of the shape files: <root>bgeo/<Folder>/<Name>.bgeo.
inline std::string& DbPbrPath() {
#ifdef DB_PBR_PATH
    static std::string root = DB_PBR_PATH;
#else
    static std::string root = "./";
#endif
    return root;
}

// ReadBgeoFileForGPU (kiri_utils.h:252-308): (x, y, z, pscale) of every point,
// pscale 0.01 where the file has none, y and z swapped with flip_yz.  A file
// that cannot be read throws (the reference dereferenced a null reader).
inline Vec_Float4 ReadBgeoFileForGPU(String Folder, String Name, bool flip_yz = false) {
    const std::string path = DbPbrPath() + "bgeo/" + Folder + "/" + Name + ".bgeo";
    size_t n = 0;
    sf_check(sf_read_bgeo(path.c_str(), flip_yz ? 1 : 0, nullptr, 0, &n), "sf_read_bgeo");
    Vec_Float4 out(n);
    sf_check(sf_read_bgeo(path.c_str(), flip_yz ? 1 : 0, reinterpret_cast<float*>(out.data()), n, &n),
             "sf_read_bgeo");
    if (n > 0) std::printf("pscale=%g\n", out[0].w);
    return out;
}

// ReadMultiBgeoFilesForGPU (kiri_utils.h:310-361): the files' points in order.
inline Vec_Float4 ReadMultiBgeoFilesForGPU(Vec_String folders, Vec_String file_names, bool flip_yz = false) {
    Vec_Float4 all;
    for (size_t k = 0; k < folders.size() && k < file_names.size(); ++k) {
        Vec_Float4 one = ReadBgeoFileForGPU(folders[k], file_names[k], flip_yz);
        std::printf("Loaded Bgeo File=%s, Number of Particles=%zu\n", file_names[k].c_str(), one.size());
        all.insert(all.end(), one.begin(), one.end());
    }
    return all;
}

// Partio2VTK (kiri_utils.h:54-250), from a bgeo file instead of a partio
// object: a binary legacy VTK unstructured grid of the particles.
inline void Partio2VTK(const std::string& vtk_path, const std::string& bgeo_path) {
    sf_check(sf_bgeo_to_vtk(bgeo_path.c_str(), vtk_path.c_str()), "sf_bgeo_to_vtk");
}

// SeepageflowVolumeData / SeepageflowMultiVolumeData (cuda_volume_emitter.cuh:43-63)
struct SeepageflowVolumeData {
    float sandMinRadius = 0.f;
    Vec_Float3 pos;
    Vec_Float3 col;
    Vec_Float mass;
    Vec_Float radius;
    Vec_SizeT label;
};
struct SeepageflowMultiVolumeData {
    float sandMinRadius = 0.f;
    Vec_Float3 pos;
    Vec_Float3 col;
    Vec_Float mass;
    Vec_Float radius;
    Vec_SizeT label;
    Vec_Float2 amcamcp;
    Vec_Float3 cda0asat;
};

class CudaVolumeEmitter {
  public:
    explicit CudaVolumeEmitter(bool enable = true) : bEnable(enable) {}
    CudaVolumeEmitter(const CudaVolumeEmitter&) = delete;
    CudaVolumeEmitter& operator=(const CudaVolumeEmitter&) = delete;

    void BuildSeepageflowShapeVolume(SeepageflowVolumeData& data, Vec_Float4 shape, float3 color, float sandDensity,
                                     bool offsetY = false, float worldLowestY = 0.f) {
        if (!bEnable) return;
        sf_volume_data d = grow(data, shape.size());
        sf_check(sf_build_seepage_shape_volume(&d, reinterpret_cast<const float*>(shape.data()), shape.size(),
                                               to_sf(color), sandDensity, offsetY ? 1 : 0, worldLowestY),
                 "sf_build_seepage_shape_volume");
        data.sandMinRadius = d.sand_min_radius;
    }
    void BuildSeepageflowShapeMultiVolume(SeepageflowMultiVolumeData& data, Vec_Float4 shape, float3 color,
                                          float sandDensity, float3 cda0asat, float2 amcamcp, bool offsetY = false,
                                          float worldLowestY = 0.f) {
        if (!bEnable) return;
        sf_volume_data d = grow(data, shape.size());
        data.cda0asat.resize(data.pos.size());
        data.amcamcp.resize(data.pos.size());
        d.cda0asat3 = reinterpret_cast<float*>(data.cda0asat.data());
        d.amcamcp2 = reinterpret_cast<float*>(data.amcamcp.data());
        sf_check(sf_build_seepage_shape_multi_volume(&d, reinterpret_cast<const float*>(shape.data()), shape.size(),
                                                     to_sf(color), sandDensity, to_sf(cda0asat), amcamcp.x,
                                                     amcamcp.y, offsetY ? 1 : 0, worldLowestY),
                 "sf_build_seepage_shape_multi_volume");
        data.sandMinRadius = d.sand_min_radius;
    }
    bool GetEmitterStatus() const { return bEnable; }

  private:
    // room for n more particles; the arrays of `data` become the C-ABI's
    template <class V>
    static sf_volume_data grow(V& data, size_t n) {
        const size_t held = data.pos.size();
        data.pos.resize(held + n);
        data.col.resize(held + n);
        data.mass.resize(held + n);
        data.radius.resize(held + n);
        data.label.resize(held + n);
        sf_volume_data d{};
        d.pos3 = reinterpret_cast<float*>(data.pos.data());
        d.col3 = reinterpret_cast<float*>(data.col.data());
        d.label = reinterpret_cast<uint64_t*>(data.label.data());
        d.mass = data.mass.data();
        d.radius = data.radius.data();
        d.size = held;
        d.capacity = held + n;
        d.sand_min_radius = data.sandMinRadius;
        return d;
    }
    bool bEnable;
};
typedef SharedPtr<CudaVolumeEmitter> CudaVolumeEmitterPtr;

class CudaSFSystem;

// Device array handed out by the library (the reference's owning CudaArray,
// cuda_array.cuh:19-61, offers the same Data() / Length()).  The library owns
// the memory; it stays valid while the system lives.
template <class T>
class CudaArray {
  public:
    CudaArray() = default;
    CudaArray(T* p, size_t len) : mData(p), mLen(len) {}
    T* Data() const { return mData; }
    size_t Length() const { return mLen; }

  private:
    T* mData = nullptr;
    size_t mLen = 0;
};

class CudaSFParticles {
  public:
    CudaSFParticles(size_t maxNum, const Vec_Float3& p, const Vec_Float3& col, const Vec_SizeT& label,
                    const Vec_Float& mass, const Vec_Float& radius, const Vec_Float3& cda0asat = {},
                    const Vec_Float2& amcamcp = {})
        : mMax(maxNum), mPos(p), mCol(col), mLabel(label), mMass(mass), mRadius(radius), mCda(cda0asat),
          mAmc(amcamcp) {}
    CudaSFParticles(const CudaSFParticles&) = delete;
    CudaSFParticles& operator=(const CudaSFParticles&) = delete;

    size_t Size() const {
        if (!mHandle) return mPos.size();
        size_t n = 0;
        sf_size(mHandle, &n, nullptr, nullptr);
        return n;
    }
    size_t MaxSize() const { return mMax; }
    // CudaParticles / CudaSFParticles getters (cuda_particles.cuh:57,
    // cuda_sf_particles.cuh:170-190): device arrays in the reference's layout
    // in the current (cell-sorted) order, null until a CudaSFSystem holds the
    // particles.  The library refreshes them after every step and takes what
    // the caller writes into pos, vel, col, label, mass, radius,
    // maxSaturation, cdA0Asat and amcAmcp in at the start of the next step.
    float3* GetPosPtr() const { return (float3*)view(SF_FIELD_POS); }
    float3* GetVelPtr() const { return (float3*)view(SF_FIELD_VEL); }
    float3* GetAccPtr() const { return (float3*)view(SF_FIELD_ACC); }
    float3* GetColPtr() const { return (float3*)view(SF_FIELD_COL); }
    size_t* GetLabelPtr() const { return (size_t*)view(SF_FIELD_LABEL); }
    float* GetRadiusPtr() const { return (float*)view(SF_FIELD_RADIUS); }
    float* GetPressurePtr() const { return (float*)view(SF_FIELD_PRESSURE); }
    float* GetDensityPtr() const { return (float*)view(SF_FIELD_DENSITY); }
    float* GetMassPtr() const { return (float*)view(SF_FIELD_MASS); }
    float* GetVoidagePtr() const { return (float*)view(SF_FIELD_VOIDAGE); }
    float* GetSaturationPtr() const { return (float*)view(SF_FIELD_SATURATION); }
    float* GetMaxSaturationPtr() const { return (float*)view(SF_FIELD_MAX_SATURATION); }
    float3* GetAvgFlowVelPtr() const { return (float3*)view(SF_FIELD_AVG_FLOW_VEL); }
    float3* GetAvgDragForcePtr() const { return (float3*)view(SF_FIELD_AVG_DRAG_FORCE); }
    float3* GetAdhesionForcePtr() const { return (float3*)view(SF_FIELD_ADHESION_FORCE); }
    float3* GetAvgAdhesionForcePtr() const { return (float3*)view(SF_FIELD_AVG_ADHESION_FORCE); }
    float3* GetCdA0AsatPtr() const { return (float3*)view(SF_FIELD_CDA0ASAT); }
    float2* GetAmcAmcpPtr() const { return (float2*)view(SF_FIELD_AMCAMCP); }

    // host copies of any field (packed float3 / float / size_t as the reference stores them)
    Vec_Float3 DownloadFloat3(int field) const {
        Vec_Float3 out(Size());
        if (!out.empty()) sf_check(sf_download(mHandle, field, out.data(), out.size() * sizeof(float3)), "sf_download");
        return out;
    }
    Vec_Float DownloadFloat(int field) const {
        Vec_Float out(Size());
        if (!out.empty()) sf_check(sf_download(mHandle, field, out.data(), out.size() * sizeof(float)), "sf_download");
        return out;
    }
    Vec_SizeT DownloadLabel() const {
        std::vector<uint64_t> l(Size());
        if (!l.empty()) sf_check(sf_download(mHandle, SF_FIELD_LABEL, l.data(), l.size() * 8), "sf_download");
        return Vec_SizeT(l.begin(), l.end());
    }
    sf_handle Handle() const { return mHandle; }

  private:
    friend class CudaSFSystem;
    void* view(int field) const {
        void* p = nullptr;
        if (!mHandle) return nullptr;
        sf_check(sf_particle_view(mHandle, field, &p), "sf_particle_view");
        return p;
    }
    void Attach(sf_handle h, bool multi) {
        const size_t n = mPos.size();
        std::vector<float> pos(3 * n), col(3 * n), cda, amc;
        std::vector<uint64_t> lab(n);
        for (size_t i = 0; i < n; ++i) {
            pos[3 * i] = mPos[i].x; pos[3 * i + 1] = mPos[i].y; pos[3 * i + 2] = mPos[i].z;
            if (i < mCol.size()) { col[3 * i] = mCol[i].x; col[3 * i + 1] = mCol[i].y; col[3 * i + 2] = mCol[i].z; }
            lab[i] = mLabel[i];
        }
        if (multi && mCda.size() == n && mAmc.size() == n) {
            cda.resize(3 * n);
            amc.resize(2 * n);
            for (size_t i = 0; i < n; ++i) {
                cda[3 * i] = mCda[i].x; cda[3 * i + 1] = mCda[i].y; cda[3 * i + 2] = mCda[i].z;
                amc[2 * i] = mAmc[i].x; amc[2 * i + 1] = mAmc[i].y;
            }
        }
        sf_check(sf_upload_particles(h, n, pos.data(), nullptr, col.data(), lab.data(), mMass.data(),
                                     mRadius.empty() ? nullptr : mRadius.data(), cda.empty() ? nullptr : cda.data(),
                                     amc.empty() ? nullptr : amc.data()),
                 "sf_upload_particles");
        mHandle = h;
    }
    size_t mMax;
    Vec_Float3 mPos, mCol;
    Vec_SizeT mLabel;
    Vec_Float mMass, mRadius;
    Vec_Float3 mCda;
    Vec_Float2 mAmc;
    sf_handle mHandle = nullptr;
};
typedef SharedPtr<CudaSFParticles> CudaSFParticlesPtr;

class CudaBoundaryParticles {
  public:
    CudaBoundaryParticles(const Vec_Float3& p, const Vec_SizeT& label = {}) : mPos(p), mLabel(label) {}
    size_t Size() const { return mPos.size(); }
    size_t MaxSize() const { return mPos.size(); }

  private:
    friend class CudaSFSystem;
    Vec_Float3 mPos;
    Vec_SizeT mLabel;
};
typedef SharedPtr<CudaBoundaryParticles> CudaBoundaryParticlesPtr;

// Uniform-grid searchers (cuda_neighbor_searcher.cuh:22-98).  The grid and
// its arrays live in the system's handle: the particle searcher builds them
// with BuildGNSearcher (sf_build_neighbors: hash, stable sort of the particle
// state by cell, cellStart); the boundary searcher's were built when the
// system took the boundary particles (CudaBaseSystem ctor, cuda_base_system.cu:26).
// GetCellStart() / GetGridIdxArray() are the handle's size_t arrays of the
// last build (numCells + 1 entries / one key per particle).
class CudaGNBaseSearcher {
  public:
    CudaGNBaseSearcher(float3 lo, float3 hi, size_t num, float cellSize)
        : mLowestPoint(lo), mHighestPoint(hi), mCellSize(cellSize), mGridSize(GridSize(lo, hi, cellSize)),
          mMaxNumOfParticles(num) {}
    virtual ~CudaGNBaseSearcher() = default;
    CudaGNBaseSearcher(const CudaGNBaseSearcher&) = delete;
    CudaGNBaseSearcher& operator=(const CudaGNBaseSearcher&) = delete;
    float3 GetLowestPoint() const { return mLowestPoint; }
    float3 GetHighestPoint() const { return mHighestPoint; }
    float GetCellSize() const { return mCellSize; }
    int3 GetGridSize() const { return mGridSize; }
    size_t* GetCellStartPtr() const { return GetCellStart().Data(); }
    const CudaArray<size_t>& GetCellStart() const {
        mCellStart = CudaArray<size_t>(index_view(cell_start_field()), num_cells() + 1);
        return mCellStart;
    }

  protected:
    friend class CudaSFSystem;
    virtual int cell_start_field() const = 0;
    size_t num_cells() const { return (size_t)mGridSize.x * (size_t)mGridSize.y * (size_t)mGridSize.z; }
    size_t* index_view(int field) const {
        if (!mHandle) return nullptr;
        void* p = nullptr;
        sf_check(sf_particle_view(mHandle, field, &p), "sf_particle_view");
        return (size_t*)p;
    }
    float3 mLowestPoint, mHighestPoint;
    float mCellSize;
    int3 mGridSize;
    size_t mMaxNumOfParticles;
    sf_handle mHandle = nullptr;  // set by the CudaSFSystem that uses this searcher
    mutable CudaArray<size_t> mCellStart, mGridIdx;
};
class CudaGNSearcher final : public CudaGNBaseSearcher {
  public:
    CudaGNSearcher(float3 lo, float3 hi, size_t num, float cellSize, SearcherParticleType type)
        : CudaGNBaseSearcher(lo, hi, num, cellSize), mType(type) {}
    SearcherParticleType GetSearcherType() const { return mType; }
    // BuildGNSearcher (cuda_neighbor_searcher.cu:36-51): hash, stable sort by
    // cell of the particle state, cellStart.  Failures are printed, as the
    // reference's KIRI_CUCALL does.
    void BuildGNSearcher(const SharedPtr<CudaSFParticles>& particles) {
        sf_handle h = particles ? particles->Handle() : nullptr;
        if (!h) {
            std::printf("BuildGNSearcher: the particles are not held by a CudaSFSystem\n");
            return;
        }
        mHandle = h;
        if (sf_build_neighbors(h) != SF_OK) std::printf("BuildGNSearcher failed: %s\n", sf_last_error());
    }
    size_t* GetGridIdxArrayPtr() const { return GetGridIdxArray().Data(); }
    const CudaArray<size_t>& GetGridIdxArray() const {
        size_t n = 0;
        if (mHandle) sf_size(mHandle, &n, nullptr, nullptr);
        mGridIdx = CudaArray<size_t>(index_view(SF_FIELD_SORTED_KEYS), n);
        return mGridIdx;
    }

  private:
    int cell_start_field() const override { return SF_FIELD_CELL_START; }
    SearcherParticleType mType;
};
class CudaGNBoundarySearcher final : public CudaGNBaseSearcher {
  public:
    using CudaGNBaseSearcher::CudaGNBaseSearcher;

  private:
    int cell_start_field() const override { return SF_FIELD_B_CELL_START; }
};
typedef SharedPtr<CudaGNSearcher> CudaGNSearcherPtr;
typedef SharedPtr<CudaGNBoundarySearcher> CudaGNBoundarySearcherPtr;

class CudaWCSphSFSolver {
  public:
    explicit CudaWCSphSFSolver(size_t num, float negativeScale = 0.f, float timeStepLimitScale = 3.f,
                               float speedOfSound = 100.f)
        : mNum(num), mOpt{negativeScale, timeStepLimitScale, speedOfSound} {}
    float GetSpeedOfSound() const { return mOpt.speed_of_sound; }
    float GetTimeStepLimitScale() const { return mOpt.time_step_limit_scale; }
    void SetSpeedOfSound(float c) { mOpt.speed_of_sound = c > 1e-6f ? c : 1e-6f; }        // cuda_wcsph_sf_solver.cpp:19-22
    void SetTimeStepLimitScale(float s) { mOpt.time_step_limit_scale = s > 0.f ? s : 0.f; } // :29-32
    const sf_solver_options& Options() const { return mOpt; }
    // CudaBaseSolver::GetNumOfSubTimeSteps (cuda_base_solver.cuh:30)
    size_t GetNumOfSubTimeSteps() const { return mNumOfSubTimeSteps; }

    // UpdateSolver (cuda_wcsph_sf_solver.cuh:36-43, cuda_wcsph_sf_solver.cpp:34-179):
    // the solver passes, the CFL sub-step count and the advect on the
    // neighbour structures of the last BuildGNSearcher.  The cell starts are
    // the handle's own (the arguments are the searchers' GetCellStart(), as
    // in the reference's call); params are used for this step; bparams must
    // describe the system's grid.  Failures are printed.
    void UpdateSolver(SharedPtr<CudaSFParticles>& particles, SharedPtr<CudaBoundaryParticles>& /*boundaries*/,
                      const CudaArray<size_t>& /*cellStart*/, const CudaArray<size_t>& /*bCellStart*/,
                      float renderInterval, CudaSeepageflowParams params, CudaBoundaryParams bparams) {
        sf_handle h = particles ? particles->Handle() : nullptr;
        if (!h) {
            std::printf("UpdateSolver: the particles are not held by a CudaSFSystem\n");
            return;
        }
        sf_int3 g{};
        if (sf_grid_size(bparams.lowest_point, bparams.highest_point, bparams.kernel_radius, &g) != SF_OK ||
            g.x != bparams.grid_size.x || g.y != bparams.grid_size.y || g.z != bparams.grid_size.z) {
            std::printf("UpdateSolver: boundary params do not describe a grid\n");
            return;
        }
        if (sf_set_params(h, &params) != SF_OK || sf_update_solver(h, renderInterval, nullptr) != SF_OK) {
            std::printf("UpdateSolver failed: %s\n", sf_last_error());
            return;
        }
        sf_num_sub_steps(h, &mNumOfSubTimeSteps);
    }

  private:
    size_t mNum;
    sf_solver_options mOpt;
    size_t mNumOfSubTimeSteps = 1;
};
typedef SharedPtr<CudaWCSphSFSolver> CudaSphSFSolverPtr;

// Inflow emitter: CudaSphEmitterParams / CUDA_SPH_EMITTER_PARAMS
// (data/cuda_sph_params.h:49-84) and CudaEmitter (emitter/cuda_emitter.cuh:19-85).
// The emitter records its configuration; CudaSFSystem hands it to the library
// (sf_set_emitter) before every UpdateSystem, which samples and emits exactly
// as the reference does, and reads its status back (it switches itself off
// at capacity).
enum CudaSphEmitterType { SQUARE = SF_EMITTER_SQUARE, CIRCLE = SF_EMITTER_CIRCLE, RECTANGLE = SF_EMITTER_RECTANGLE };
struct CudaSphEmitterParams {
    bool enable = false;
    bool run = false;
    float3 emit_pos{0.f, 0.f, 0.f};
    float3 emit_vel{0.f, 0.f, 0.f};
    float3 emit_col{0.f, 0.f, 0.f};
    CudaSphEmitterType emit_type = SQUARE;
    float emit_radius = 0.f;
    float emit_width = 0.f;
    float emit_height = 0.f;
};
inline CudaSphEmitterParams CUDA_SPH_EMITTER_PARAMS{};

class CudaEmitter {
  public:
    CudaEmitter() : CudaEmitter(float3{0.f, 0.f, 0.f}, float3{1.f, 0.f, 0.f}, false) {}
    CudaEmitter(float3 emitPosition, float3 emitVelocity, bool enable)
        : mPos(emitPosition), mVel(emitVelocity), bEnable(enable) {}
    CudaEmitter(const CudaEmitter&) = delete;
    CudaEmitter& operator=(const CudaEmitter&) = delete;

    void SetEmitterStatus(bool enable) { bEnable = enable; }
    void UpdateEmitterVelocity(float3 emitVelocity) { mVel = emitVelocity; }
    // the library samples with the water particle radius, as the CudaSFSystem
    // ctor calls these (cuda_sf_system.cu:36-58)
    void BuildSquareEmitter(float, float emitterRadius) { build(SF_EMITTER_SQUARE, emitterRadius, 0.f, 0.f); }
    void BuildCircleEmitter(float, float emitterRadius) { build(SF_EMITTER_CIRCLE, emitterRadius, 0.f, 0.f); }
    void BuildRectangleEmitter(float, float emitterWidth, float emitterHeight) {
        build(SF_EMITTER_RECTANGLE, 0.f, emitterWidth, emitterHeight);
    }
    bool GetEmitterStatus() const { return bEnable; }
    size_t GetNumOfEmitterPoints() const { return mPoints; }
    float3 GetEmitterPosition() const { return mPos; }
    float3 GetEmitterVelocity() const { return mVel; }

    // CudaSFSystem side: push the configuration, pull the status
    void Push(sf_handle h) const {
        sf_emitter_params e{};
        e.enable = (bEnable && bBuild) ? 1 : 0;
        e.run = CUDA_SPH_EMITTER_PARAMS.run ? 1 : 0;
        e.emit_pos = sf_float3{mPos.x, mPos.y, mPos.z};
        e.emit_vel = sf_float3{mVel.x, mVel.y, mVel.z};
        const float3 c = CUDA_SPH_EMITTER_PARAMS.emit_col;
        e.emit_col = sf_float3{c.x, c.y, c.z};
        e.emit_type = mType;
        e.emit_radius = mRadius;
        e.emit_width = mWidth;
        e.emit_height = mHeight;
        if (sf_set_emitter(h, &e) != SF_OK) throw std::runtime_error(std::string("sf_set_emitter: ") + sf_last_error());
    }
    void Pull(sf_handle h) {
        int enabled = 0;
        size_t points = 0;
        sf_get_emitter(h, &enabled, &points, nullptr);
        if (bEnable && bBuild && !enabled) bEnable = false;  // switched off at capacity
        mPoints = points;
    }

  private:
    void build(int type, float radius, float width, float height) {
        mType = type;
        mRadius = radius;
        mWidth = width;
        mHeight = height;
        bBuild = true;
    }
    float3 mPos, mVel;
    bool bEnable, bBuild = false;
    int mType = SF_EMITTER_SQUARE;
    float mRadius = 0.f, mWidth = 0.f, mHeight = 0.f;
    size_t mPoints = 0;
};
typedef SharedPtr<CudaEmitter> CudaEmitterPtr;

class CudaSFSystem {
  public:
    CudaSFSystem(CudaSFParticlesPtr& particles, CudaBoundaryParticlesPtr& boundaryParticles,
                 CudaSphSFSolverPtr& solver, CudaGNSearcherPtr& searcher, CudaGNBoundarySearcherPtr& boundarySearcher,
                 CudaEmitterPtr& emitter, bool adaptiveSubTimeStep = false)
        : mParticles(std::move(particles)), mBoundaries(std::move(boundaryParticles)), mSolver(std::move(solver)),
          mSearcher(std::move(searcher)), mBoundarySearcher(std::move(boundarySearcher)),
          mEmitter(emitter), bAdaptiveSubTimeStep(adaptiveSubTimeStep) {
        const bool multi = CUDA_SEEPAGEFLOW_PARAMS.sf_type == SF_TYPE_MULTI_SF;
        sf_check(sf_create(&CUDA_SEEPAGEFLOW_PARAMS, &CUDA_BOUNDARY_PARAMS, &mSolver->Options(),
                           mParticles->MaxSize(), -1, &mHandle),
                 "sf_create");
        mParticles->Attach(mHandle, multi);
        std::vector<float> bp(3 * mBoundaries->mPos.size());
        std::vector<uint64_t> bl(mBoundaries->mPos.size(), 0);
        for (size_t i = 0; i < mBoundaries->mPos.size(); ++i) {
            bp[3 * i] = mBoundaries->mPos[i].x;
            bp[3 * i + 1] = mBoundaries->mPos[i].y;
            bp[3 * i + 2] = mBoundaries->mPos[i].z;
            if (i < mBoundaries->mLabel.size()) bl[i] = mBoundaries->mLabel[i];
        }
        sf_check(sf_upload_boundary(mHandle, mBoundaries->mPos.size(), bp.data(), bl.data()), "sf_upload_boundary");
        if (mSearcher) mSearcher->mHandle = mHandle;
        if (mBoundarySearcher) mBoundarySearcher->mHandle = mHandle;
        if (mEmitter && CUDA_SPH_EMITTER_PARAMS.enable) {  // cuda_sf_system.cu:36-58
            const float r = CUDA_SEEPAGEFLOW_PARAMS.sph_particle_radius;
            switch (CUDA_SPH_EMITTER_PARAMS.emit_type) {
            case SQUARE: mEmitter->BuildSquareEmitter(r, CUDA_SPH_EMITTER_PARAMS.emit_radius); break;
            case CIRCLE: mEmitter->BuildCircleEmitter(r, CUDA_SPH_EMITTER_PARAMS.emit_radius); break;
            case RECTANGLE:
                mEmitter->BuildRectangleEmitter(r, CUDA_SPH_EMITTER_PARAMS.emit_width,
                                                CUDA_SPH_EMITTER_PARAMS.emit_height);
                break;
            }
        }
        if (mEmitter) {
            mEmitter->Push(mHandle);
            mEmitter->Pull(mHandle);
        }
    }
    CudaSFSystem(const CudaSFSystem&) = delete;
    CudaSFSystem& operator=(const CudaSFSystem&) = delete;
    ~CudaSFSystem() { sf_destroy(mHandle); }

    // CudaBaseSystem::UpdateSystem (cuda_base_system.cu:29-56): the global
    // parameter structs are passed by value to every step, failures are
    // printed and the step returns its measured time in ms.
    float UpdateSystem(float renderInterval) {
        float ms = 0.f;
        if (mEmitter) mEmitter->Push(mHandle);  // status, velocity and the run switch may have changed
        if (sf_set_params(mHandle, &CUDA_SEEPAGEFLOW_PARAMS) != SF_OK ||
            sf_update_system(mHandle, renderInterval, &ms) != SF_OK)
            std::printf("UpdateSystem failed: %s\n", sf_last_error());
        if (mEmitter) mEmitter->Pull(mHandle);
        return ms;
    }
    // One frame of main.cpp's Update() loop (main.cpp:235-259) run by the
    // library (sf_update_frame): the same state as calling UpdateSystem from
    // the loop, with the loop's bookkeeping on the device.  Returns the
    // number of sub-steps.
    size_t UpdateFrame(float renderInterval) {
        size_t k = 0;
        float ms = 0.f;
        if (mEmitter) mEmitter->Push(mHandle);
        if (sf_set_params(mHandle, &CUDA_SEEPAGEFLOW_PARAMS) != SF_OK ||
            sf_update_frame(mHandle, renderInterval, bAdaptiveSubTimeStep ? 1 : 0, &k, &ms) != SF_OK)
            std::printf("UpdateFrame failed: %s\n", sf_last_error());
        if (mEmitter) mEmitter->Pull(mHandle);
        return k;
    }
    size_t GetNumOfSubTimeSteps() const {
        size_t k = 1;
        sf_num_sub_steps(mHandle, &k);
        return k;
    }
    bool GetAdaptiveSubTimeStep() const { return bAdaptiveSubTimeStep; }
    size_t NumOfParticles() const { return mParticles->Size(); }
    size_t MaxNumOfParticles() const { return mParticles->MaxSize(); }
    CudaSFParticlesPtr GetSFParticles() const { return mParticles; }
    CudaSphSFSolverPtr GetSolver() const { return mSolver; }
    CudaGNSearcherPtr GetSearcher() const { return mSearcher; }
    CudaGNBoundarySearcherPtr GetBoundarySearcher() const { return mBoundarySearcher; }
    CudaBoundaryParticlesPtr GetBoundaries() const { return mBoundaries; }
    sf_handle Handle() const { return mHandle; }

  private:
    CudaSFParticlesPtr mParticles;
    CudaBoundaryParticlesPtr mBoundaries;
    CudaSphSFSolverPtr mSolver;
    CudaGNSearcherPtr mSearcher;
    CudaGNBoundarySearcherPtr mBoundarySearcher;
    CudaEmitterPtr mEmitter;
    bool bAdaptiveSubTimeStep;
    sf_handle mHandle = nullptr;
};
typedef SharedPtr<CudaSFSystem> CudaSFSystemPtr;

// UInt2Str4Digit (kiri_utils.h:36-42)
inline std::string UInt2Str4Digit(unsigned input) {
    char output[16];
    std::snprintf(output, sizeof(output), "%04u", input);
    return std::string(output);
}

// ExportBgeoFileCUDA (kiri_utils.h:364-448): folder/file_name.bgeo with
// position, Cd, pscale and label of n particles held in device arrays of the
// reference's layout (the particles' Get*Ptr() views).
inline void ExportBgeoFileCUDA(String folder, String file_name, const float3* pos, const float3* col,
                               const float* radius, const size_t* label, const size_t particlesNum) {
    if (sf_export_bgeo_device(folder.c_str(), file_name.c_str(), (const float*)pos, (const float*)col, radius,
                              (const uint64_t*)label, particlesNum) != SF_OK)
        std::printf("ExportBgeoFileCUDA failed: %s\n", sf_last_error());
}

// The same file straight from the particle state (no views needed).
inline void ExportBgeoFileCUDA(const std::string& folder, const std::string& file_name,
                               const CudaSFParticlesPtr& particles) {
    if (sf_export_bgeo(particles->Handle(), folder.c_str(), file_name.c_str()) != SF_OK)
        std::printf("ExportBgeoFileCUDA failed: %s\n", sf_last_error());
}

}  // namespace KIRI
