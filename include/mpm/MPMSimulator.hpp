// MPMSimulator.hpp -- C++ mirror of the reference simulator's interface over
// the C ABI of libmpm_hip (header-only).
//
// Reference: JerryYan97/MPM_CUDA src/simulator/MPMSimulator.cuh:19-102
// (ObjInitInfo, MPMSimulator), src/simulator/material/Elasiticity.cuh:10-46
// (MaterialType, Material), src/utilities/FilesIO.h:10 (OutputBego).
// Same names, argument meaning and call order; errors throw
// std::runtime_error carrying mpm_last_error() where the reference prints
// and calls exit(1).  glm::vec3 is replaced by mpm::vec3 (three floats).
#pragma once

#include <array>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mpm.h"

namespace mpm {

enum MaterialType { JELLO = MPM_JELLO, SNOW = MPM_SNOW, WATER = MPM_WATER, SAND = 3 };

inline void throw_if(int rc)
{
    if (rc != MPM_OK) throw std::runtime_error(std::string("mpm: ") + mpm_last_error());
}

// Material(E, nu, rho, type): Lame constants from Young's modulus and
// Poisson's ratio (Elasiticity.cuh:30-45).
class Material {
public:
    double mLambda = 0.0, mMu = 0.0, mDensity = 0.0;
    MaterialType mType = JELLO;
    mpm_material c{};

    Material() { std::memset(&c, 0, sizeof c); }
    Material(double iYM, double iPR, double iDensity, MaterialType iType) : mDensity(iDensity), mType(iType)
    {
        if (iType == SAND) throw std::runtime_error("mpm: SAND has no constitutive model");
        throw_if(mpm_material_init(&c, iYM, iPR, iDensity, (int)iType));
        mMu = iYM / (2 * (1 + iPR));
        mLambda = iYM * iPR / ((1 + iPR) * (1 - 2 * iPR));
    }
};

struct vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
    vec3() = default;
    explicit vec3(float s) : x(s), y(s), z(s) {}
    vec3(float a, float b, float d) : x(a), y(b), z(d) {}
};

// ObjInitInfo (MPMSimulator.cuh:19-27); `shape` selects a procedural shape
// instead of the OBJ file (MPM_SHAPE_BOX / MPM_SHAPE_SPHERE with shapeParam).
struct ObjInitInfo {
    std::array<double, 3> initVel{0.0, 0.0, 0.0};
    vec3 initTranslation{0.f, 0.f, 0.f};
    vec3 initRotationAxis{1.f, 0.f, 0.f};
    vec3 initScale{1.f, 1.f, 1.f};
    float initRotationDegree = 0.f;
    std::string objPath;
    Material mMaterial;
    int shape = MPM_SHAPE_MESH;
    std::array<double, 6> shapeParam{0, 0, 0, 1, 1, 1};
    long long maxParticles = 0;
};

// Grid (MPMSimulator.cuh:54-64): the geometry the reference's viewer reads
// (main.cpp:171-176).  Node data stays on the device; read it with
// mpm_read_grid on handle().
struct Grid {
    double h = 0.0;
    unsigned int nodeNumDimX = 0, nodeNumDimY = 0, nodeNumDimZ = 0;
    std::array<double, 3> originCorner{0.0, 0.0, 0.0};
};

// ParticleGroup (MPMSimulator.cuh:30-52): counts and per-particle constants;
// the particle planes stay on the device (mpm_read_particles).
struct ParticleGroup {
    unsigned int particleNum = 0;
    double mParticleMass = 0.0;
    double mParticleVolume = 0.0;
    Material mMaterial;
};

class MPMSimulator {
public:
    double max_dt = 0.0;        // MPMSimulator.cuh:70
    double adp_dt = 0.0;        // :85 (the reference never adapts it; with a dt controller the last dt)
    double current_time = 0.0;  // :86
    int current_frame = 0;      // :87
    std::vector<ParticleGroup> mParticlesGroupsVec;  // :89
    Grid mGrid{};                                     // :90

    // MPMSimulator.cuh:92-96.  `cfg` (optional) carries the settings the
    // reference hard-codes: gravity, wall band, execution form, ...
    MPMSimulator(double gap, double dt, unsigned nodeNumDimX, unsigned nodeNumDimY, unsigned nodeNumDimZ,
                 unsigned particleNumPerCell, std::vector<ObjInitInfo> &objInitInfoVec,
                 const mpm_config *cfg = nullptr)
    {
        mpm_config c;
        if (cfg) c = *cfg; else mpm_config_default(&c);
        c.h = gap; c.dt = dt;
        c.nx = (int)nodeNumDimX; c.ny = (int)nodeNumDimY; c.nz = (int)nodeNumDimZ;
        c.ppc = (int)particleNumPerCell;
        max_dt = adp_dt = dt;
        adaptive_ = c.dt_mode != MPM_DT_FIXED;
        mGrid.h = gap;
        mGrid.nodeNumDimX = nodeNumDimX; mGrid.nodeNumDimY = nodeNumDimY; mGrid.nodeNumDimZ = nodeNumDimZ;
        for (int k = 0; k < 3; ++k) mGrid.originCorner[k] = c.origin[k];
        throw_if(mpm_create(&c, &sim_));
        for (auto &info : objInitInfoVec) {
            mpm_object o;
            std::memset(&o, 0, sizeof o);
            o.shape = info.shape;
            o.obj_path = info.objPath.c_str();
            for (int k = 0; k < 6; ++k) o.shape_param[k] = info.shapeParam[k];
            for (int k = 0; k < 3; ++k) o.init_vel[k] = info.initVel[k];
            o.translation[0] = info.initTranslation.x; o.translation[1] = info.initTranslation.y;
            o.translation[2] = info.initTranslation.z;
            o.scale[0] = info.initScale.x; o.scale[1] = info.initScale.y; o.scale[2] = info.initScale.z;
            o.rotation_axis[0] = info.initRotationAxis.x; o.rotation_axis[1] = info.initRotationAxis.y;
            o.rotation_axis[2] = info.initRotationAxis.z;
            o.rotation_deg = info.initRotationDegree;
            o.material = info.mMaterial.c;
            o.max_particles = info.maxParticles;
            int32_t gidx = -1;
            int rc = mpm_add_object(sim_, &o, &gidx);
            if (rc != MPM_OK) {
                std::string msg = mpm_last_error();
                mpm_destroy(sim_);
                sim_ = nullptr;
                throw std::runtime_error("mpm: " + msg);
            }
            ParticleGroup pg;
            pg.particleNum = (unsigned int)mpm_group_size(sim_, gidx);
            pg.mMaterial = info.mMaterial;
            const double vol = c.volume_mode == MPM_VOLUME_PPC ? gap * gap * gap / c.ppc : gap * gap * gap;
            pg.mParticleVolume = vol;                        // MPMSimulator.cpp:263
            pg.mParticleMass = vol * info.mMaterial.mDensity; // MPMSimulator.cpp:264
            mParticlesGroupsVec.push_back(pg);
        }
    }
    MPMSimulator(const MPMSimulator &) = delete;
    MPMSimulator &operator=(const MPMSimulator &) = delete;
    ~MPMSimulator() { mpm_destroy(sim_); }

    // MPMCUDA.cu:793-925
    void step()
    {
        throw_if(mpm_step(sim_, adaptive_ ? max_dt : adp_dt));
        if (adaptive_) adp_dt = mpm_last_dt(sim_);
        current_time = mpm_current_time(sim_);
        current_frame = (int)mpm_current_frame(sim_);
    }

    int totalParticlesNum() const { return (int)mpm_total_particles(sim_); }

    // MPMSimulator.cpp:362-371
    void getGLParticlesPos(std::vector<float> &oPosVec)
    {
        oPosVec.resize((size_t)totalParticlesNum() * 3);
        throw_if(mpm_get_positions(sim_, oPosVec.data()));
    }

    mpm_sim *handle() { return sim_; }

private:
    mpm_sim *sim_ = nullptr;
    bool adaptive_ = false;
};

// FilesIO.cpp:6-23
inline void OutputBego(const std::string &fileNamePath, const std::vector<float> &outPosData)
{
    throw_if(mpm_write_bgeo(fileNamePath.c_str(), outPosData.data(), (int64_t)(outPosData.size() / 3)));
}

}  // namespace mpm
