// fluidsystem.cpp -- FluidSystem (include/fluidsystem.h) over the C ABI.
// Host C++ only: compiles with g++ or hipcc, no HIP headers needed.

#include "fluidsystem.h"

#include <cmath>
#include <iostream>

FluidSystem::FluidSystem() : m_isInit(false), m_scene(nullptr), m_dirty(true), m_numPoints(0), m_status(LF_OK) {
    lf_params_default(&m_params, 1);
}

FluidSystem::~FluidSystem() { clear(); }

bool FluidSystem::check(int rc, const char *what) {
    m_status = rc;
    if (rc != LF_OK) {
        std::cerr << "FluidSystem::" << what << "() failed (" << rc << "): " << lf_last_error() << "\n";
        return false;
    }
    return true;
}

void FluidSystem::clear() {
    if (m_scene) lf_destroy(m_scene);
    m_scene = nullptr;
    m_isInit = false;
    m_numPoints = 0;
}

// FluidSystem::init (fluidsystem.cu:196-265): defaults, then allocation.
void FluidSystem::init(const uint &_numPoints, const uint &_res) {
    clear();
    lf_params p;
    if (!check(lf_params_default(&p, _res), "init")) return;
    // Keep gravity chosen before init (the reference fixes it at construction).
    for (int a = 0; a < 3; ++a) p.gravity[a] = m_params.gravity[a];
    m_params = p;
    if (!check(lf_create(&m_params, _numPoints, &m_scene), "init")) {
        m_scene = nullptr;
        return;
    }
    m_numPoints = _numPoints;
    m_dirty = false;
    m_isInit = true;
    debugPrint();
}

void FluidSystem::setup(const uint &_numPoints, const uint &_res) {
    init(_numPoints, _res);
    setAdhesion(0.0f);
    setViscosity(0.0f);
    setSurfaceTension(0.0f);
    setVDamp(0.5f);
    createFluidBlock(0, _numPoints, 0.1f, 0.1f, 0.1f, 0.9f, 0.9f, 0.9f);
}

void FluidSystem::setupFromData(const std::vector<float3> &points, const std::vector<float3> &velocity,
                                const uint &_res) {
    init((uint)points.size(), _res);
    if (!m_isInit) return;
    if (velocity.size() < points.size()) {
        std::cerr << "FluidSystem::setupFromData() - " << velocity.size() << " velocities for " << points.size()
                  << " points\n";
        m_status = LF_ERR_INVALID;
        return;
    }
    // Like the reference, mass and particle size keep init's defaults of 1.
    check(lf_upload(m_scene, 0, (uint32_t)points.size(), &points[0].x, &velocity[0].x, nullptr), "setupFromData");
}

void FluidSystem::setupScene(const lf_params &params, const std::vector<float3> &points,
                             const std::vector<float3> &velocity) {
    clear();
    m_params = params;
    if (!check(lf_create(&m_params, (uint32_t)points.size(), &m_scene), "setupScene")) {
        m_scene = nullptr;
        return;
    }
    m_numPoints = (uint)points.size();
    m_isInit = true;
    m_dirty = false;
    const float *vel = velocity.size() >= points.size() && !velocity.empty() ? &velocity[0].x : nullptr;
    check(lf_upload(m_scene, 0, m_numPoints, &points[0].x, vel, nullptr), "setupScene");
}

void FluidSystem::createFluidBlock(const uint &firstPtIdx, const uint &numPoints, const float &min_x,
                                   const float &min_y, const float &min_z, const float &max_x, const float &max_y,
                                   const float &max_z) {
    if (!m_isInit) return;
    if ((firstPtIdx + numPoints) > m_numPoints) {
        std::cerr << "FluidSystem::createFluidBlock() - attempting to make " << numPoints
                  << " points but you only have enough space for " << m_numPoints - firstPtIdx << "\n!";
        return;
    }
    if (numPoints == 0) return;
    // The block's mass follows the current rest density (fluidsystem.cu:105).
    if (m_dirty) {
        if (!check(lf_set_params(m_scene, &m_params), "createFluidBlock")) return;
        m_dirty = false;
    }
    const float mn[3] = {min_x, min_y, min_z}, mx[3] = {max_x, max_y, max_z};
    uint32_t written = 0;
    float mass = 0.0f, psize = 0.0f;
    // The lattice is generated on the device (the reference writes it point
    // by point from the host, fluidsystem.cu:113-126).
    if (!check(lf_setup_fluid_block(m_scene, firstPtIdx, numPoints, mn, mx, &written, &mass, &psize),
               "createFluidBlock"))
        return;
    // The block sets the (single) particle mass and size (fluidsystem.cu:105,108).
    setMass(mass);
    setParticleSize(psize);
}

void FluidSystem::exportToData(std::vector<float3> &points, std::vector<float3> &colour,
                               std::vector<float3> &velocity) {
    if (!m_isInit) return;
    if (points.size() < m_numPoints) points.resize(m_numPoints);
    if (colour.size() < m_numPoints) colour.resize(m_numPoints);
    if (velocity.size() < m_numPoints) velocity.resize(m_numPoints);
    check(lf_download(m_scene, LF_ORDER_CELL, &points[0].x, &colour[0].x, &velocity[0].x, nullptr, nullptr),
          "exportToData");
}

void FluidSystem::exportById(std::vector<float3> &points, std::vector<float3> &colour, std::vector<float3> &velocity) {
    if (!m_isInit) return;
    points.resize(m_numPoints);
    colour.resize(m_numPoints);
    velocity.resize(m_numPoints);
    check(lf_download(m_scene, LF_ORDER_ID, &points[0].x, &colour[0].x, &velocity[0].x, nullptr, nullptr),
          "exportById");
}

float FluidSystem::getPscale() const { return m_params.particle_size; }

void FluidSystem::advance(const float &dt, const uint &substeps) {
    if (!m_isInit) return;
    if (m_dirty) {
        if (!check(lf_set_params(m_scene, &m_params), "advance")) return;
        m_dirty = false;
    }
    if (!check(lf_step(m_scene, dt, substeps), "advance")) return;
    check(lf_sync(m_scene), "advance");
}

#define LF_SETTER(name, field)                  \
    void FluidSystem::name(const float &v) {    \
        if (m_params.field != v) m_dirty = true; \
        m_params.field = v;                     \
    }
LF_SETTER(setRestDensity, rest_density)
LF_SETTER(setParticleSize, particle_size)
LF_SETTER(setGasConstant, gas_constant)
LF_SETTER(setXSPHScale, xsph_scale)
LF_SETTER(setSurfaceTension, surface_tension)
LF_SETTER(setViscosity, viscosity)
LF_SETTER(setAdhesion, adhesion)
LF_SETTER(setVDamp, vdamp)
LF_SETTER(setGamma, gamma)
LF_SETTER(setMass, mass)
#undef LF_SETTER

void FluidSystem::setLimits(const float &_velLimit, const float &_accLimit) {
    m_params.vel_limit = _velLimit;
    m_params.acc_limit = _accLimit;
    m_dirty = true;
}

void FluidSystem::setGravity(const float &x, const float &y, const float &z) {
    m_params.gravity[0] = x;
    m_params.gravity[1] = y;
    m_params.gravity[2] = z;
    m_dirty = true;
}

// FluidParams::debugPrint (fluidparams.cu:232-240), printed by init.
void FluidSystem::debugPrint() const {
    const double pi = 3.141592654, h = m_params.h;
    const double poly6 = 315.0 / (64.0 * pi * std::pow(h, 9.0));
    const double grad = -945.0 / (32.0 * pi * std::pow(h, 9.0));
    std::cout << "FluidParams:\n";
    std::cout << "res: " << m_params.grid[0] << "x" << m_params.grid[1] << "x" << m_params.grid[2]
              << ", numCells: " << (unsigned long long)m_params.grid[0] * m_params.grid[1] * m_params.grid[2]
              << ", numParticles: " << m_numPoints << "\n";
    std::cout << "restDensity: " << m_params.rest_density << ", gasConstant: " << m_params.gas_constant
              << ", mass: " << m_params.mass << " invMass: " << 1.0f / m_params.mass << "\n";
    std::cout << "h: " << m_params.h << ", m_vdamp: " << m_params.vdamp
              << ", particleSize: " << m_params.particle_size << "\n";
    std::cout << "Adhesion: " << m_params.adhesion << ", Viscosity: " << m_params.viscosity
              << ", SurfaceTension: " << m_params.surface_tension << "\n";
    std::cout << "kernel coefficients: " << (float)poly6 << ", " << (float)grad << ", " << (float)grad << "\n";
}
