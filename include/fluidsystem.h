/*
 * fluidsystem.h -- the FluidSystem C++ host API on MI355X.
 *
 * Same class name, members, signatures and defaults as the reference's
 * libfluid/include/fluidsystem.h:16-126, so callers (and subclasses such as
 * the double-dam-break scene) compile unchanged.  It is implemented over the
 * C ABI in lf_abi.h; device memory is owned by an lf_scene handle instead of
 * thrust::device_vector members, and parameters live per scene instead of in
 * one global __constant__ symbol, so several FluidSystem objects can coexist.
 *
 * float3: a plain {x,y,z} struct unless HIP's vector types were included
 * first (both are 12-byte, standard-layout).
 */
#ifndef FLUIDSYSTEM_H
#define FLUIDSYSTEM_H

#include <sys/types.h>

#include <vector>

#include "fluiddefines.h"
#include "lf_abi.h"

#if !defined(HIP_INCLUDE_HIP_AMD_DETAIL_HIP_VECTOR_TYPES_H) && !defined(LF_FLOAT3_DEFINED)
#define LF_FLOAT3_DEFINED
struct float3 {
    float x, y, z;
};
inline float3 make_float3(float x, float y, float z) {
    float3 v;
    v.x = x;
    v.y = y;
    v.z = z;
    return v;
}
#endif
static_assert(sizeof(float3) == 3 * sizeof(float), "float3 must be three packed floats");

class FluidSystem {
public:
    typedef enum { STATIC = 1, DYNAMIC = 2 } ParticleState;

    /// Construct an empty fluid system
    FluidSystem();

    /// Destruct our fluid system
    virtual ~FluidSystem();

    /// Initialise a relatively standard dambreak simulation (fluidsystem.cu:132-147)
    virtual void setup(const uint &_numPoints, const uint &_res);

    /// Initialise from data generated externally (fluidsystem.cu:179-189)
    void setupFromData(const std::vector<float3> &points, const std::vector<float3> &velocity,
                       const uint &_res = 32);

    /// Copy positions, density colour and velocity to the host, in the current
    /// cell-sorted order (fluidsystem.cu:154-173).  The reference copies into
    /// vectors the caller sized to the particle count (main.cpp:119-121);
    /// here a vector shorter than that is grown to it first (longer ones keep
    /// their size and tail).
    void exportToData(std::vector<float3> &points, std::vector<float3> &colour, std::vector<float3> &velocity);

    /// Retrieve the current particle scale
    float getPscale() const;

    /// Progress the simulation by dt in `substeps` substeps (fluidsystem.cu:273-399)
    void advance(const float &dt = DEFAULT_TIMESTEP, const uint &substeps = 10);

    void setRestDensity(const float &_restDensity);
    void setParticleSize(const float &_particleSize);
    void setGasConstant(const float &_gasConstant);
    void setLimits(const float &_velLimit, const float &_accLimit);
    void setXSPHScale(const float &_xsphScale);
    void setSurfaceTension(const float &_surfaceTension);
    void setViscosity(const float &_viscosity);
    void setAdhesion(const float &_adhesion);
    void setVDamp(const float &_vdamp);
    void setGamma(const float &_gamma);
    void setMass(const float &_mass);

    /* ---- additions (not in the reference) ---- */
    /// Scene with an explicit box, grid, h and gravity; points/velocity by particle id.
    void setupScene(const lf_params &params, const std::vector<float3> &points,
                    const std::vector<float3> &velocity);
    void setGravity(const float &x, const float &y, const float &z);
    /// Copy the state in particle-id order.
    void exportById(std::vector<float3> &points, std::vector<float3> &colour, std::vector<float3> &velocity);
    uint numParticles() const { return m_numPoints; }
    const lf_params &params() const { return m_params; }
    lf_scene *scene() { return m_scene; }
    /// Status of the last API call (LF_OK or an LF_ERR_* code).
    int lastStatus() const { return m_status; }

protected:
    /// Keep track of whether the simulation is ready to start
    bool m_isInit;

    /// Allocate the scene for _numPoints particles on a res^3 unit-cube grid (fluidsystem.cu:196-265)
    void init(const uint &_numPoints, const uint &_res);

    /// Release the scene
    void clear();

    /// Construct a block of fluid particles (fluidsystem.cu:66-127)
    void createFluidBlock(const uint &firstPtIdx, const uint &numPoints, const float &min_x, const float &min_y,
                          const float &min_z, const float &max_x, const float &max_y, const float &max_z);

private:
    bool check(int rc, const char *what);
    void debugPrint() const;

    lf_scene *m_scene;
    lf_params m_params;
    bool m_dirty;
    uint m_numPoints;
    int m_status;
};

#endif  // FLUIDSYSTEM_H
