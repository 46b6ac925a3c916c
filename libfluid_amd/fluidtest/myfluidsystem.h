// myfluidsystem.h -- the double-dam-break scene as a FluidSystem subclass,
// same interface as the reference's test/src/myfluidsystem.h:9-25.
#ifndef MYFLUIDSYSTEM_H
#define MYFLUIDSYSTEM_H

#include "fluidsystem.h"

class MyFluidSystem : public FluidSystem {
public:
    MyFluidSystem() : FluidSystem() {}
    ~MyFluidSystem() {}

    /// Two blocks in opposite corners of the unit cube (test/src/myfluidsystem.cpp:6-30)
    void setupDoubleDamBreak(const uint &_numPoints = 10000, const uint &_res = 32, const float &_adhesion = 0.1f,
                             const float &_viscosity = 0.2f, const float &_stension = 1.0f,
                             const float &_vdamp = 0.5f) noexcept;
};

#endif  // MYFLUIDSYSTEM_H
