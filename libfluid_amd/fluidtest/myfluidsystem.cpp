// myfluidsystem.cpp -- double dam break (test/src/myfluidsystem.cpp:6-30).
#include "myfluidsystem.h"

void MyFluidSystem::setupDoubleDamBreak(const uint &_numPoints, const uint &_res, const float &_adhesion,
                                        const float &_viscosity, const float &_stension,
                                        const float &_vdamp) noexcept {
    init(_numPoints, _res);
    setAdhesion(_adhesion);
    setViscosity(_viscosity);
    setSurfaceTension(_stension);
    setVDamp(_vdamp);
    createFluidBlock(0, _numPoints / 2, 0.0f, 0.0f, 0.0f, 0.33f, 0.9f, 0.33f);
    createFluidBlock(_numPoints / 2, _numPoints - _numPoints / 2, 0.66f, 0.0f, 0.66f, 1.0f, 0.9f, 1.0f);
}
