/*
 * fluiddefines.h -- compile-time defaults of the FluidSystem API.
 * Same macro names and values as the reference's
 * common/include/fluiddefines.h:4-25, so callers written against it compile
 * unchanged.
 */
#ifndef FLUIDDEFINES_H
#define FLUIDDEFINES_H

#include <limits.h>

#define DEFAULT_TIMESTEP 0.01f
#define DEFAULT_REST_DENSITY 1.0f
#define DEFAULT_GAS_CONSTANT 1.0f

#define MAX_BOUNDING_PLANES 8

/* Cell key of a particle outside the grid. */
#define NULL_HASH UINT_MAX

#ifndef CUDART_PI_F
#define CUDART_PI_F 3.141592654f
#endif

#define DEFAULT_SURFACE_TENSION_FORCE 1.0f
#define DEFAULT_VISCOSITY_FORCE 0.1f
#define DEFAULT_ADHESION_FORCE 1.0f

#define DEFAULT_BOUNDARY_DENSITY 2600

#endif /* FLUIDDEFINES_H */
