// lf_host.cpp -- host-only parts of the C ABI: error reporting, parameter
// defaults and validation, scene-input generators and the .geo writer.
// Compiled with -ffp-contract=off so the FP32 lattice loops round exactly as
// written.

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "lf_internal.h"

namespace lf {

static thread_local std::string g_error;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

void clear_error() { g_error.clear(); }

// SplitMix64 (Steele, Lea, Flood 2014): the seeded generator for every
// synthetic scene, so the GPU path and the oracle start from identical bytes.
static inline uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void block_axes(uint32_t n, const float mn[3], const float mx[3], float rest_density, std::vector<float> axes[3],
                float *mass, float *particle_size) {
    const float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    const float volume = dx * dy * dz;
    const float part_volume = volume / float(n);
    const float approx = std::cbrt(part_volume);
    const float rx = dx / approx + 0.5f, ry = dy / approx + 0.5f, rz = dz / approx + 0.5f;
    const float diam = std::min(std::min(dx / rx, dy / ry), dz / rz);
    for (int a = 0; a < 3; ++a) {
        axes[a].clear();
        // The loops stop at n points in total, so no axis needs more than n values.
        for (float v = mn[a] + 0.5f * diam; v <= mx[a] && axes[a].size() < n; v += diam) axes[a].push_back(v);
    }
    if (mass) *mass = part_volume * rest_density;
    if (particle_size) *particle_size = diam * 0.5f;
}

// Uniform double in [0, 1) with 53 random bits.
static inline double u01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace lf

extern "C" {

int lf_version(void) { return LF_ABI_VERSION; }

const char *lf_last_error(void) { return lf::g_error.c_str(); }

int lf_params_default(lf_params *p, uint32_t res) {
    if (!p) return lf::fail(LF_ERR_INVALID, "lf_params_default: null params");
    if (res == 0) return lf::fail(LF_ERR_INVALID, "lf_params_default: res must be positive");
    std::memset(p, 0, sizeof(*p));
    // FluidSystem::init (fluidsystem.cu:202-235) over FluidParams' constructor
    // defaults (fluidparams.cuh:81-97) for the coefficients init leaves alone.
    p->rest_density = 1000.0f;
    p->gas_constant = 100.0f;
    p->gamma = 7.0f;
    p->mass = 1.0f;
    p->particle_size = 1.0f;
    p->xsph_scale = 0.5f;
    p->surface_tension = 1.0f;  // DEFAULT_SURFACE_TENSION_FORCE
    p->viscosity = 0.1f;        // DEFAULT_VISCOSITY_FORCE
    p->adhesion = 1.0f;         // DEFAULT_ADHESION_FORCE
    p->vel_limit = 5.0f;
    p->acc_limit = 150.0f;
    p->vdamp = 0.0f;
    p->gravity[0] = 0.0f;
    p->gravity[1] = -9.8f;
    p->gravity[2] = 0.0f;
    // setRes: h = 1/res on the unit cube (fluidparams.cu:51-62).
    p->h = 1.0f / float(res);
    for (int a = 0; a < 3; ++a) {
        p->inv_cell[a] = float(res);
        p->grid[a] = res;
        p->box[a] = 1.0f;
    }
    return LF_OK;
}

int lf_params_check(const lf_params *p) {
    if (!p) return lf::fail(LF_ERR_INVALID, "null params");
    if (!(std::isfinite(p->h) && p->h > 0.0f)) return lf::fail(LF_ERR_INVALID, "h must be positive and finite");
    if (!(std::isfinite(p->particle_size) && p->particle_size > 0.0f))
        return lf::fail(LF_ERR_INVALID, "particle_size must be positive");
    if (!(std::isfinite(p->mass) && p->mass > 0.0f)) return lf::fail(LF_ERR_INVALID, "mass must be positive");
    uint64_t cells = 1;
    for (int a = 0; a < 3; ++a) {
        if (p->grid[a] == 0) return lf::fail(LF_ERR_INVALID, "grid[%d] must be positive", a);
        if (!(std::isfinite(p->inv_cell[a]) && p->inv_cell[a] > 0.0f))
            return lf::fail(LF_ERR_INVALID, "inv_cell[%d] must be positive", a);
        if (!(std::isfinite(p->box[a]) && p->box[a] > 0.0f))
            return lf::fail(LF_ERR_INVALID, "box[%d] must be positive", a);
        // Cells no smaller than h: the 27-cell stencil then holds every pair within h.
        if ((double)p->h * (double)p->inv_cell[a] > 1.0 + 1e-6)
            return lf::fail(LF_ERR_INVALID, "cell size along axis %d is smaller than h", a);
        // The grid covers the box, so clamped particles always have a cell.
        if ((double)p->grid[a] < (double)p->box[a] * (double)p->inv_cell[a] - 1e-4)
            return lf::fail(LF_ERR_INVALID, "grid[%d] does not cover the box", a);
        cells *= p->grid[a];
    }
    if (cells >= (1ull << 31)) return lf::fail(LF_ERR_INVALID, "too many cells (%llu)", (unsigned long long)cells);
    const float coeffs[] = {p->rest_density, p->gas_constant, p->gamma, p->xsph_scale, p->surface_tension,
                            p->viscosity, p->adhesion, p->vel_limit, p->acc_limit, p->vdamp,
                            p->gravity[0], p->gravity[1], p->gravity[2]};
    for (float v : coeffs)
        if (!std::isfinite(v)) return lf::fail(LF_ERR_INVALID, "non-finite coefficient");
    return LF_OK;
}

// FluidSystem::createFluidBlock (fluidsystem.cu:66-127), FP32 throughout.
int lf_fluid_block(uint32_t n, const float mn[3], const float mx[3], float rest_density, float *pos_xyz,
                   uint32_t *written, float *mass, float *particle_size) {
    if (!mn || !mx || (!pos_xyz && n)) return lf::fail(LF_ERR_INVALID, "lf_fluid_block: null argument");
    if (n == 0) return lf::fail(LF_ERR_INVALID, "lf_fluid_block: zero points");
    std::vector<float> ax[3];
    lf::block_axes(n, mn, mx, rest_density, ax, mass, particle_size);
    uint32_t cnt = 0;
    for (float x : ax[0])
        for (float y : ax[1])
            for (float z : ax[2])
                if (cnt < n) {
                    pos_xyz[3 * cnt] = x;
                    pos_xyz[3 * cnt + 1] = y;
                    pos_xyz[3 * cnt + 2] = z;
                    ++cnt;
                }
    if (written) *written = cnt;
    return LF_OK;
}

int lf_jittered_lattice(uint32_t nx, uint32_t ny, uint32_t nz, float dx, const float origin[3], float jitter,
                        uint64_t seed, float *pos_xyz) {
    if (!origin || !pos_xyz) return lf::fail(LF_ERR_INVALID, "lf_jittered_lattice: null argument");
    uint64_t s = seed;
    size_t k = 0;
    for (uint32_t i = 0; i < nx; ++i)
        for (uint32_t j = 0; j < ny; ++j)
            for (uint32_t l = 0; l < nz; ++l) {
                const uint32_t idx[3] = {i, j, l};
                for (int a = 0; a < 3; ++a) {
                    const double u = 2.0 * lf::u01(s) - 1.0;
                    pos_xyz[k++] = (float)((double)origin[a] + ((double)idx[a] + 0.5) * (double)dx + u * (double)jitter);
                }
            }
    return LF_OK;
}

int lf_uniform_points(uint32_t n, const float lo[3], const float hi[3], uint64_t seed, float *pos_xyz) {
    if (!lo || !hi || (!pos_xyz && n)) return lf::fail(LF_ERR_INVALID, "lf_uniform_points: null argument");
    uint64_t s = seed;
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            float v = (float)((double)lo[a] + lf::u01(s) * ((double)hi[a] - (double)lo[a]));
            if (v >= hi[a]) v = std::nextafter(hi[a], lo[a]);
            pos_xyz[3 * (size_t)i + a] = v;
        }
    return LF_OK;
}

// dumpToGeo (test/src/main.cpp:10-74): same header, point lines, primitive
// block and default ostream float formatting.
int lf_write_geo(const char *filebase, uint32_t cnt, const float *pos, const float *col, const float *vel,
                 uint32_t n, float pscale) {
    if (!filebase || (n && (!pos || !col || !vel))) return lf::fail(LF_ERR_INVALID, "lf_write_geo: null argument");
    char fname[4096];
    std::snprintf(fname, sizeof(fname), "%s.%04d.geo", filebase, (int)cnt);
    std::ofstream file(fname);
    if (!file.is_open()) return lf::fail(LF_ERR_INVALID, "failed to open file %s", fname);
    std::stringstream ss;
    ss << "PGEOMETRY V5\n";
    ss << "NPoints " << (size_t)n << " NPrims 1\n";
    ss << "NPointGroups 0 NPrimGroups 1\n";
    ss << "NPointAttrib 3  NVertexAttrib 0 NPrimAttrib 2 NAttrib 0\n";
    ss << "PointAttrib \n";
    ss << "Cd 3 float 1 1 1\n";
    ss << "v 3 float 0 0 0\n";
    ss << "pscale 1 float 1\n";
    for (uint32_t i = 0; i < n; ++i) {
        const float *p = pos + 3 * (size_t)i, *c = col + 3 * (size_t)i, *v = vel + 3 * (size_t)i;
        ss << p[0] << " " << p[1] << " " << p[2] << " 1 ";
        ss << "(" << c[0] << " " << c[1] << " " << c[2] << " " << v[0] << " " << v[1] << " " << v[2] << " " << pscale
           << ")\n";
    }
    ss << "PrimitiveAttrib\n";
    ss << "generator 1 index 1 location1\n";
    ss << "dopobject 1 index 1 /obj/AutoDopNetwork:1\n";
    ss << "Part " << (size_t)n << " ";
    for (size_t i = 0; i < n; ++i) ss << i << " ";
    ss << " [0\t0]\n";
    ss << "box_object1 unordered\n";
    ss << "1 1\n";
    ss << "beginExtra\n";
    ss << "endExtra\n";
    file << ss.rdbuf();
    file.close();
    if (!file) return lf::fail(LF_ERR_INVALID, "failed writing %s", fname);
    return LF_OK;
}

}  // extern "C"
