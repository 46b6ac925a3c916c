// lf_oracle.cpp -- FP64 CPU restatement of one WCSPH substep.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ and __graft_entry__.smoke() as the
// reference the GPU results are compared against, and by the cpu_baseline
// leg of bench.py as the timed CPU baseline.  The product never links it.
//
// Written from the formulas of the reference (citations are to
// NCCA/libfluid), not from its source text.  State is kept in particle-id
// order; the stages that decide membership (cell key, pair predicate) are
// evaluated in FP32 when ORC_PIN_F32 is set so that, on an FP32 state, they
// reproduce the GPU's integer results exactly.
//
// Deliberate, documented differences from the reference (all shared with the
// GPU path):
//   * tie order inside a cell is particle id (the reference keeps Thrust's
//     previous order, fluidsystem.cu:309);
//   * XSPH reads the pre-XSPH velocity of every neighbour (Jacobi); the
//     reference updates velocity in place while other blocks read it
//     (fluidphysics.cu:400,407);
//   * particles outside the grid get zero density/pressure/normal/forces
//     instead of whatever stale values their array slot held
//     (hash.cu:20-26, fluidsystem.cu:372-386);
//   * adhesion is skipped when its coefficient is exactly 0 (the reference
//     multiplies 0 by 1/(1-z)^3, which is NaN only when that overflows).

#include "lf_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t kNull = 0xFFFFFFFFu;
// CUDART_PI_F as the reference defines it (common/include/fluiddefines.h:16).
constexpr double kPi = 3.141592654;
// Finite-difference step of boundaryNormalSDF (libfluid/cudasrc/fluidutil.cu:106).
const double kSdfEps = (double)0.0001f;
// Polynomial scale of boundaryDensity (libfluid/cudasrc/fluidkernel.cu:65).
const double kBdScale = (double)0.0039062f;
// Cohesion offset (libfluid/cudasrc/fluidkernel.cu:37).
const double kCohOffset = (double)0.0156f;

// FP64 view of the FP32 parameters.  Derived constants follow
// FluidParams::setRes and the setters (libfluid/cudasrc/fluidparams.cu:51-213),
// recomputed in double from the FP32 base values.
struct P64 {
    double h, h2_pred, hh, h4, h6, h8, h9, invh9, halfh;
    double C6, Cg6, Cgs;
    double rho0, inv_rho0, K, gamma, inv_gamma;
    double m, psize, inv_psize, xsph, sigma, nu, alpha;
    double vlim, alim, vdamp, g[3], half_box[3];
    double eps;
    float h2f, inv_cell_f[3];
    int grid[3];
    uint32_t ncells;
};

P64 derive(const orc_params *p) {
    P64 d;
    d.h = (double)p->h;
    float h2f = p->h * p->h;  // FluidParamData::m_h2 (fluidparams.cu:61)
    d.h2f = h2f;
    d.h2_pred = (double)h2f;
    d.hh = d.h * d.h;
    d.h4 = d.hh * d.hh;
    d.h6 = d.h4 * d.hh;
    d.h8 = d.h4 * d.h4;
    d.h9 = d.h8 * d.h;
    d.invh9 = 1.0 / d.h9;
    d.halfh = 0.5 * d.h;
    // Kernel coefficients (fluidparams.cu:80-86).
    d.C6 = 315.0 / (64.0 * kPi * d.h9);
    d.Cg6 = -945.0 / (32.0 * kPi * d.h9);
    d.Cgs = -45.0 / (kPi * d.h6);
    d.rho0 = p->rest_density;
    d.inv_rho0 = (p->rest_density == 0.0f) ? 1.0 : 1.0 / (double)p->rest_density;  // fluidparams.cu:97
    d.K = p->gas_constant;
    d.gamma = p->gamma;
    d.inv_gamma = (p->gamma == 0.0f) ? 1.0 : 1.0 / (double)p->gamma;  // fluidparams.cu:106
    d.m = p->mass;
    d.psize = p->particle_size;
    d.inv_psize = 1.0 / (double)p->particle_size;
    d.xsph = p->xsph_scale;
    d.sigma = p->surface_tension;
    d.nu = p->viscosity;
    d.alpha = p->adhesion;
    d.vlim = p->vel_limit;
    d.alim = p->acc_limit;
    d.vdamp = p->vdamp;
    for (int a = 0; a < 3; ++a) {
        d.g[a] = p->gravity[a];
        d.half_box[a] = 0.5 * (double)p->box[a];
        d.inv_cell_f[a] = p->inv_cell[a];
        d.grid[a] = (int)p->grid[a];
    }
    d.eps = (double)FLT_MIN;  // FluidParams ctor: m_eps = numeric_limits<float>::min() (fluidparams.cu:48)
    d.ncells = p->grid[0] * p->grid[1] * p->grid[2];
    return d;
}

// ---- cell keys (hash.cu:10-27, fluidutil.cu:15-39) -------------------------

// grid_from_point: floor(x * res) per axis, in FP32.
inline uint32_t key_f32(const P64 &d, float x, float y, float z) {
    float fx = std::floor(x * d.inv_cell_f[0]);
    float fy = std::floor(y * d.inv_cell_f[1]);
    float fz = std::floor(z * d.inv_cell_f[2]);
    // cell_from_grid(int3): inside [0,res) on every axis, else NULL_HASH.
    if (!(fx >= 0.0f && fx < (float)d.grid[0] && fy >= 0.0f && fy < (float)d.grid[1] &&
          fz >= 0.0f && fz < (float)d.grid[2]))
        return kNull;
    uint32_t ix = (uint32_t)fx, iy = (uint32_t)fy, iz = (uint32_t)fz;
    return ix + (uint32_t)d.grid[0] * (iy + (uint32_t)d.grid[1] * iz);
}

inline uint32_t key_f64(const P64 &d, const double *x) {
    double f[3];
    for (int a = 0; a < 3; ++a) {
        f[a] = std::floor(x[a] * (double)d.inv_cell_f[a]);
        if (!(f[a] >= 0.0 && f[a] < (double)d.grid[a])) return kNull;
    }
    return (uint32_t)f[0] + (uint32_t)d.grid[0] * ((uint32_t)f[1] + (uint32_t)d.grid[1] * (uint32_t)f[2]);
}

// Squared distance in FP32 with an explicit FMA order; the GPU kernels use
// the same expression, so membership decided on it is identical.
inline float rr_f32(const float *a, const float *b) {
    float dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    return std::fma(dz, dz, std::fma(dy, dy, dx * dx));
}

inline double rr_f64(const double *a, const double *b) {
    double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    return dx * dx + dy * dy + dz * dz;
}

// ---- smoothing kernels (fluidkernel.cu) ------------------------------------

inline double poly6(const P64 &d, double r) {  // fluidkernel.cu:5-8
    if (r >= d.h) return 0.0;
    double t = d.hh - r * r;
    return d.C6 * t * t * t;
}
inline double grad_poly6(const P64 &d, double r) {  // fluidkernel.cu:10-13
    if (r >= d.h) return 0.0;
    double t = d.hh - r * r;
    return d.Cg6 * r * t * t;
}
inline double grad_spiky(const P64 &d, double r) {  // fluidkernel.cu:25-28
    if (r >= d.h) return 0.0;
    double t = d.h - r;
    return d.Cgs * t * t;
}
// cohesionKernel (fluidkernel.cu:35-38).  Operator precedence there makes
// the coefficient multiply only the (r > h/2) test, so it never scales the
// result: r >= h -> 0; r > h/2 -> (h-r)^3 r^3; else 2(h-r)^3 r^3 - 0.0156 h^6.
inline double cohesion(const P64 &d, double r) {
    if (r >= d.h) return 0.0;
    double t = d.h - r;
    double base = t * t * t * r * r * r;
    return (r > d.halfh) ? base : 2.0 * base - d.h6 * kCohOffset;
}
// boundaryDensity (fluidkernel.cu:45-67).
inline double boundary_density(const P64 &d, double dist) {
    if (dist > d.h || dist < 0.0) return 0.0;
    double c = -dist, c2 = c * c, c3 = c2 * c, c4 = c3 * c, c6 = c3 * c3;
    double poly = c3 * (35.0 * c6 - 180.0 * c4 * d.hh + 378.0 * c2 * d.h4 - 420.0 * d.h6) +
                  315.0 * c * d.h8 + 128.0 * d.h9;
    return 1.0 - kBdScale * d.invh9 * poly;
}

// ---- box SDF (fluidutil.cu:85-112), box [0, L] --------------------------------

inline double sdf(const P64 &d, const double *x) {
    double q[3];
    for (int a = 0; a < 3; ++a) q[a] = std::fabs(x[a] - d.half_box[a]) - d.half_box[a];
    double inside = std::min(std::max(q[0], std::max(q[1], q[2])), 0.0);
    double o0 = std::max(q[0], 0.0), o1 = std::max(q[1], 0.0), o2 = std::max(q[2], 0.0);
    return inside + std::sqrt(o0 * o0 + o1 * o1 + o2 * o2);
}

inline void sdf_normal(const P64 &d, const double *x, double *n) {
    double v[3];
    for (int a = 0; a < 3; ++a) {
        double xp[3] = {x[0], x[1], x[2]}, xm[3] = {x[0], x[1], x[2]};
        xp[a] += kSdfEps;
        xm[a] -= kSdfEps;
        v[a] = sdf(d, xp) - sdf(d, xm);
    }
    double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int a = 0; a < 3; ++a) n[a] = v[a] / len;
}

// ---- cell table ------------------------------------------------------------

struct CellTable {
    std::vector<uint32_t> start;  // ncells + 2
    std::vector<uint32_t> order;  // slot -> id, sorted by (key, id)
};

// Counting sort over the buckets 0..ncells (bucket ncells = out of grid),
// filled in id order, so ties are broken by id.
void build_cells(const P64 &d, const std::vector<uint32_t> &keys, uint32_t n, CellTable &ct) {
    uint32_t nb = d.ncells + 1;
    ct.start.assign(nb + 1, 0);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t b = keys[i] == kNull ? d.ncells : keys[i];
        ct.start[b + 1]++;
    }
    for (uint32_t b = 0; b < nb; ++b) ct.start[b + 1] += ct.start[b];
    std::vector<uint32_t> fill(ct.start.begin(), ct.start.end() - 1);
    ct.order.assign(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t b = keys[i] == kNull ? d.ncells : keys[i];
        ct.order[fill[b]++] = i;
    }
}

// Calls fn(j) for every particle j in the clipped 3x3x3 stencil around `key`
// (the loops at fluidphysics.cu:30-53; the stencil is clipped at grid faces).
template <class Fn>
inline void for_stencil(const P64 &d, const CellTable &ct, uint32_t key, Fn &&fn) {
    int gx = d.grid[0], gy = d.grid[1], gz = d.grid[2];
    int cx = (int)(key % (uint32_t)gx);
    int cy = (int)((key / (uint32_t)gx) % (uint32_t)gy);
    int cz = (int)(key / ((uint32_t)gx * (uint32_t)gy));
    for (int ox = (cx == 0 ? 0 : -1); ox <= (cx == gx - 1 ? 0 : 1); ++ox)
        for (int oy = (cy == 0 ? 0 : -1); oy <= (cy == gy - 1 ? 0 : 1); ++oy)
            for (int oz = (cz == 0 ? 0 : -1); oz <= (cz == gz - 1 ? 0 : 1); ++oz) {
                uint32_t c = (uint32_t)(cx + ox) + (uint32_t)gx * ((uint32_t)(cy + oy) + (uint32_t)gy * (uint32_t)(cz + oz));
                for (uint32_t s = ct.start[c]; s < ct.start[c + 1]; ++s) fn(ct.order[s]);
            }
}

struct Vec3 {
    double x, y, z;
};

}  // namespace

extern "C" {

int orc_version(void) { return 1; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

uint32_t orc_num_cells(const orc_params *p) { return p->grid[0] * p->grid[1] * p->grid[2]; }

int orc_hash(const orc_params *p, const float *pos, uint32_t n, uint32_t *keys, uint32_t *counts) {
    P64 d = derive(p);
    if (counts) std::memset(counts, 0, sizeof(uint32_t) * (d.ncells + 1));
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t k = key_f32(d, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        keys[i] = k;
        if (counts) counts[k == kNull ? d.ncells : k]++;
    }
    return 0;
}

// Stable sort of ids by key (std::stable_sort over id order == (key, id) order).
int orc_sort(const uint32_t *keys, uint32_t n, uint32_t *order) {
    std::vector<uint32_t> ids(n);
    for (uint32_t i = 0; i < n; ++i) ids[i] = i;
    std::stable_sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    std::memcpy(order, ids.data(), sizeof(uint32_t) * n);
    return 0;
}

// Exclusive scan of the per-bucket counts (fluidsystem.cu:316), plus the end
// sentinel.  start[ncells] is the first out-of-grid slot; start[ncells+1] = n.
int orc_cell_start(const orc_params *p, const uint32_t *keys, uint32_t n, uint32_t *start) {
    P64 d = derive(p);
    uint32_t nb = d.ncells + 1;
    std::vector<uint32_t> cnt(nb, 0);
    for (uint32_t i = 0; i < n; ++i) cnt[keys[i] == kNull ? d.ncells : keys[i]]++;
    uint32_t acc = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        start[b] = acc;
        acc += cnt[b];
    }
    start[nb] = acc;
    return 0;
}

int64_t orc_neighbours(const orc_params *p, const float *pos, uint32_t n, uint32_t *counts, uint32_t *list,
                       uint64_t cap) {
    P64 d = derive(p);
    std::vector<uint32_t> keys(n);
    for (uint32_t i = 0; i < n; ++i) keys[i] = key_f32(d, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    CellTable ct;
    build_cells(d, keys, n, ct);
    uint64_t total = 0;
    std::vector<uint32_t> tmp;
    for (uint32_t i = 0; i < n; ++i) {
        tmp.clear();
        if (keys[i] != kNull) {
            for_stencil(d, ct, keys[i], [&](uint32_t j) {
                if (j != i && rr_f32(&pos[3 * i], &pos[3 * j]) <= d.h2f) tmp.push_back(j);
            });
        }
        std::sort(tmp.begin(), tmp.end());
        counts[i] = (uint32_t)tmp.size();
        if (total + tmp.size() > cap) return -1;
        if (list) std::memcpy(list + total, tmp.data(), tmp.size() * sizeof(uint32_t));
        total += tmp.size();
    }
    return (int64_t)total;
}

int64_t orc_neighbours_brute(const orc_params *p, const float *pos, uint32_t n, int stencil, uint32_t *counts,
                             uint32_t *list, uint64_t cap) {
    P64 d = derive(p);
    std::vector<uint32_t> keys(n);
    for (uint32_t i = 0; i < n; ++i) keys[i] = key_f32(d, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    auto coords = [&](uint32_t k, int *c) {
        c[0] = (int)(k % (uint32_t)d.grid[0]);
        c[1] = (int)((k / (uint32_t)d.grid[0]) % (uint32_t)d.grid[1]);
        c[2] = (int)(k / ((uint32_t)d.grid[0] * (uint32_t)d.grid[1]));
    };
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t c = 0;
        if (keys[i] != kNull) {
            int ci[3];
            coords(keys[i], ci);
            for (uint32_t j = 0; j < n; ++j) {
                if (j == i || keys[j] == kNull) continue;
                if (!(rr_f32(&pos[3 * i], &pos[3 * j]) <= d.h2f)) continue;
                if (stencil) {
                    int cj[3];
                    coords(keys[j], cj);
                    if (std::abs(cj[0] - ci[0]) > 1 || std::abs(cj[1] - ci[1]) > 1 || std::abs(cj[2] - ci[2]) > 1)
                        continue;
                }
                if (total >= cap) return -1;
                if (list) list[total] = j;
                ++total;
                ++c;
            }
        }
        counts[i] = c;
    }
    return (int64_t)total;
}

int orc_step(const orc_params *p, double dt, uint32_t n, double *pos, double *vel, double *acc, uint32_t flags,
             orc_out *out) {
    const P64 d = derive(p);
    const bool pin = (flags & ORC_PIN_F32) != 0;

    // FP32 copy of the positions for pinned membership decisions.
    std::vector<float> pf;
    if (pin) {
        pf.resize(3 * (size_t)n);
        for (size_t k = 0; k < 3 * (size_t)n; ++k) pf[k] = (float)pos[k];
    }
    auto in_range = [&](uint32_t i, uint32_t j, bool strict_eps, double &rr) -> bool {
        rr = rr_f64(&pos[3 * i], &pos[3 * j]);
        if (pin) {
            float r32 = rr_f32(&pf[3 * i], &pf[3 * j]);
            return r32 <= d.h2f && (!strict_eps || r32 > FLT_MIN);
        }
        return rr <= d.h2_pred && (!strict_eps || rr > d.eps);
    };

    // 1. hash (fluidsystem.cu:304-306).
    std::vector<uint32_t> keys(n);
    for (uint32_t i = 0; i < n; ++i)
        keys[i] = pin ? key_f32(d, pf[3 * i], pf[3 * i + 1], pf[3 * i + 2]) : key_f64(d, &pos[3 * i]);
    // 2. sort by (key, id) and 3. cell start table (fluidsystem.cu:309-316).
    CellTable ct;
    build_cells(d, keys, n, ct);

    std::vector<double> rho(n, 0.0), prs(n, 0.0), cs(n, 0.0);
    std::vector<Vec3> nrm(n, Vec3{0, 0, 0});
    std::vector<Vec3> fP(n, Vec3{0, 0, 0}), fV(n, Vec3{0, 0, 0}), fT(n, Vec3{0, 0, 0}), fA(n, Vec3{0, 0, 0});

    // 4. density and Tait EOS (computeDensity, fluidphysics.cu:8-84).
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint32_t i = (uint32_t)ii;
        if (keys[i] == kNull) continue;
        double sum = 0.0;
        for_stencil(d, ct, keys[i], [&](uint32_t j) {
            double rr;
            if (j != i && in_range(i, j, false, rr)) sum += d.m * poly6(d, std::sqrt(rr));
        });
        double dist = std::fabs(sdf(d, &pos[3 * i]));
        double corrected = sum;
        if (dist < d.h) corrected += d.rho0 * (1.0 - boundary_density(d, dist));
        rho[i] = (sum == 0.0) ? d.rho0 : corrected;
        double ratio = corrected * d.inv_rho0;
        prs[i] = (corrected == 0.0) ? 0.0 : (d.K * d.rho0 * d.inv_gamma) * (std::pow(ratio, d.gamma) - 1.0);
        cs[i] = d.K * std::pow(ratio, d.gamma - 1.0);
    }

    // 5. normals (computeNormals, fluidphysics.cu:154-210).
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint32_t i = (uint32_t)ii;
        if (keys[i] == kNull) continue;
        double s[3] = {0, 0, 0};
        for_stencil(d, ct, keys[i], [&](uint32_t j) {
            double rr;
            if (j != i && rho[j] > d.eps && in_range(i, j, true, rr)) {
                double r = std::sqrt(rr);
                double w = d.m * grad_poly6(d, r) / (rho[j] * r);
                for (int a = 0; a < 3; ++a) s[a] += w * (pos[3 * i + a] - pos[3 * j + a]);
            }
        });
        nrm[i] = Vec3{d.h * s[0], d.h * s[1], d.h * s[2]};
    }

    // 6. forces (computeAllForces, fluidphysics.cu:215-352).
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint32_t i = (uint32_t)ii;
        if (keys[i] == kNull) continue;
        double P[3] = {0, 0, 0}, V[3] = {0, 0, 0}, T[3] = {0, 0, 0};
        const double *xi = &pos[3 * i], *vi = &vel[3 * i];
        const double ni[3] = {nrm[i].x, nrm[i].y, nrm[i].z};
        for_stencil(d, ct, keys[i], [&](uint32_t j) {
            double rr;
            if (!(j != i && rho[j] > d.eps && in_range(i, j, true, rr))) return;
            const double *xj = &pos[3 * j], *vj = &vel[3 * j];
            double rv[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};  // x_j - x_i
            double r = std::sqrt(rr), inv_r = 1.0 / r;
            double ux = -((vi[0] - vj[0]) * rv[0] + (vi[1] - vj[1]) * rv[1] + (vi[2] - vj[2]) * rv[2]);
            double Pi = 0.0;
            if (ux < 0.0) {
                double mu = (d.h * ux) / (rr + 0.01 * d.h2_pred);
                Pi = (-d.nu * (cs[i] + cs[j]) * mu) / (rho[i] + rho[j]);
            }
            double mid = prs[j] / (rho[i] * rho[i]) + prs[i] / (rho[j] * rho[j]);
            double common = d.m * grad_spiky(d, r) * inv_r;
            double coh = -d.m * d.m * cohesion(d, r) * inv_r;
            double tscale = 2.0 * d.rho0 / (rho[i] + rho[i]);
            const double nj[3] = {nrm[j].x, nrm[j].y, nrm[j].z};
            for (int a = 0; a < 3; ++a) {
                P[a] += common * rv[a] * mid;
                V[a] += common * rv[a] * (mid + Pi);
                double curv = -d.sigma * d.m * (ni[a] - nj[a]);
                T[a] += tscale * (coh * rv[a] + curv);
            }
        });
        fP[i] = Vec3{P[0], P[1], P[2]};
        fV[i] = Vec3{d.nu * V[0], d.nu * V[1], d.nu * V[2]};
        fT[i] = Vec3{d.sigma * T[0], d.sigma * T[1], d.sigma * T[2]};
        if (d.alpha != 0.0) {
            double z = -std::min(sdf(d, xi), 0.0) * 0.25 * d.inv_psize;
            if (z < 1.0) {
                double nb[3];
                sdf_normal(d, xi, nb);
                double one_z = 1.0 - z;
                double s = d.alpha * rho[i] * d.inv_rho0 / (one_z * one_z * one_z);
                fA[i] = Vec3{s * nb[0], s * nb[1], s * nb[2]};
            }
        }
    }

    // 7. leapfrog integration of every particle (fluidintegrator.cu:33-70).
    std::vector<double> sdf_after(n);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint32_t i = (uint32_t)ii;
        double *x = &pos[3 * i], *v = &vel[3 * i], *a = &acc[3 * i];
        double F[3] = {fP[i].x + fT[i].x + fA[i].x + fV[i].x, fP[i].y + fT[i].y + fA[i].y + fV[i].y,
                       fP[i].z + fT[i].z + fA[i].z + fV[i].z};
        double an[3];
        for (int k = 0; k < 3; ++k) an[k] = F[k] * d.m + d.g[k];
        double s = an[0] * an[0] + an[1] * an[1] + an[2] * an[2];
        if (s > d.alim * d.alim) {
            double sc = d.alim / std::sqrt(s);
            for (int k = 0; k < 3; ++k) an[k] *= sc;
        }
        for (int k = 0; k < 3; ++k) {
            x[k] += v[k] * dt + 0.5 * a[k] * (dt * dt);
            v[k] += 0.5 * (a[k] + an[k]) * dt;
        }
        s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (s > d.vlim * d.vlim) {
            double sc = d.vlim / std::sqrt(s);
            for (int k = 0; k < 3; ++k) v[k] *= sc;
        }
        // boundaryCheckSDF (fluidintegrator.cu:11-26)
        double dd = sdf(d, x);
        sdf_after[i] = dd;
        if (dd > -d.psize) {
            double nb[3];
            sdf_normal(d, x, nb);
            double vn = nb[0] * v[0] + nb[1] * v[1] + nb[2] * v[2];
            for (int k = 0; k < 3; ++k) {
                x[k] -= (dd + d.psize) * nb[k];
                v[k] -= (1.0 + d.vdamp) * vn * nb[k];
            }
        }
        for (int k = 0; k < 3; ++k) a[k] = an[k];
    }

    if (out && out->vel_half) std::memcpy(out->vel_half, vel, sizeof(double) * 3 * (size_t)n);

    // 8. XSPH (correctVelocityXSPH, fluidphysics.cu:355-409), with the cell
    // table of this substep and the integrated positions; Jacobi update.
    if (pin)
        for (size_t k = 0; k < 3 * (size_t)n; ++k) pf[k] = (float)pos[k];
    std::vector<double> px;  // integrated positions as FP32 storage holds them (ORC_XSPH_F32_POS)
    if (flags & ORC_XSPH_F32_POS) {
        px.resize(3 * (size_t)n);
        for (size_t k = 0; k < 3 * (size_t)n; ++k) px[k] = (double)(float)pos[k];
    }
    std::vector<double> vh(vel, vel + 3 * (size_t)n);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint32_t i = (uint32_t)ii;
        if (keys[i] == kNull) continue;
        double s[3] = {0, 0, 0};
        for_stencil(d, ct, keys[i], [&](uint32_t j) {
            double rr;
            if (j != i && in_range(i, j, false, rr) && (rho[i] + rho[j]) > d.eps) {
                if (!px.empty()) rr = rr_f64(&px[3 * i], &px[3 * j]);
                double c = poly6(d, std::sqrt(rr)) * ((2.0 * d.m) / (rho[i] + rho[j]));
                for (int a = 0; a < 3; ++a) s[a] += c * (vh[3 * j + a] - vh[3 * i + a]);
            }
        });
        for (int a = 0; a < 3; ++a) vel[3 * i + a] = vh[3 * i + a] + d.xsph * s[a];
    }

    if (out) {
        if (out->keys) std::memcpy(out->keys, keys.data(), sizeof(uint32_t) * n);
        if (out->order) std::memcpy(out->order, ct.order.data(), sizeof(uint32_t) * n);
        if (out->start) std::memcpy(out->start, ct.start.data(), sizeof(uint32_t) * (d.ncells + 2));
        for (uint32_t i = 0; i < n; ++i) {
            if (out->density) out->density[i] = rho[i];
            if (out->pressure) out->pressure[i] = prs[i];
            if (out->sound_speed) out->sound_speed[i] = cs[i];
            if (out->sdf_after) out->sdf_after[i] = sdf_after[i];
            const Vec3 *src[5] = {&nrm[i], &fP[i], &fV[i], &fT[i], &fA[i]};
            double *dst[5] = {out->normal, out->f_pressure, out->f_viscosity, out->f_tension, out->f_adhesion};
            for (int q = 0; q < 5; ++q)
                if (dst[q]) {
                    dst[q][3 * i] = src[q]->x;
                    dst[q][3 * i + 1] = src[q]->y;
                    dst[q][3 * i + 2] = src[q]->z;
                }
        }
    }
    return 0;
}

double orc_poly6(const orc_params *p, double r) { return poly6(derive(p), r); }
double orc_grad_poly6(const orc_params *p, double r) { return grad_poly6(derive(p), r); }
double orc_grad_spiky(const orc_params *p, double r) { return grad_spiky(derive(p), r); }
double orc_cohesion(const orc_params *p, double r) { return cohesion(derive(p), r); }
double orc_boundary_density(const orc_params *p, double dist) { return boundary_density(derive(p), dist); }
double orc_sdf(const orc_params *p, const double *x) { return sdf(derive(p), x); }
void orc_sdf_normal(const orc_params *p, const double *x, double *n) { sdf_normal(derive(p), x, n); }

// DensityColourOperator (fluidutil.cu:67-80).
void orc_colour(const orc_params *p, double density, double *rgb) {
    P64 d = derive(p);
    double ratio = density * d.inv_rho0;
    rgb[0] = 1.0;
    rgb[1] = 0.0;
    rgb[2] = 0.0;
    if (ratio < 1.0) {
        rgb[0] = 0.0;
        rgb[1] = ratio;
        rgb[2] = 1.0 - ratio;
    } else if (ratio < 2.0) {
        rgb[0] = ratio - 1.0;
        rgb[1] = 2.0 - ratio;
        rgb[2] = 0.0;
    }
}

void orc_colours(const orc_params *p, uint32_t n, const double *density, double *rgb) {
    for (uint32_t i = 0; i < n; ++i) orc_colour(p, density[i], rgb + 3 * (size_t)i);
}

}  // extern "C"
