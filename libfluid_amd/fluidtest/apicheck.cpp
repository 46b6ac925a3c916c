// apicheck -- drives the FluidSystem C++ class (include/fluidsystem.h) through
// the entry points the reference driver does not use, for the tests: setup()
// (fluidsystem.cu:132-147) and setupFromData() (fluidsystem.cu:179-189), then
// advance() and exportToData().
//
//   apicheck setup    <n> <res> <steps> <dt> <out>
//   apicheck fromdata <in> <res> <steps> <dt> <out>
//
// <in>:  uint32 n, then n float3 points and n float3 velocities.
// <out>: uint32 n, float pscale, float mass, then exportToData's points,
//        colour and velocity (cell order), then the points by particle id.

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "fluidsystem.h"

namespace {

bool write_out(const char *path, FluidSystem &fs) {
    const uint n = fs.numParticles();
    std::vector<float3> p(n), c(n), v(n), pid, cid, vid;
    fs.exportToData(p, c, v);
    fs.exportById(pid, cid, vid);
    if (fs.lastStatus() != LF_OK) return false;
    FILE *f = std::fopen(path, "wb");
    if (!f) return false;
    const uint32_t n32 = n;
    const float ps = fs.getPscale(), mass = fs.params().mass;
    bool ok = std::fwrite(&n32, 4, 1, f) == 1 && std::fwrite(&ps, 4, 1, f) == 1 && std::fwrite(&mass, 4, 1, f) == 1;
    for (const std::vector<float3> *a : {&p, &c, &v, &pid}) ok = ok && std::fwrite(a->data(), 12, n, f) == n;
    return std::fclose(f) == 0 && ok;
}

bool read_in(const char *path, std::vector<float3> &p, std::vector<float3> &v) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    uint32_t n = 0;
    bool ok = std::fread(&n, 4, 1, f) == 1;
    if (ok) {
        p.resize(n);
        v.resize(n);
        ok = std::fread(p.data(), 12, n, f) == n && std::fread(v.data(), 12, n, f) == n;
    }
    std::fclose(f);
    return ok;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 7) {
        std::cerr << "usage: apicheck setup <n> <res> <steps> <dt> <out>\n"
                     "       apicheck fromdata <in> <res> <steps> <dt> <out>\n";
        return 2;
    }
    const std::string mode = argv[1];
    const uint res = (uint)std::strtoul(argv[3], nullptr, 10);
    const uint steps = (uint)std::strtoul(argv[4], nullptr, 10);
    const float dt = std::strtof(argv[5], nullptr);
    FluidSystem fs;
    if (mode == "setup") {
        fs.setup((uint)std::strtoul(argv[2], nullptr, 10), res);
    } else if (mode == "fromdata") {
        std::vector<float3> p, v;
        if (!read_in(argv[2], p, v)) {
            std::cerr << "apicheck: cannot read " << argv[2] << "\n";
            return 1;
        }
        fs.setupFromData(p, v, res);
    } else {
        std::cerr << "apicheck: unknown mode " << mode << "\n";
        return 2;
    }
    if (fs.lastStatus() != LF_OK) return 1;
    for (uint k = 0; k < steps; ++k) {
        fs.advance(dt, 1);
        if (fs.lastStatus() != LF_OK) return 1;
    }
    if (!write_out(argv[6], fs)) {
        std::cerr << "apicheck: cannot write " << argv[6] << "\n";
        return 1;
    }
    return 0;
}
