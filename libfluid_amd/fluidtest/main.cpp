// fluidtest -- command-line driver: double dam break, one .geo frame per
// step, per-step wall time around advance().  Same flags, defaults, console
// lines and output files as the reference driver (test/src/main.cpp:76-158);
// a small parser replaces TCLAP.  Extra flag: --nogeo skips the .geo files.

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "myfluidsystem.h"

namespace {

struct Args {
    unsigned num = 10000, grid = 16, substeps = 1, numsteps = 500;
    float adhesion = 0.1f, viscosity = 0.2f, tension = 1.0f, damp = 0.5f, timestep = 0.002f;
    std::string filebase;
    bool have_file = false, nogeo = false;
};

void usage() {
    std::cout << "fluidtest: a command line application to generate a SPH dam break simulation using the libfluid "
                 "library.\n"
                 "  -n, --num <unsigned int>       Number of particles (default 10000)\n"
                 "  -g, --grid <unsigned int>      Resolution of grid for spacial partitioning on one axis (default 16)\n"
                 "  -a, --adhesion <float>         Fluid adhesion (default 0.1)\n"
                 "  -v, --viscosity <float>        Fluid viscosity (default 0.2)\n"
                 "  -s, --tension <float>          Surface tension (default 1.0)\n"
                 "  -d, --damp <float>             Velocity collision damping (default 0.5)\n"
                 "  -t, --timestep <float>         Simulation time step (default 0.002)\n"
                 "  -p, --substeps <unsigned int>  Number of substeps in each time step (default 1)\n"
                 "  -q, --numsteps <unsigned int>  Number of overall simulation steps (default 500)\n"
                 "  -f, --filebase <string>        Base path and filename for output geometry files (required)\n"
                 "      --nogeo                    Do not write .geo files\n"
                 "Example usage: fluidtest -f geo/sim\n";
}

bool parse(int argc, char **argv, Args &a) {
    for (int i = 1; i < argc; ++i) {
        std::string k = argv[i];
        if (k == "-h" || k == "--help") {
            usage();
            std::exit(EXIT_SUCCESS);
        }
        if (k == "--nogeo") {
            a.nogeo = true;
            continue;
        }
        if (i + 1 >= argc) {
            std::cerr << "error: missing value for " << k << "\n";
            return false;
        }
        const char *v = argv[++i];
        char *end = nullptr;
        if (k == "-n" || k == "--num") a.num = (unsigned)std::strtoul(v, &end, 10);
        else if (k == "-g" || k == "--grid") a.grid = (unsigned)std::strtoul(v, &end, 10);
        else if (k == "-a" || k == "--adhesion") a.adhesion = std::strtof(v, &end);
        else if (k == "-v" || k == "--viscosity") a.viscosity = std::strtof(v, &end);
        else if (k == "-s" || k == "--tension") a.tension = std::strtof(v, &end);
        else if (k == "-d" || k == "--damp") a.damp = std::strtof(v, &end);
        else if (k == "-t" || k == "--timestep") a.timestep = std::strtof(v, &end);
        else if (k == "-p" || k == "--substeps") a.substeps = (unsigned)std::strtoul(v, &end, 10);
        else if (k == "-q" || k == "--numsteps") a.numsteps = (unsigned)std::strtoul(v, &end, 10);
        else if (k == "-f" || k == "--filebase") {
            a.filebase = v;
            a.have_file = true;
            continue;
        } else {
            std::cerr << "error: unknown argument " << k << "\n";
            return false;
        }
        if (!end || *end != '\0') {
            std::cerr << "error: bad value '" << v << "' for " << k << "\n";
            return false;
        }
    }
    if (!a.have_file) {
        std::cerr << "error: Required argument missing: filebase\n";
        return false;
    }
    return true;
}

}  // namespace

int main(int argc, char *argv[]) {
    Args args;
    if (!parse(argc, argv, args)) {
        usage();
        return EXIT_FAILURE;
    }

    MyFluidSystem splash;
    splash.setupDoubleDamBreak(args.num, args.grid, args.adhesion, args.viscosity, args.tension, args.damp);
    if (splash.lastStatus() != LF_OK) return EXIT_FAILURE;

    std::vector<float3> points(args.num), colour(args.num), velocity(args.num);
    const float dt = args.timestep;
    const uint substeps = args.substeps;
    const float pscale = splash.getPscale();
    double t_tot = 0.0;

    for (uint i = 1; i <= args.numsteps; ++i) {
        std::cout << "---Step " << i << ", Timestep " << dt * float(i) << ", Substeps " << substeps << " pscale "
                  << pscale << " ";
        auto t1 = std::chrono::high_resolution_clock::now();
        splash.advance(dt, substeps);
        auto t2 = std::chrono::high_resolution_clock::now();
        std::chrono::duration<double, std::milli> ms = t2 - t1;
        std::cout << " step time " << ms.count() << "ms----\n";
        t_tot += ms.count();
        if (splash.lastStatus() != LF_OK) return EXIT_FAILURE;

        if (!args.nogeo) {
            splash.exportToData(points, colour, velocity);
            if (lf_write_geo(args.filebase.c_str(), i, &points[0].x, &colour[0].x, &velocity[0].x,
                             (uint32_t)points.size(), pscale) != LF_OK) {
                std::cout << lf_last_error() << '\n';
                return EXIT_FAILURE;
            }
        }
    }
    std::cout << "TOTAL TIME " << t_tot << "\n";
    return EXIT_SUCCESS;
}
