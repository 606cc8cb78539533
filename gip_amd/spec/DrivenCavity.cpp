// Lid-driven cavity on one MI355X: the counterpart of the reference's
// DrivenCavity.cpp:1-43 with the HIP specification plugged into the
// templated integrator.
//   DrivenCavityHIP <name> [--steps N] [--gamma G] [--pcg-maxit K] [--pcg-tol T]
//                          [--pcg-mode reference|fused|mixed] [--full] [--root DIR] [--device D]
//                          [--timing]
// Reads InputFiles/... (one partition, <name>_1p-0) below DIR (default: the
// working directory) and prints one line per step in the reference's format.
// --timing times every step with the HIP timer specification (a channel
// started before and stopped after the step, GIP_Timer.h:6-26) and prints
// its last, summed and mean times; each phase of a step is also a profiler
// range (roctx) named after it.
#include <cstdlib>
#include <cstring>
#include <exception>
#include <iostream>

#include "GIP_DNS.h"
#include "HIP.h"
#include "gip_markers.h"

typedef GIP_MatrixHIP Matrix;
typedef GIP_VectorHIP Vector;
typedef GIP_NonLinearHIP NonLinear;
typedef GIP_ArchHIP Architecture;

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cout << "DrivenCavityHIP <Mesh_name> [--steps N] [--gamma G] [--pcg-maxit K] [--pcg-tol T] "
                     "[--pcg-mode reference|fused|mixed] [--full] [--root DIR] [--device D] [--timing]"
                  << std::endl;
        return 0;
    }
    GIP_Parameters p;
    p.setInt("steps", 10);
    p.setDouble("gamma", 1e-3);
    p.setInt("print", 1);
    std::string root;
    bool timing = false;
    for (int a = 2; a < argc; ++a) {
        std::string k = argv[a];
        auto next = [&]() -> const char * {
            if (a + 1 >= argc) {
                std::cerr << "missing value for " << k << std::endl;
                std::exit(1);
            }
            return argv[++a];
        };
        if (k == "--steps") p.setInt("steps", std::atoi(next()));
        else if (k == "--gamma") p.setDouble("gamma", std::atof(next()));
        else if (k == "--pcg-maxit") p.setInt("pcg_maxIt", std::atoi(next()));
        else if (k == "--pcg-tol") p.setDouble("pcg_tol", std::atof(next()));
        else if (k == "--pcg-mode") p.setString("pcg_mode", next());
        else if (k == "--full") p.setInt("print", 2);
        else if (k == "--root") root = next();
        else if (k == "--device") p.setInt("device", std::atoi(next()));
        else if (k == "--timing") timing = true;
        else {
            std::cerr << "unknown option " << k << std::endl;
            return 1;
        }
    }
    try {
        GIP_DNS<Matrix, Vector, Architecture, NonLinear> driven(argv[1]);
        driven.setRoot(root);
        driven.configure(p);
        driven.setUp();
        if (!timing) {
            driven.integrate();
            return 0;
        }
        const int steps = p.getInt("steps");
        driven.setPhaseHook(gip_phase_ranges);
        GIP_TimerHIP timer(&driven.node());
        for (int k = 0; k < steps; ++k) {
            GIP_StepInfo info;
            timer.startChannel();
            driven.step(&info);
            timer.stopAllChannels();
        }
        std::printf("timer steps %d last_s %.9e sum_s %.9e avg_s %.9e\n", steps, timer.getTimeG(), timer.getSumG(),
                    timer.getTimeAvg());
    } catch (const std::exception &e) {  // unreadable or unsupported input files
        std::cerr << e.what() << std::endl;
        return 2;
    }
    return 0;
}
