// Profiler ranges for the five phases of a time step: a roctx range is
// opened at each phase's start boundary and closed at its end boundary, so
// `rocprofv3 --marker-trace` shows predictor / poisson / correction /
// mass_fluxes / cfl around the kernels of each step (phases of
// GIP_DNS.h:308-414).  Used as the integrator's phase hook.
#ifndef _gip_markers_
#define _gip_markers_

#include <rocprofiler-sdk-roctx/roctx.h>

inline void gip_phase_ranges(int boundary) {
    static const char *const names[5] = {"predictor", "poisson", "correction", "mass_fluxes", "cfl"};
    if (boundary > 0) roctxRangePop();
    if (boundary < 5) roctxRangePushA(names[boundary]);
}

#endif
