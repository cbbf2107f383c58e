// Host-side internals shared by the C-ABI implementation files.
#pragma once

#include <string>
#include <vector>

#include "dem_mph.h"

extern "C" void dmph_params_defaults(dmph_params *p);

namespace dmph {

// lines of the last .data file that matched no key (the reference logs them, source/DEM-MPH.cpp:661-663)
extern std::vector<std::string> g_invalid_data_lines;

struct VtkFields {
    int n;
    const int *type, *ncount;
    const double *pos, *vel, *kn, *etan, *kt, *etat, *alpha, *drag, *vs, *virial;
};

int write_prof(const char *path, double time, int n, double d, const double *dmin, const double *dmax, const int *type,
               const double *pos, const double *vel);
int write_vtk(const char *path, const VtkFields &f);

}  // namespace dmph
