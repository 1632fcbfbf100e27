// The device-resident pressure equation of a DeviceFv object, for the solver's system assembly.
#pragma once

#include <cfdgpu.h>

namespace cfdg {

struct FvPressureArrays {
    int n_cells;
    int n_internal_faces;
    const double* diag;   // device arrays of the last pressure equation (null before one)
    const double* upper;
    const double* lower;
    const double* source;
};

FvPressureArrays fv_pressure_arrays(const cfdg_fv* fv);

} // namespace cfdg
