// bifurcationHOST.h -- C++ host API, same names and argument lists as
// Bifurcation/include/bifurcationHOST.h:16-50.  Results go to OUT_FILE_PATH in
// the reference's text format.  The system and semantics come from the
// process defaults (chaos_set_default_options in chaoslibrary.h); failures
// throw std::runtime_error (the reference prints and exit()s).
#ifndef CHAOS_BIFURCATION_HOST_H
#define CHAOS_BIFURCATION_HOST_H

#include <string>

namespace Bifurcation {

void bifurcation1D(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                   const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                   const int writableVar, const double maxValue, const double transientTime, const double* values,
                   const int amountOfValues, const int preScaller, std::string OUT_FILE_PATH);

void bifurcation2D(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                   const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                   const int writableVar, const double maxValue, const double transientTime, const double* values,
                   const int amountOfValues, const int preScaller, const double eps, std::string OUT_FILE_PATH);

}  // namespace Bifurcation

#endif
