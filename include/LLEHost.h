// LLEHost.h -- C++ host API, same name and argument list as
// LLE/include/LLEHost.h:25-41.  Writes the reference's CSV; throws
// std::runtime_error on failure.
#ifndef CHAOS_LLE_HOST_H
#define CHAOS_LLE_HOST_H

#include <string>

namespace LLE {

void LLE2D(const double tMax, const double NT, const int nPts, const double h, const double eps,
           const double* initialConditions, const int amountOfInitialConditions, const double* ranges,
           const int* indicesOfMutVars, const int writableVar, const double maxValue, const double transientTime,
           const double* values, const int amountOfValues, std::string OUT_FILE_PATH);

}  // namespace LLE

#endif
