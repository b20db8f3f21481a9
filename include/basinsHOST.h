// basinsHOST.h -- C++ host API, same name and argument list as
// Basins/include/basinsHOST.h:15-30.  Writes OUT_FILE_PATH and its _1/_2/_3
// side files (plus _4 under CONFIG semantics); throws std::runtime_error on
// failure and never resets the device (the reference ends with
// cudaDeviceReset, Basins/src/basinsCUDA.cu:811).
#ifndef CHAOS_BASINS_HOST_H
#define CHAOS_BASINS_HOST_H

#include <string>

namespace Basins {

void basinsOfAttraction_2(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                          const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                          const int writableVar, const double maxValue, const double transientTime,
                          const double* values, const int amountOfValues, const int preScaller, const double eps,
                          std::string OUT_FILE_PATH);

}  // namespace Basins

#endif
