// host_api.cpp -- the reference's C++ host API over the C-ABI
// (Bifurcation/src/bifurcationHOST.cpp:5-74, LLE/src/LLEHost.cpp:8-42,
// Basins/src/basinsHOST.cpp:5-40 forward their arguments the same way).
#include <stdexcept>
#include <string>

#include "../../include/LLEHost.h"
#include "../../include/basinsHOST.h"
#include "../../include/bifurcationHOST.h"
#include "../../include/legacyHOST.h"
#include "../../include/chaoslibrary.h"

namespace {
void check(int rc)
{
    if (rc != CHAOS_OK) throw std::runtime_error(chaos_last_error());
}
}  // namespace

namespace Bifurcation {

void bifurcation1D(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                   const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                   const int writableVar, const double maxValue, const double transientTime, const double* values,
                   const int amountOfValues, const int preScaller, std::string OUT_FILE_PATH)
{
    check(chaos_bifurcation1D(nullptr, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                              indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                              preScaller, OUT_FILE_PATH.c_str()));
}

void bifurcation2D(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                   const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                   const int writableVar, const double maxValue, const double transientTime, const double* values,
                   const int amountOfValues, const int preScaller, const double eps, std::string OUT_FILE_PATH)
{
    check(chaos_bifurcation2D(nullptr, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                              indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                              preScaller, eps, OUT_FILE_PATH.c_str()));
}

}  // namespace Bifurcation

namespace LLE {

void LLE2D(const double tMax, const double NT, const int nPts, const double h, const double eps,
           const double* initialConditions, const int amountOfInitialConditions, const double* ranges,
           const int* indicesOfMutVars, const int writableVar, const double maxValue, const double transientTime,
           const double* values, const int amountOfValues, std::string OUT_FILE_PATH)
{
    check(chaos_LLE2D(nullptr, tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions, ranges,
                      indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                      OUT_FILE_PATH.c_str()));
}

}  // namespace LLE

namespace Basins {

void basinsOfAttraction_2(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                          const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                          const int writableVar, const double maxValue, const double transientTime,
                          const double* values, const int amountOfValues, const int preScaller, const double eps,
                          std::string OUT_FILE_PATH)
{
    check(chaos_basinsOfAttraction_2(nullptr, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                                     indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                                     preScaller, eps, OUT_FILE_PATH.c_str()));
}

}  // namespace Basins

namespace old_library {

void TimeDomainCalculation(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                           const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                           const int writableVar, const double maxValue, const double transientTime,
                           const double* values, const int amountOfValues, const int preScaller,
                           std::string OUT_FILE_PATH)
{
    check(chaos_TimeDomainCalculation(nullptr, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                                      indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                                      preScaller, OUT_FILE_PATH.c_str()));
}

void bifurcation1DIC(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                     const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                     const int writableVar, const double maxValue, const double transientTime, const double* values,
                     const int amountOfValues, const int preScaller, std::string OUT_FILE_PATH)
{
    check(chaos_bifurcation1DIC(nullptr, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                                indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                                preScaller, OUT_FILE_PATH.c_str()));
}

void bifurcation2DIC(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                     const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                     const int writableVar, const double maxValue, const double transientTime, const double* values,
                     const int amountOfValues, const int preScaller, const double eps, std::string OUT_FILE_PATH)
{
    check(chaos_bifurcation2DIC(nullptr, tMax, nPts, h, amountOfInitialConditions, initialConditions, ranges,
                                indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,
                                preScaller, eps, OUT_FILE_PATH.c_str()));
}

#define CHAOS_LEGACY_LLE_STYLE(NAME)                                                                              \
    void NAME(const double tMax, const double NT, const int nPts, const double h, const double eps,                \
              const double* initialConditions, const int amountOfInitialConditions, const double* ranges,          \
              const int* indicesOfMutVars, const int writableVar, const double maxValue,                           \
              const double transientTime, const double* values, const int amountOfValues,                          \
              std::string OUT_FILE_PATH)                                                                           \
    {                                                                                                              \
        check(chaos_##NAME(nullptr, tMax, NT, nPts, h, eps, initialConditions, amountOfInitialConditions, ranges,  \
                           indicesOfMutVars, writableVar, maxValue, transientTime, values, amountOfValues,         \
                           OUT_FILE_PATH.c_str()));                                                                \
    }

CHAOS_LEGACY_LLE_STYLE(LLE1D)
CHAOS_LEGACY_LLE_STYLE(LLE1DIC)
CHAOS_LEGACY_LLE_STYLE(LLE2DIC)
CHAOS_LEGACY_LLE_STYLE(LS1D)
CHAOS_LEGACY_LLE_STYLE(LS2D)

#undef CHAOS_LEGACY_LLE_STYLE

}  // namespace old_library
