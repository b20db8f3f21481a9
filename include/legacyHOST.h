// legacyHOST.h -- the reference's legacy C++ entry points that sweep the
// integrator (namespace old_library, debri/hostLibrary.cuh), same names,
// argument lists and output files.  The system is chosen at run time
// (chaos_set_default_options) instead of by the legacy library's hard-coded
// calcDiscreteModel.  Failures throw std::runtime_error.
#pragma once

#include <string>

namespace old_library {

// debri/hostLibrary.cuh:69-83
void TimeDomainCalculation(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                           const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                           const int writableVar, const double maxValue, const double transientTime,
                           const double* values, const int amountOfValues, const int preScaller,
                           std::string OUT_FILE_PATH);

// debri/hostLibrary.cuh:109-123
void bifurcation1DIC(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                     const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                     const int writableVar, const double maxValue, const double transientTime, const double* values,
                     const int amountOfValues, const int preScaller, std::string OUT_FILE_PATH);

// debri/hostLibrary.cuh:219-234
void bifurcation2DIC(const double tMax, const int nPts, const double h, const int amountOfInitialConditions,
                     const double* initialConditions, const double* ranges, const int* indicesOfMutVars,
                     const int writableVar, const double maxValue, const double transientTime, const double* values,
                     const int amountOfValues, const int preScaller, const double eps, std::string OUT_FILE_PATH);

// debri/hostLibrary.cuh:241-256
void LLE1D(const double tMax, const double NT, const int nPts, const double h, const double eps,
           const double* initialConditions, const int amountOfInitialConditions, const double* ranges,
           const int* indicesOfMutVars, const int writableVar, const double maxValue, const double transientTime,
           const double* values, const int amountOfValues, std::string OUT_FILE_PATH);

// debri/hostLibrary.cuh:263-278
void LLE1DIC(const double tMax, const double NT, const int nPts, const double h, const double eps,
             const double* initialConditions, const int amountOfInitialConditions, const double* ranges,
             const int* indicesOfMutVars, const int writableVar, const double maxValue, const double transientTime,
             const double* values, const int amountOfValues, std::string OUT_FILE_PATH);

// debri/hostLibrary.cuh:339-354
void LLE2DIC(const double tMax, const double NT, const int nPts, const double h, const double eps,
             const double* initialConditions, const int amountOfInitialConditions, const double* ranges,
             const int* indicesOfMutVars, const int writableVar, const double maxValue, const double transientTime,
             const double* values, const int amountOfValues, std::string OUT_FILE_PATH);

// debri/hostLibrary.cuh:377-392
void LS1D(const double tMax, const double NT, const int nPts, const double h, const double eps,
          const double* initialConditions, const int amountOfInitialConditions, const double* ranges,
          const int* indicesOfMutVars, const int writableVar, const double maxValue, const double transientTime,
          const double* values, const int amountOfValues, std::string OUT_FILE_PATH);

// debri/hostLibrary.cuh:416-431
void LS2D(const double tMax, const double NT, const int nPts, const double h, const double eps,
          const double* initialConditions, const int amountOfInitialConditions, const double* ranges,
          const int* indicesOfMutVars, const int writableVar, const double maxValue, const double transientTime,
          const double* values, const int amountOfValues, std::string OUT_FILE_PATH);

}  // namespace old_library
