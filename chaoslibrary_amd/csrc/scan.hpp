// scan.hpp -- device exclusive prefix sum.
#pragma once

#include "runtime.hpp"

namespace chaos {

// d_out[i] = sum_{k<i} d_in[k]; returns the total.  Synchronises st.
long long exclusive_scan(hipStream_t st, const long long* d_in, long long* d_out, long long n);

}  // namespace chaos
