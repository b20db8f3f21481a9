// dispatch.hpp -- run-time system id -> compile-time System<S> instantiation.
#pragma once

#include <type_traits>

#include "runtime.hpp"
#include "systems.hpp"

namespace chaos {

template <int S>
using SysTag = std::integral_constant<int, S>;

// Calls f(SysTag<S>{}) for the system id; every kernel is instantiated once
// per system so the step is inlined with constant NX/NA.
template <class F>
void dispatch_system(int sys, F&& f)
{
    switch (sys) {
    case SYS_CHAMELEON: f(SysTag<SYS_CHAMELEON>{}); break;
    case SYS_ROSSLER_CD: f(SysTag<SYS_ROSSLER_CD>{}); break;
    case SYS_BASINS: f(SysTag<SYS_BASINS>{}); break;
    case SYS_BASINS_2: f(SysTag<SYS_BASINS_2>{}); break;
    case SYS_DAMIR: f(SysTag<SYS_DAMIR>{}); break;
    case SYS_ROSSLER_RK4: f(SysTag<SYS_ROSSLER_RK4>{}); break;
    case SYS_LORENZ_RK4: f(SysTag<SYS_LORENZ_RK4>{}); break;
    default: fail(CHAOS_E_INVALID, "unknown system id %d", sys);
    }
}

constexpr int kMaxNX = 4;
constexpr int kMaxNA = 8;

}  // namespace chaos
