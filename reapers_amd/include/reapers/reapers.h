// reapers.h -- umbrella header of the REAPERS C++ API on the MI355X backend.
//
// Mirrors the reference's inc/reapers.h:22-123: namespace REAPERS, DefFpType
// (float unless REAPERS_FP64), complex = std::complex, SumOps = the backend's
// term list, State, EvolutionAlgorithm::Krylov, Model::{FermionOp, SYK}.
// The default backend (DefImpl) is GPUImpl over libreapers_hip.so's C ABI;
// link with -lreapers_hip.  Defining REAPERS_NOGPU selects CPUImpl instead
// (inc/reapers.h:96-100), and such a program needs neither the HIP library
// nor a GPU.  CPUImpl is also available next to GPUImpl in a GPU build, with
// explicit State conversions between the two.  Host code only: no HIP
// headers are needed to use it.
#pragma once

#include <algorithm>
#include <cassert>
#include <chrono>
#include <cmath>
#include <complex>
#include <concepts>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <iostream>
#include <map>
#include <memory>
#include <numeric>
#include <optional>
#include <random>
#include <set>
#include <sstream>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "reapers_hip.h"

#define _REAPERS_H_

namespace REAPERS {

template<typename T>
concept RealScalar = std::floating_point<T>;

#if defined(REAPERS_FP64) && defined(REAPERS_FP32)
#error "You can only define one of REAPERS_FP64 or REAPERS_FP32"
#elif defined(REAPERS_FP64)
using DefFpType = double;
#else
using DefFpType = float;
#endif

template<RealScalar FpType = DefFpType>
using complex = std::complex<FpType>;

template<RealScalar FpType = DefFpType>
constexpr FpType epsilon() { return std::numeric_limits<FpType>::epsilon(); }

template<typename T>
concept ComplexScalar = std::convertible_to<T, std::complex<double>>;

template<typename T>
concept ScalarType = RealScalar<T> || ComplexScalar<T>;

#include "except.h"
#include "ops.h"
#include "expm.h"
#include "hostmat.h"
#include "sumops.h"
#include "cpuimpl.h"
#ifdef REAPERS_NOGPU
using DefImpl = CPUImpl;
#else
#include "gpuimpl.h"
#include "dense.h"
using DefImpl = GPUImpl;
#endif
#include "blkops.h"

template<RealScalar FpType = DefFpType>
using SumOps = DefImpl::SumOps<FpType>;

#include "state.h"
#include "schedule.h"
#include "otoc.h"
#ifndef REAPERS_NOGPU
#include "batch.h"
#endif

namespace Model {
#include "fermions.h"
#include "syk.h"
#include "xxz.h"
}

}  // namespace REAPERS
