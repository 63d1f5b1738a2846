// timing.hpp -- optional per-kernel device timing: a pair of HIP events
// recorded on the launching stream around every kernel, resolved on demand
// (the device-side counterpart of the reference's host-clock
// HaloExchangeTimings, include/decomposition.hpp:30-38).
#pragma once

#include <hip/hip_runtime.h>

namespace dsdx::timing {

enum Kernel {
    kMigrate = 0,
    kClassify,
    kScan,
    kScatter,
    kReverse,
    kReverseFields,
    kForward,
    kRefresh,
    kGhostOwners,
    kMigrateClassify,
    kKernelCount
};

bool enabled();

// Records a start event on construction and a stop event on destruction when
// timing is enabled; does nothing otherwise.
class Scope {
public:
    Scope(Kernel kernel, hipStream_t stream);
    ~Scope();
    Scope(const Scope&) = delete;
    Scope& operator=(const Scope&) = delete;

private:
    Kernel kernel_;
    hipStream_t stream_;
    hipEvent_t start_ = nullptr;
};

}  // namespace dsdx::timing
