// timing.hpp -- optional per-kernel device timing: the start and stop
// timestamps of each kernel's own dispatch (hipExtLaunchKernel events),
// resolved on demand (the device-side counterpart of the reference's
// host-clock HaloExchangeTimings, include/decomposition.hpp:30-38).  A timed
// dispatch starts several microseconds after the previous kernel ends, so a
// throughput measurement runs with timing off and times kernels in a pass of
// its own.
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
    kCopy,
    kGridMigrate,
    kGridScatter,
    kGridExchange,
    kPack,
    kUnpack,
    kPackReverse,
    kAccumulate,
    kLoopback,
    kKernelCount
};

bool enabled();

// Opens a timed region on this thread when timing is enabled and the stream
// is not being captured: the first kernel launched inside it (through
// claim()) carries the region's event pair.  A region that launches nothing
// records nothing.
class Scope {
public:
    Scope(Kernel kernel, hipStream_t stream);
    ~Scope();
    Scope(const Scope&) = delete;
    Scope& operator=(const Scope&) = delete;

private:
    friend void claim(hipEvent_t* start, hipEvent_t* stop);
    Kernel kernel_;
    hipEvent_t start_ = nullptr;
    hipEvent_t stop_ = nullptr;
    bool launched_ = false;
    Scope* outer_ = nullptr;
};

// The event pair of the innermost open Scope on this thread, once; null
// events otherwise (the launch is then untimed).
void claim(hipEvent_t* start, hipEvent_t* stop);

}  // namespace dsdx::timing
