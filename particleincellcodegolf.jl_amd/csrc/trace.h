// trace.h -- ROCTx ranges around the C-ABI entry points, so a
// `rocprofv3 --marker-trace` run shows each call (pic_run, pic2d_run, ...)
// around the kernels it launched.  Without a profiler attached a range costs
// two calls into the ROCTx stub.
#pragma once
#include <rocprofiler-sdk-roctx/roctx.h>

namespace pic {
struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};
}  // namespace pic

#define PIC_TRACE(name) const ::pic::TraceRange pic_trace_range_(name)
