// trace.h -- roctx ranges around the driver's module calls, so a profile
// (rocprofv3 --marker-trace) shows which module each kernel ran under.  The
// reference labels its kernels for Kokkos-Tools (e.g. "ForceLJNeigh::compute",
// src/force_types/force_lj_neigh_impl.h:114); these ranges are the same
// labels at module granularity.
#ifndef EMD_TRACE_H
#define EMD_TRACE_H
#include <rocprofiler-sdk-roctx/roctx.h>

class TraceRange {
public:
  explicit TraceRange(const char* name) { roctxRangePush(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};
#endif
