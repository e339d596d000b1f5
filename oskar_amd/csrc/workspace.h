/*
 * Device workspaces of the correlator kernels, one per (device, stream,
 * kind).  Calls on different streams get different buffers, so concurrent
 * correlations on two streams do not share partial sums; calls on one
 * stream are ordered by the stream and reuse its buffer.  A buffer grows
 * (hipFree, which waits for the device, then hipMalloc) when a call needs
 * more; buffers live until the process ends.
 */
#ifndef OSKAR_AMD_WORKSPACE_H_
#define OSKAR_AMD_WORKSPACE_H_

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace oskar_amd {

enum WorkspaceKind
{
    WS_XCORR_SPLIT = 0,      /* direct correlator: FP64 split-K partials */
    WS_FACT_SOURCES,         /* shared-beam factorised: stations, sources */
    WS_FACT_PARTIALS,        /* shared-beam factorised: FP32 part tiles */
    WS_FACT_STATIONS,        /* per-station factorised: everything */
    WS_KINDS
};

/* A buffer of at least `bytes` for this (kind, device, stream), or NULL
 * with *err set. */
void* ws_acquire(int kind, int dev, hipStream_t stream, size_t bytes,
        hipError_t* err);

} /* namespace oskar_amd */

#endif /* OSKAR_AMD_WORKSPACE_H_ */
