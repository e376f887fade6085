// pzl_launch.h -- kernel launches of a batch, enqueued or recorded for a graph.
//
// A batch's segments and correlations are replayed from a graph built with
// hipGraphAddKernelNode rather than captured from a stream: a capture in one
// thread changes how the runtime treats HIP calls made meanwhile by the
// other threads of the process (the request loop's workers, the candidate
// stage), while building a graph touches no stream.  While a thread records
// (pzl_recorder set), the launch helpers of the search append each kernel --
// its address, geometry and single by-value argument block -- instead of
// enqueueing it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

struct PzlKernelRec {
    const void *func;  // the kernel (as hipLaunchKernel takes it)
    dim3 grid, block;
    size_t smem;
    std::vector<unsigned char> arg;  // the argument block
};

// Set by a thread that records a batch (the engine), null otherwise.
extern thread_local std::vector<PzlKernelRec> *pzl_recorder;

// Launch func<<<g, b, smem, st>>>(arg), or record it.
template <typename A>
hipError_t pzl_launch(const void *func, dim3 g, dim3 b, size_t smem, hipStream_t st, const A &arg)
{
    if (pzl_recorder) {
        PzlKernelRec r{func, g, b, smem, std::vector<unsigned char>(sizeof(A))};
        std::memcpy(r.arg.data(), &arg, sizeof(A));
        pzl_recorder->push_back(std::move(r));
        return hipSuccess;
    }
    void *args[1] = {(void *)&arg};
    return hipLaunchKernel(func, g, b, args, smem, st);
}
