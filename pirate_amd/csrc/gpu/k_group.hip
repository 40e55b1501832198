// Chunk counter of a replayed HIP graph: the graph's last node adds one, so the next replay's
// kernels (which read the counter instead of a launch argument) see the next chunk index.

#include <hip/hip_runtime.h>

namespace pirate_amd {

__global__ void k_advance_counter(long *counter)
{
    if (threadIdx.x == 0)
        counter[0] = counter[0] + 1;
}

void launch_advance_counter(long *counter, hipStream_t st)
{
    hipLaunchKernelGGL(k_advance_counter, dim3(1), dim3(64), 0, st, counter);
}

}  // namespace pirate_amd
