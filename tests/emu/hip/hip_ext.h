// CPU emulation of hipExtLaunchKernelGGL: the launch runs like hipLaunchKernelGGL; the
// start and stop events (timing only) are ignored.
#pragma once
#include "hip_runtime.h"

#define hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, start, stop, flags, ...) \
    ::emu::launch(dim3(grid), dim3(block), [&]() { kernel(__VA_ARGS__); })
