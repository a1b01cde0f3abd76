// psx_launch.h - launch sizing shared by the HIP translation units (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace psx {
namespace {

int device_cus() {
    int device = 0, cus = 0;
    if (hipGetDevice(&device) != hipSuccess) return 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) return 1;
    return cus;
}

// Blocks of `kernel` the whole device holds at once.  CDNA3/4 compute units
// have 4 SIMDs with 512 VGPRs per lane each (granule 8), at most 8 waves per
// SIMD and 160 KiB of LDS; the runtime's occupancy query is taken when it
// reports more.
uint32_t resident_blocks(const void *kernel, int block) {
    int api = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&api, kernel, block, 0) != hipSuccess) api = 0;
    hipFuncAttributes attr{};
    int own = 1;
    if (hipFuncGetAttributes(&attr, kernel) == hipSuccess && attr.numRegs > 0) {
        const int vgprs = (attr.numRegs + 7) / 8 * 8;
        int waves_per_simd = 512 / vgprs;
        waves_per_simd = waves_per_simd > 8 ? 8 : waves_per_simd;
        const int waves_per_block = (block + 63) / 64;
        int blocks = 4 * waves_per_simd / waves_per_block;
        if (attr.sharedSizeBytes > 0) {
            const int by_lds = static_cast<int>((160u * 1024u) / attr.sharedSizeBytes);
            blocks = blocks < by_lds ? blocks : by_lds;
        }
        own = blocks > 0 ? blocks : 1;
    }
    const int per_cu = api > own ? api : own;
    return static_cast<uint32_t>(per_cu) * static_cast<uint32_t>(device_cus());
}

}  // namespace
}  // namespace psx
