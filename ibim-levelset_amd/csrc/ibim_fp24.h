// ibim_fp24.h -- addressing of the FP16 M2L stage's per-pair result buffer.
//
// The V- and X-list GEMMs of the FP16 mixed matvec store one row of 4 p^3
// results per interaction (pair), padded to YS values (a multiple of 32), as
// FP24: four consecutive values packed into three dwords.  Pair rows reach a
// few times 10^7 on large surfaces (about 5.6M rows already overflow a 32-bit
// dword offset at p = 5), so every offset is formed in 64 bits here, on the
// host and on the device alike.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ibim {

// dwords of one pair row of YS values
__host__ __device__ constexpr size_t fp24_row_dwords(int ys) { return (size_t)(ys / 4) * 3; }

// dword offset of value `v` (a multiple of 4) of pair row `pair`
__host__ __device__ constexpr size_t fp24_offset(int64_t pair, int ys, int v) {
    return (size_t)pair * fp24_row_dwords(ys) + (size_t)(v / 4) * 3;
}

// dwords of a buffer holding `rows` pair rows
__host__ __device__ constexpr size_t fp24_buffer_dwords(int64_t rows, int ys) {
    return (size_t)rows * fp24_row_dwords(ys);
}

}  // namespace ibim
