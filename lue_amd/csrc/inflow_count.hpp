// Inflow count: number of upstream neighbours per cell.
//
// Reference: source/framework/algorithm/include/lue/framework/algorithm/inflow_count.hpp:9-42 and
// definition/inflow_count.hpp:97-726. LUE counts, per partition, the in-partition neighbours draining
// into a cell and adds the neighbours of other partitions through channels; the total is the number of
// in-array neighbours whose ldd points at the cell. On one device the partition boundary does not
// exist, so each cell gathers its 8 neighbours directly. No-data ldd cells get the count's no-data
// value (255 for uint8, policy::no_data_value). Flow into a no-data cell is not counted.
#pragma once

#include "routing_common.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lue_amd {

    // Tile of 64 x 16 cells per 256-thread workgroup; ldd with halo staged in LDS.
    constexpr int ic_tw = 64;
    constexpr int ic_th = 16;

    __global__ void __launch_bounds__(256) inflow_count_kernel(
        std::uint8_t const* __restrict__ ldd, std::int64_t rows, std::int64_t cols, std::uint8_t* __restrict__ count)
    {
        constexpr int hw = ic_tw + 2;
        constexpr int hn = (ic_tw + 2) * (ic_th + 2);
        __shared__ std::uint8_t s_ldd[hn];
        std::int64_t const r0 = static_cast<std::int64_t>(blockIdx.y) * ic_th;
        std::int64_t const c0 = static_cast<std::int64_t>(blockIdx.x) * ic_tw;
        for (int i = threadIdx.x; i < hn; i += blockDim.x)
        {
            std::int64_t const r = r0 + i / hw - 1;
            std::int64_t const c = c0 + i % hw - 1;
            s_ldd[i] = (r >= 0 && r < rows && c >= 0 && c < cols) ? ldd[r * cols + c] : no_data_ldd;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < ic_tw * ic_th; i += blockDim.x)
        {
            int const lr = i / ic_tw;
            int const lc = i % ic_tw;
            std::int64_t const r = r0 + lr;
            std::int64_t const c = c0 + lc;
            if (r >= rows || c >= cols)
                continue;
            int const h = (lr + 1) * hw + (lc + 1);
            std::uint8_t n = 0;
            if (s_ldd[h] == no_data_ldd)
            {
                n = 255;
            }
            else
            {
                for (int k = 0; k < 8; ++k)
                {
                    std::uint8_t const code = s_ldd[h + neighbour_dr(k) * hw + neighbour_dc(k)];
                    n += (code != no_data_ldd && ldd_direction(code) == opposite(k)) ? 1 : 0;
                }
            }
            count[r * cols + c] = n;
        }
    }

}  // namespace lue_amd
