// D8 flow direction from an elevation raster (the step before the routing path).
//
// LUE's d8_flow_direction (source/framework/algorithm/include/lue/framework/algorithm/definition/
// d8_flow_direction.hpp:12-143) with its value policies (value_policies/d8_flow_direction.hpp:9-57): a
// no-data (NaN) cell gets the no-data ldd (255); any other cell starts as a sink (5) with drop 0 and takes
// each neighbour, in the order N, W, E, S, NW, NE, SW, SE, whose raw drop (centre - neighbour, no distance
// weighting) exceeds the best so far; no-data neighbours are skipped. Outside the array the window holds
// numeric_limits<T>::min() (FillHaloWithConstantValue), so an edge cell above that value drains off the
// array where that is its steepest drop. Partition borders play no role: the raster is one array.
//
// A 3x3 stencil that reads sizeof(T) and writes 1 byte per cell: HBM-bound. One workgroup per 64 x 32
// tile; the tile and its one-cell halo go through LDS once (the halo rows of neighbouring tiles are L2 hits
// or re-reads: 2 rows in 32); each thread decides 4 adjacent cells of a row and stores them as one 32-bit
// word where the row allows.
#pragma once

#include "routing_common.hpp"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace lue_amd {

    constexpr int d8_tw = 64;
    constexpr int d8_th = 32;
    constexpr int d8_threads = 256;

    template<typename T>
    __device__ inline T d8_halo_value();

    template<>
    __device__ inline float d8_halo_value<float>()
    {
        return FLT_MIN;
    }

    template<>
    __device__ inline double d8_halo_value<double>()
    {
        return DBL_MIN;
    }

    template<typename T>
    __global__ void __launch_bounds__(d8_threads)
        d8_flow_direction_kernel(T const* elevation, std::int64_t rows, std::int64_t cols, std::uint8_t* out)
    {
        constexpr int hw = d8_tw + 2;
        constexpr int hh = d8_th + 2;
        __shared__ T s[hh][hw];
        std::int64_t const r0 = static_cast<std::int64_t>(blockIdx.y) * d8_th;
        std::int64_t const c0 = static_cast<std::int64_t>(blockIdx.x) * d8_tw;
        auto const g_elev = as_global(elevation);
        T const halo = d8_halo_value<T>();
        for (int i = threadIdx.x; i < hh * hw; i += d8_threads)
        {
            int const hr = i / hw;
            int const hc = i % hw;
            std::int64_t const r = r0 + hr - 1;
            std::int64_t const c = c0 + hc - 1;
            s[hr][hc] = (r >= 0 && r < rows && c >= 0 && c < cols) ? g_elev[r * cols + c] : halo;
        }
        __syncthreads();

        // (row offset, column offset) and ldd code in LUE's order
        constexpr int dr[8] = {-1, 0, 0, 1, -1, -1, 1, 1};
        constexpr int dc[8] = {0, -1, 1, 0, -1, 1, -1, 1};
        constexpr std::uint8_t code[8] = {8, 4, 6, 2, 7, 9, 1, 3};
        constexpr int groups = d8_tw / 4;  // 4-cell groups per tile row
        for (int g = threadIdx.x; g < d8_th * groups; g += d8_threads)
        {
            int const lr = g / groups;
            int const lc = (g % groups) * 4;
            std::int64_t const r = r0 + lr;
            if (r >= rows)
                continue;
            std::uint32_t packed = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                T const centre = s[lr + 1][lc + j + 1];
                std::uint8_t direction = no_data_ldd;
                if (!isnan(centre))
                {
                    direction = 5;
                    T drop{0};
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                    {
                        T const neighbour = s[lr + 1 + dr[k]][lc + j + 1 + dc[k]];
                        T const other_drop = centre - neighbour;
                        if (!isnan(neighbour) && other_drop > drop)
                        {
                            drop = other_drop;
                            direction = code[k];
                        }
                    }
                }
                packed |= static_cast<std::uint32_t>(direction) << (8 * j);
            }
            std::int64_t const c = c0 + lc;
            std::int64_t const at = r * cols + c;
            auto const g_out = as_global(out);
            if (c + 4 <= cols && (at & 3) == 0)
            {
                *reinterpret_cast<global_ptr<std::uint32_t>>(g_out + at) = packed;
            }
            else
            {
                for (int j = 0; j < 4 && c + j < cols; ++j)
                    g_out[at + j] = static_cast<std::uint8_t>(packed >> (8 * j));
            }
        }
    }

}  // namespace lue_amd
