// upstream, downstream and downstream_distance: per-cell operations along the ldd.
//
// LUE's upstream (source/framework/algorithm/include/lue/framework/algorithm/definition/upstream.hpp:36-150)
// and downstream (definition/downstream.hpp:36-155) with the flow-accumulation value policies
// (flow_accumulation.hpp:47-84):
//   - a no-data ldd (255) or material in the cell gives no-data (NaN for floating point, numeric_limits
//     min for signed and max for unsigned integers: policy/default_value_policies.hpp:19-29,
//     policy/detect_no_data_by_value.hpp:25-50);
//   - upstream: the sum, starting from 0, of the material of the neighbours that drain into the cell, in the
//     window's row-major order (the order fixes the floating point result); a no-data material among them
//     gives no-data. Cells outside the array never drain into it (FlowDirectionHalo points away);
//   - downstream: the material of the downstream neighbour (no-data if that is no-data, 0 beyond the array's
//     edge: FillHaloWithConstantValue{0}); the cell's own material for a pit or any code that is no direction.
// Integer sums wrap (two's complement).
//
// downstream_distance (definition/downstream_distance.hpp:13-75): cell_size for N, W, E, S, sqrt(2) * cell_size
// (rounded in T, computed on the host as the reference computes it) for the diagonals, 0 for a pit or unknown
// code; with value policies (the Python binding, python/source/algorithm/routing_operation/
// downstream_distance.cpp:14-23) a no-data ldd gives NaN. A streaming map: 1 byte in, sizeof(T) out per cell.
//
// upstream and downstream are 3x3 gathers that read 1 + sizeof(T) and write sizeof(T) bytes per cell: HBM-bound. One
// workgroup per 64 x 16 tile; ldd and material with a one-cell halo go through LDS once; a wavefront
// writes one 64-cell tile row.
#pragma once

#include "routing_common.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <type_traits>

namespace lue_amd {

    constexpr int nb_tw = 64;
    constexpr int nb_th = 16;
    constexpr int nb_threads = 256;

    template<typename T>
    __host__ __device__ constexpr T material_no_data()
    {
        if constexpr (std::is_floating_point_v<T>)
            return std::numeric_limits<T>::quiet_NaN();
        else if constexpr (std::is_signed_v<T>)
            return std::numeric_limits<T>::min();
        else
            return std::numeric_limits<T>::max();
    }

    template<typename T>
    __device__ inline bool material_is_no_data(T v)
    {
        if constexpr (std::is_floating_point_v<T>)
            return isnan(v);
        else
            return v == material_no_data<T>();
    }

    template<typename T>
    __device__ inline T material_add(T a, T b)
    {
        if constexpr (std::is_floating_point_v<T>)
            return a + b;
        else
        {
            using U = std::make_unsigned_t<T>;
            return static_cast<T>(static_cast<U>(static_cast<U>(a) + static_cast<U>(b)));
        }
    }

    template<typename T, bool Upstream>
    __global__ void __launch_bounds__(nb_threads) flow_neighbour_kernel(
        std::uint8_t const* ldd, T const* material, std::int64_t rows, std::int64_t cols, T* out)
    {
        constexpr int hw = nb_tw + 2;
        constexpr int hh = nb_th + 2;
        __shared__ std::uint8_t s_ldd[hh][hw];
        __shared__ T s_mat[hh][hw];
        std::int64_t const r0 = static_cast<std::int64_t>(blockIdx.y) * nb_th;
        std::int64_t const c0 = static_cast<std::int64_t>(blockIdx.x) * nb_tw;
        auto const g_ldd = as_global(ldd);
        auto const g_mat = as_global(material);
        for (int i = threadIdx.x; i < hh * hw; i += nb_threads)
        {
            int const hr = i / hw;
            int const hc = i % hw;
            std::int64_t const r = r0 + hr - 1;
            std::int64_t const c = c0 + hc - 1;
            bool const inside = r >= 0 && r < rows && c >= 0 && c < cols;
            // outside: code 0 drains nowhere, material 0 (the halo values)
            s_ldd[hr][hc] = inside ? g_ldd[r * cols + c] : std::uint8_t{0};
            s_mat[hr][hc] = inside ? g_mat[r * cols + c] : T{0};
        }
        __syncthreads();

        int const lc = threadIdx.x % nb_tw;
        auto const g_out = as_global(out);
        for (int lr = threadIdx.x / nb_tw; lr < nb_th; lr += nb_threads / nb_tw)
        {
            std::int64_t const r = r0 + lr;
            std::int64_t const c = c0 + lc;
            if (r >= rows || c >= cols)
                continue;
            int const y = lr + 1;
            int const x = lc + 1;
            std::uint8_t const code = s_ldd[y][x];
            T const own = s_mat[y][x];
            T result;
            if (code == no_data_ldd || material_is_no_data(own))
                result = material_no_data<T>();
            else if constexpr (Upstream)
            {
                // (row offset, column offset, code that drains into the centre) in window order
                constexpr int window[8][3] = {{-1, -1, 3}, {-1, 0, 2}, {-1, 1, 1}, {0, -1, 6},
                                              {0, 1, 4},   {1, -1, 9}, {1, 0, 8},  {1, 1, 7}};
                T sum{0};
                bool no_data = false;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                {
                    if (s_ldd[y + window[k][0]][x + window[k][1]] == window[k][2])
                    {
                        T const v = s_mat[y + window[k][0]][x + window[k][1]];
                        no_data = no_data || material_is_no_data(v);
                        sum = material_add(sum, v);
                    }
                }
                result = no_data ? material_no_data<T>() : sum;
            }
            else
            {
                if (code >= 1 && code <= 9 && code != 5)
                {
                    int const dr = code <= 3 ? 1 : (code >= 7 ? -1 : 0);
                    int const dc = code % 3 == 1 ? -1 : (code % 3 == 0 ? 1 : 0);
                    T const v = s_mat[y + dr][x + dc];
                    result = material_is_no_data(v) ? material_no_data<T>() : v;
                }
                else
                    result = own;
            }
            g_out[r * cols + c] = result;
        }
    }

    // 4 cells per thread and step: one 32-bit ldd load, 4 stores of T (n % 4 cells at the end by thread 0 of
    // block 0)
    template<typename T>
    __global__ void __launch_bounds__(256) downstream_distance_kernel(
        std::uint8_t const* ldd, std::int64_t n, T cell_size, T diagonal, int mark_no_data, T* out)
    {
        auto distance = [&](std::uint32_t code) -> T
        {
            if (mark_no_data && code == no_data_ldd)
                return material_no_data<T>();
            if (code == 7 || code == 9 || code == 1 || code == 3)
                return diagonal;
            if (code == 8 || code == 4 || code == 6 || code == 2)
                return cell_size;
            return T{0};
        };
        std::int64_t const groups = n / 4;
        std::int64_t const stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
        auto const g_ldd = reinterpret_cast<global_ptr<std::uint32_t const>>(as_global(ldd));
        auto const g_out = as_global(out);
        for (std::int64_t g = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < groups;
             g += stride)
        {
            std::uint32_t const codes = g_ldd[g];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                g_out[4 * g + j] = distance((codes >> (8 * j)) & 0xffu);
        }
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            for (std::int64_t i = 4 * groups; i < n; ++i)
                g_out[i] = distance(as_global(ldd)[i]);
        }
    }

}  // namespace lue_amd
