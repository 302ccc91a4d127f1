// Seeded synthetic inputs for the routing workloads, generated on the device.
//
// Counter-based: every value is a pure function of (seed, stream, cell index), through splitmix64
// finalisers, so the same arrays can be produced on the host (lue_amd/synthetic.py) bit for bit.
// Floating point here is evaluated without contraction (no FMA), so host and device round alike.
//
// DEM (the workload definition, not a LUE algorithm):
//     e(r, c) = (R - r) + 50 |sin(pi c / 256)| + 0.9 u(r, c),   u ~ U[0, 1)
// |sin(pi c / 256)| depends on c mod 256 only; the 256 values come from a host table, so both sides
// use the same numbers. The ldd is the steepest descent over in-array neighbours, drop / distance with
// distance 1 or sqrt(2), ties resolved by the first neighbour in the order N, NE, E, SE, S, SW, W, NW;
// the bottom row is an outlet (5). e(r, c) - e(r + 1, c) >= 0.1, so no interior cell is a pit.
#pragma once

#include "routing_common.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lue_amd {

    __host__ __device__ inline std::uint64_t splitmix_finalise(std::uint64_t z)
    {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }

    __host__ __device__ inline std::uint64_t hash3(std::uint64_t seed, std::uint64_t stream, std::uint64_t idx)
    {
        std::uint64_t const key = splitmix_finalise(seed + 0x9E3779B97F4A7C15ull * (stream + 1));
        return splitmix_finalise(key ^ (idx * 0xD1B54A32D192ED03ull + 0x9E3779B97F4A7C15ull));
    }

    // U[0, 1) with 53 random bits
    __host__ __device__ inline double unit_f64(std::uint64_t h)
    {
        return static_cast<double>(h >> 11) * 0x1.0p-53;
    }

    constexpr int dem_tw = 64;
    constexpr int dem_th = 16;

    __device__ inline double dem_value(
        std::uint64_t seed, std::int64_t rows, std::int64_t cols, std::int64_t r, std::int64_t c, double const* sin_table)
    {
#pragma clang fp contract(off)
        double const u = unit_f64(hash3(seed, 0, static_cast<std::uint64_t>(r * cols + c)));
        double const base = static_cast<double>(rows - r);
        double const ridge = 50.0 * sin_table[c & 255];
        double const noise = 0.9 * u;
        return (base + ridge) + noise;
    }

    __global__ void __launch_bounds__(256) dem_ldd_kernel(
        std::uint64_t seed, std::int64_t rows, std::int64_t cols, double const* __restrict__ sin_table,
        std::uint8_t* __restrict__ ldd)
    {
#pragma clang fp contract(off)
        constexpr int hw = dem_tw + 2;
        constexpr int hn = (dem_tw + 2) * (dem_th + 2);
        __shared__ double s_e[hn];
        std::int64_t const r0 = static_cast<std::int64_t>(blockIdx.y) * dem_th;
        std::int64_t const c0 = static_cast<std::int64_t>(blockIdx.x) * dem_tw;
        for (int i = threadIdx.x; i < hn; i += blockDim.x)
        {
            std::int64_t const r = r0 + i / hw - 1;
            std::int64_t const c = c0 + i % hw - 1;
            s_e[i] = (r >= 0 && r < rows && c >= 0 && c < cols) ? dem_value(seed, rows, cols, r, c, sin_table) : 0.0;
        }
        __syncthreads();
        // ldd codes of the directions N, NE, E, SE, S, SW, W, NW
        constexpr std::uint8_t code_of[8] = {8, 9, 6, 3, 2, 1, 4, 7};
        double const inv_sqrt2 = 0.70710678118654752440;
        for (int i = threadIdx.x; i < dem_tw * dem_th; i += blockDim.x)
        {
            int const lr = i / dem_tw;
            int const lc = i % dem_tw;
            std::int64_t const r = r0 + lr;
            std::int64_t const c = c0 + lc;
            if (r >= rows || c >= cols)
                continue;
            std::uint8_t code = 5;
            if (r != rows - 1)
            {
                double const e0 = s_e[(lr + 1) * hw + (lc + 1)];
                double best = 0.0;
                for (int k = 0; k < 8; ++k)
                {
                    std::int64_t const r2 = r + neighbour_dr(k);
                    std::int64_t const c2 = c + neighbour_dc(k);
                    if (r2 < 0 || r2 >= rows || c2 < 0 || c2 >= cols)
                        continue;
                    double const e1 = s_e[(lr + 1 + neighbour_dr(k)) * hw + (lc + 1 + neighbour_dc(k))];
                    double const diff = e0 - e1;
                    double const drop = (k & 1) ? diff * inv_sqrt2 : diff;
                    if (drop > best)
                    {
                        best = drop;
                        code = code_of[k];
                    }
                }
            }
            ldd[r * cols + c] = code;
        }
    }

    // The synthetic DEM itself (dem_value), rounded to T.
    template<typename T>
    __global__ void __launch_bounds__(256) dem_kernel(
        std::uint64_t seed, std::int64_t rows, std::int64_t cols, double const* __restrict__ sin_table,
        T* __restrict__ out)
    {
        std::int64_t const n = rows * cols;
        std::int64_t const stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
        for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
            out[i] = static_cast<T>(dem_value(seed, rows, cols, i / cols, i % cols, sin_table));
    }

    // lo + (hi - lo) u, rounded to T
    template<typename T>
    __global__ void __launch_bounds__(256) uniform_kernel(
        std::uint64_t seed, std::uint64_t stream, double lo, double hi, std::int64_t n, T* __restrict__ out)
    {
#pragma clang fp contract(off)
        std::int64_t const stride = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
        double const width = hi - lo;
        for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        {
            double const u = unit_f64(hash3(seed, stream, static_cast<std::uint64_t>(i)));
            double const scaled = width * u;
            out[i] = static_cast<T>(lo + scaled);
        }
    }

}  // namespace lue_amd
