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
// A 3x3 stencil that reads sizeof(T) and writes 1 byte per cell: HBM-bound. Two kernels:
//   - d8_strip_kernel (rows a multiple of 4 cells long): a workgroup streams a strip of 1024 columns x 32
//     rows; each thread holds 4 adjacent columns of three rows in registers (16-byte loads), takes its left
//     and right neighbours from the adjacent lanes (DPP wave shifts; the wavefront's edge lanes load theirs),
//     keeps the next rows' loads in flight while it decides row r, and stores its 4 cells as one 32-bit word.
//     Each elevation value is read once, plus 2 rows in 32 for the strip borders; each difference between
//     two cells is computed once and used by both.
//   - d8_flow_direction_kernel (any shape): 64 x 32 tiles with a one-cell halo in LDS.
#pragma once

#include "routing_common.hpp"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <type_traits>

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
    __device__ inline T d8_max(T a, T b)  // IEEE maxNum: a quiet NaN operand yields the other one
    {
        if constexpr (std::is_same_v<T, float>)
            return fmaxf(a, b);
        else
            return fmax(a, b);
    }

    // LUE's choice for one cell from its 8 raw drops: the first drop in the order N, W, E, S, NW, NE, SW, SE
    // that exceeds all earlier ones and 0. That is the first drop equal to the maximum of 0 and the drops,
    // if that maximum is above 0 (else a sink, 5); a NaN drop (no-data neighbour) is never above anything,
    // and maxNum skips it. A no-data centre gives 255.
    template<typename T>
    __device__ inline std::uint32_t d8_pick(T centre, T n, T w, T e, T s, T nw, T ne, T sw, T se)
    {
        T const m = d8_max(d8_max(d8_max(T{0}, n), d8_max(w, e)),
                           d8_max(d8_max(s, nw), d8_max(d8_max(ne, sw), se)));
        std::uint32_t d = 3;
        d = sw == m ? 1u : d;
        d = ne == m ? 9u : d;
        d = nw == m ? 7u : d;
        d = s == m ? 2u : d;
        d = e == m ? 6u : d;
        d = w == m ? 4u : d;
        d = n == m ? 8u : d;
        d = m > T{0} ? d : 5u;
        return isnan(centre) ? std::uint32_t{no_data_ldd} : d;
    }

    __device__ inline float d8_wave_shr1(float v)  // lane i gets lane i - 1 (lane 0: 0)
    {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, false));
    }

    __device__ inline float d8_wave_shl1(float v)  // lane i gets lane i + 1 (lane 63: 0)
    {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, false));
    }

    __device__ inline double d8_wave_shr1(double v)
    {
        int const lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xF, 0xF, false);
        int const hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xF, 0xF, false);
        return __hiloint2double(hi, lo);
    }

    __device__ inline double d8_wave_shl1(double v)
    {
        int const lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xF, 0xF, false);
        int const hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xF, 0xF, false);
        return __hiloint2double(hi, lo);
    }

    constexpr int d8_strip_cols = 4 * d8_threads;  // columns per workgroup
    constexpr int d8_strip_rows = 32;

    // One row of a thread's 4 columns as loaded (v) and its neighbours (left of v[0], right of v[3]).
    template<typename T>
    struct D8Row
    {
        T v[4];
        T left;
        T right;
    };

    // The drops of a row's 4 cells towards the row below: S, SE and SW. The row below reuses them negated as
    // its N, NW and NE drops: in IEEE arithmetic b - a == -(a - b) exactly (rounding is symmetric), so every
    // difference between two cells is computed once.
    template<typename T>
    struct D8Down
    {
        T s[4];
        T se[4];
        T sw[4];
    };

    // Issue the loads of row r (all lanes, row-uniform bounds): the 4 values, and at the wavefront's edge
    // lanes the neighbour column the adjacent lane of another wavefront holds.
    template<typename T>
    __device__ inline void d8_load(T const* elevation, std::int64_t rows, std::int64_t cols, std::int64_t r,
                                   std::int64_t c, bool active, int lane, D8Row<T>& row)
    {
        T const halo = d8_halo_value<T>();
        if (r < 0 || r >= rows || !active)
        {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                row.v[j] = halo;
            row.left = halo;
            row.right = halo;
            return;
        }
        T const* const p = elevation + (r * cols + c);  // 16-byte aligned: cols % 4 == 0
        if constexpr (std::is_same_v<T, float>)
        {
            float4 const q = *reinterpret_cast<float4 const*>(p);
            row.v[0] = q.x;
            row.v[1] = q.y;
            row.v[2] = q.z;
            row.v[3] = q.w;
        }
        else
        {
            double2 const q0 = *reinterpret_cast<double2 const*>(p);
            double2 const q1 = *reinterpret_cast<double2 const*>(p + 2);
            row.v[0] = q0.x;
            row.v[1] = q0.y;
            row.v[2] = q1.x;
            row.v[3] = q1.y;
        }
        row.left = (lane == 0 && c > 0) ? p[-1] : halo;
        row.right = (lane == 63 && c + 4 < cols) ? p[4] : halo;
    }

    // Neighbours from the adjacent lanes (all lanes execute the shifts); the edge lanes keep what they loaded.
    template<typename T>
    __device__ inline void d8_neighbours(std::int64_t cols, std::int64_t c, int lane, D8Row<T>& row)
    {
        T const from_left = d8_wave_shr1(row.v[3]);
        T const from_right = d8_wave_shl1(row.v[0]);
        if (lane != 0 && c > 0)
            row.left = from_left;
        if (lane != 63)
            row.right = c + 4 < cols ? from_right : d8_halo_value<T>();
    }

    template<typename T>
    __device__ inline void d8_down(D8Row<T> const& b, D8Row<T> const& c, D8Down<T>& q)
    {
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            q.s[j] = b.v[j] - c.v[j];
            q.se[j] = b.v[j] - (j == 3 ? c.right : c.v[j + 1]);
            q.sw[j] = b.v[j] - (j == 0 ? c.left : c.v[j - 1]);
        }
    }

    // The 4 codes of centre row b (a above, c below, p = the drops of row a towards b), packed into a word;
    // leaves the drops of b towards c in q.
    template<typename T>
    __device__ inline std::uint32_t d8_row_codes(
        D8Row<T> const& a, D8Row<T> const& b, D8Row<T> const& c, D8Down<T> const& p, D8Down<T>& q)
    {
        d8_down(b, c, q);
        T h[5];  // h[j] = left-of-j minus j: the W drop of cell j is -h[j], its E drop h[j + 1]
        h[0] = b.left - b.v[0];
#pragma unroll
        for (int j = 1; j < 4; ++j)
            h[j] = b.v[j - 1] - b.v[j];
        h[4] = b.v[3] - b.right;
        std::uint32_t packed = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            T const nw = j == 0 ? b.v[0] - a.left : -p.se[j - 1];
            T const ne = j == 3 ? b.v[3] - a.right : -p.sw[j + 1];
            std::uint32_t const code =
                d8_pick(b.v[j], -p.s[j], -h[j], h[j + 1], q.s[j], nw, ne, q.sw[j], q.se[j]);
            packed |= code << (8 * j);
        }
        return packed;
    }

    // Rows whose length is a multiple of 4 (16-byte aligned rows of 4-cell groups). Depth rows load ahead
    // of the row being decided; a ring of Depth + 3 row buffers and two sets of carried drops rotate through
    // a loop unrolled over their common period, so the rotation costs no register moves.
    template<typename T, int Rows = d8_strip_rows, int Depth = 3>
    __global__ void __launch_bounds__(d8_threads)
        d8_strip_kernel(T const* elevation, std::int64_t rows, std::int64_t cols, std::uint8_t* out)
    {
        constexpr int ring = Depth + 3;
        constexpr int period = ring % 2 == 0 ? ring : 2 * ring;
        int const lane = threadIdx.x & 63;
        std::int64_t const c = (static_cast<std::int64_t>(blockIdx.x) * d8_threads + threadIdx.x) * 4;
        bool const active = c < cols;
        std::int64_t const r0 = static_cast<std::int64_t>(blockIdx.y) * Rows;
        std::int64_t const r1 = r0 + Rows < rows ? r0 + Rows : rows;
        D8Row<T> row[ring];  // row[k % ring] holds row r0 - 1 + k
        D8Down<T> down[2];   // down[k % 2]: the drops of row r0 - 1 + k towards the row below
#pragma unroll
        for (int k = 0; k < Depth + 2; ++k)
            d8_load(elevation, rows, cols, r0 - 1 + k, c, active, lane, row[k]);
        d8_neighbours(cols, c, lane, row[0]);
        d8_neighbours(cols, c, lane, row[1]);
        d8_down(row[0], row[1], down[0]);
        auto const g_out = as_global(out);
        for (std::int64_t r = r0; r < r1; r += period)
        {
#pragma unroll
            for (int k = 0; k < period; ++k)
            {
                if (r + k >= r1)
                    break;
                // centre row r + k is row[(k + 1) % ring]; row r + k + Depth + 1 goes into the free buffer
                d8_load(elevation, rows, cols, r + k + Depth + 1, c, active, lane, row[(k + Depth + 2) % ring]);
                D8Row<T>& below = row[(k + 2) % ring];
                d8_neighbours(cols, c, lane, below);
                std::uint32_t const packed =
                    d8_row_codes(row[k % ring], row[(k + 1) % ring], below, down[k % 2], down[(k + 1) % 2]);
                if (active)
                    *reinterpret_cast<global_ptr<std::uint32_t>>(g_out + ((r + k) * cols + c)) = packed;
            }
        }
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
                int const y = lr + 1;
                int const x = lc + j + 1;
                T const centre = s[y][x];
                std::uint32_t const code =
                    d8_pick(centre, centre - s[y - 1][x], centre - s[y][x - 1], centre - s[y][x + 1],
                            centre - s[y + 1][x], centre - s[y - 1][x - 1], centre - s[y - 1][x + 1],
                            centre - s[y + 1][x - 1], centre - s[y + 1][x + 1]);
                packed |= code << (8 * j);
            }
            std::int64_t const c = c0 + lc;
            std::int64_t const at = r * cols + c;
            auto const g_out = as_global(out);
            // whole-word store where the four bytes lie in the row and at a 4-byte aligned address
            if (c + 4 <= cols && ((reinterpret_cast<std::uintptr_t>(out) + static_cast<std::uintptr_t>(at)) & 3u) == 0)
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
