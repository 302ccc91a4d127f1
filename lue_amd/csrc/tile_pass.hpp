// Round 0 of a replay under the chain schedule (the "tile pass"): every tile solves its local cells (the cells
// whose whole upstream area lies in the tile) on the in-tile chain layout recorded for the flow network, writes
// their outputs, and stages the inputs of its pending cells at their chain positions for the chain phase
// (chains.hpp).
//
// Reference: the intra-partition phase of LUE's accumulating router
// (source/framework/algorithm/include/lue/framework/algorithm/definition/accumulating_router.hpp:789-899),
// with a tile in the role of a partition.
//
// Everything that depends on the ldd alone is part of the recorded program (tile_layout.hpp): per layout
// position the local cell, its chain start and its light children (upstream neighbours other than the chain
// predecessor); per tile the pending bitmap and the list of its local cells that drain into another tile, each
// with the side slot it feeds there; per pending cell its local index and the mask of its upstream neighbours
// that the tile solves. A replay therefore reads no ldd and derives nothing from it: per cell it loads the raw
// inputs and 3 bytes of layout, per pending cell 8 bytes of staging entries, and spends its instructions on
// the cell functions.
#pragma once

#include "block_solve.hpp"
#include "cells.hpp"
#include "routing_common.hpp"
#include "tile_layout.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace lue_amd {

    template<int OP, typename T, int TW, int TH>
    struct TilePassLds
    {
        using G = Tile<TW, TH>;
        static constexpr bool compensated = std::is_same_v<T, float>;
        static constexpr bool two_outputs = is_split_op(OP);

        CellInput<OP, T> in[G::n];
        T hi[G::n];                          // outflow (FP32 path: the FP64 value rounded, lo the rest)
        float lo[compensated ? G::n : 1];
        T rem[two_outputs ? G::n : 1];
        alignas(8) std::uint16_t order[G::n];  // local cell at each position | order_start_bit
        alignas(4) std::uint8_t light[G::n];   // light children of the cell at each position
    };

    // Offset of neighbour k of a tile cell in the tile's row-major cell index.
    template<int TW>
    __device__ inline int tile_neighbour_offset(int k)
    {
        return neighbour_dr(k) * TW + neighbour_dc(k);
    }

    // FP64 sum, in neighbour order, of the outflows of the tile cells in `mask` around cell i.
    template<int OP, typename T, int TW, int TH>
    __device__ inline double tile_neighbour_sum(TilePassLds<OP, T, TW, TH> const& s, int i, unsigned mask)
    {
        double sum = 0.0;
        while (mask != 0u)
        {
            int const k = __builtin_ctz(mask);
            mask &= mask - 1u;
            int const u = (i + tile_neighbour_offset<TW>(k)) & (TW * TH - 1);  // in bounds whatever the record
            double v = static_cast<double>(s.hi[u]);
            if constexpr (TilePassLds<OP, T, TW, TH>::compensated)
                v += static_cast<double>(s.lo[u]);
            sum += v;
        }
        return sum;
    }

    template<int OP, typename T, int TW, int TH>
    constexpr int tile_pass_workgroups_per_cu()
    {
        int const by_lds = static_cast<int>((160 * 1024) / sizeof(TilePassLds<OP, T, TW, TH>));
        int const by_waves = 32 / sweep_waves;
        int const n = by_lds < by_waves ? by_lds : by_waves;
        return n < 1 ? 1 : n;
    }

    template<int OP, typename T, int TW, int TH>
    constexpr int tile_pass_waves_per_simd()
    {
        return tile_pass_workgroups_per_cu<OP, T, TW, TH>() * sweep_waves / 4;
    }

    // One tile; all sweep_threads threads of the workgroup call it.
    template<int OP, typename T, int TW, int TH>
    __device__ inline void tile_pass_tile(SweepArgs<T> const& a, TilePassLds<OP, T, TW, TH>& s, std::int64_t tile)
    {
        using G = Tile<TW, TH>;
        using Lds = TilePassLds<OP, T, TW, TH>;
        constexpr bool compensated = Lds::compensated;
        constexpr bool two_outputs = Lds::two_outputs;
        constexpr int nthreads = sweep_threads;
        constexpr int per_thread = G::n / nthreads;
        static_assert(G::n % nthreads == 0 && (G::n & (G::n - 1)) == 0 && G::n <= 4096,
                      "tiles of a power of two cells, a multiple of the workgroup, at most 4096");

        int const tid = threadIdx.x;
        int const lane_id = tid & 63;
        Catchment cat;
        CatchmentIO<T> io;
        tile_catchment(a, tile, cat, io);
        auto const g_out = as_global(io.out);
        auto const g_rem = as_global(io.rem);
        // tile coordinates (32-bit: a catchment has fewer than 2^32 tiles, at most 2^31 rows and columns); a
        // lane addresses the tile's elements from its first one with tile_offset
        std::uint32_t const local = static_cast<std::uint32_t>(tile - cat.tile_base);
        std::uint32_t const tiles_x = static_cast<std::uint32_t>(cat.tiles_x);
        std::uint32_t const ty = local / tiles_x;
        std::uint32_t const tx = local - ty * tiles_x;
        std::uint32_t const rows32 = static_cast<std::uint32_t>(cat.rows);
        std::uint32_t const cols32 = static_cast<std::uint32_t>(cat.cols);
        std::uint32_t const r0u = ty * TH;
        std::uint32_t const c0u = tx * TW;
        std::int64_t const origin = static_cast<std::int64_t>(r0u) * cat.cols + c0u;

        // 1. loads: staging range, level bounds (lane l: level l of this wavefront), layout, raw inputs
        std::uint32_t const stage_begin = as_global(a.stage_off)[tile];
        std::uint32_t const stage_end = as_global(a.stage_off)[tile + 1];
        std::uint32_t const lo_begin = as_global(a.lo_off)[tile];
        std::uint32_t const lo_end = as_global(a.lo_off)[tile + 1];
        int split_lo = 0;
        int split_hi = 0;
        {
            int const wave = tid >> 6;
            auto const split = as_global(a.tile_split + tile * tile_split_entries);
            if (lane_id < tile_max_levels)
            {
                split_lo = split[lane_id * (sweep_waves + 1) + wave];
                split_hi = split[lane_id * (sweep_waves + 1) + wave + 1];
            }
        }
        int const levels = min(static_cast<int>(as_global(a.tile_nlevels)[tile]), tile_max_levels);
        auto const order = as_global(a.tile_order + tile * G::n);
        auto const light = as_global(a.tile_flags + tile * G::n);
        RawInput<OP, T> raw[per_thread];
#pragma unroll
        for (int j = 0; j < per_thread; ++j)
        {
            int const i = tid + j * nthreads;
            std::uint32_t const lr = static_cast<std::uint32_t>(i / TW);
            std::uint32_t const lc = static_cast<std::uint32_t>(i % TW);
            if (r0u + lr < rows32 && c0u + lc < cols32)
                load_raw<OP, T>(a, io, origin + tile_offset(lr, cols32, lc), raw[j]);
            else
                raw[j] = RawInput<OP, T>{};  // outside the array: staged but never read
        }
        if constexpr (per_thread == 4)
        {
            // positions 4 tid .. 4 tid + 3: one 8-byte and one 4-byte load (a tile's layout starts at a multiple
            // of 2 n bytes)
            std::uint64_t const o4 = reinterpret_cast<global_ptr<std::uint64_t const>>(order)[tid];
            std::uint32_t const f4 = reinterpret_cast<global_ptr<std::uint32_t const>>(light)[tid];
            reinterpret_cast<std::uint64_t*>(s.order)[tid] = o4;
            reinterpret_cast<std::uint32_t*>(s.light)[tid] = f4;
        }
        else
        {
            std::uint16_t o[per_thread];
            std::uint8_t f[per_thread];
#pragma unroll
            for (int j = 0; j < per_thread; ++j)
            {
                o[j] = order[tid + j * nthreads];
                f[j] = light[tid + j * nthreads];
            }
#pragma unroll
            for (int j = 0; j < per_thread; ++j)
            {
                s.order[tid + j * nthreads] = o[j];
                s.light[tid + j * nthreads] = f[j];
            }
        }
        // the tile's pending bitmap words of this thread's cells, its staging entries and its side-slot entry,
        // all in flight with the inputs (used after the levels)
        std::uint32_t const npending = stage_end - stage_begin;
        bool const feeds = lo_begin + static_cast<std::uint32_t>(tid) < lo_end;
        std::uint64_t const lo_entry = feeds ? as_global(a.lo_entry)[lo_begin + tid] : 0ull;
        std::uint32_t cell_of[per_thread];
        std::uint32_t pos_of[per_thread];
        std::uint32_t pword[per_thread];
#pragma unroll
        for (int j = 0; j < per_thread; ++j)
        {
            std::uint32_t const rank = static_cast<std::uint32_t>(tid + j * nthreads);
            cell_of[j] = rank < npending ? as_global(a.stage_cell)[stage_begin + rank] : 0u;
            pos_of[j] = rank < npending ? as_global(a.stage_pos)[stage_begin + rank] : 0u;
            pword[j] = as_global(a.pending_bits)[tile * G::words + ((tid + j * nthreads) >> 5)];
        }
        bool const dx_scalar = io.dx == nullptr;
#pragma unroll
        for (int j = 0; j < per_thread; ++j)
        {
            int const i = tid + j * nthreads;
            // cells the layout does not solve keep NaN: those of no-data ldd cells are written as such
            // (mark_no_data, accumulating_router.hpp:866-872), pending cells are not written here
            s.hi[i] = quiet_nan<T>();
            if constexpr (two_outputs)
                s.rem[i] = quiet_nan<T>();
            stage_raw<OP, T>(a, raw[j], s.in[i], dx_scalar);
        }
        __syncthreads();

        // 2. levels of the in-tile chains: wavefront w solves the positions [split[l][w], split[l][w + 1]) in
        // blocks of 64; the light children of a chain lie in lower levels, the heavy child is the previous
        // position of the chain
        for (int l = 0; l < levels; ++l)
        {
            int const lo = __builtin_amdgcn_readlane(split_lo, l);
            int const hi = __builtin_amdgcn_readlane(split_hi, l);
            BlockState st;
            for (int blk = lo; blk < hi; blk += 64)
            {
                int const p = blk + lane_id;
                int const nb = hi - blk < 64 ? hi - blk : 64;
                ChainLane<OP, T> d{};
                int i = 0;
                d.valid = p < hi;
                if (d.valid)
                {
                    std::uint16_t const o = s.order[p & (G::n - 1)];
                    i = o & (G::n - 1);
                    d.start = (o & order_start_bit) != 0;
                    d.extra = tile_neighbour_sum<OP, T, TW, TH>(s, i, s.light[p & (G::n - 1)]);
                    d.in = s.in[i];
                }
                double value;
                T rem{};
                solve_block<OP, T>(a, d, nb, st, value, rem);
                if (d.valid)
                {
                    T const h = static_cast<T>(value);
                    s.hi[i] = h;
                    if constexpr (compensated)
                        s.lo[i] = isnan(value) ? 0.0f : static_cast<float>(value - static_cast<double>(h));
                    if constexpr (two_outputs)
                        s.rem[i] = rem;
                }
            }
            __syncthreads();
        }

        // 3. write-back of every cell of the array that is not pending (local cells, and NaN for no-data
        // ldd cells), coalesced; the FP64 outflows of the local cells that feed side slots in other tiles (at
        // most one per perimeter cell: one per thread)
        static_assert(G::perimeter <= nthreads, "one side-slot source per thread");
#pragma unroll
        for (int j = 0; j < per_thread; ++j)
        {
            int const i = tid + j * nthreads;
            std::uint32_t const lr = static_cast<std::uint32_t>(i / TW);
            std::uint32_t const lc = static_cast<std::uint32_t>(i % TW);
            std::uint32_t const bit = 1u << (i & 31);
            if (r0u + lr < rows32 && c0u + lc < cols32 && (pword[j] & bit) == 0u)
            {
                std::int64_t const g = origin + tile_offset(lr, cols32, lc);
                g_out[g] = s.hi[i];
                if constexpr (two_outputs)
                    g_rem[g] = s.rem[i];
            }
        }
        if (feeds)
        {
            int const i = static_cast<int>(lo_entry & (G::n - 1));
            double v = static_cast<double>(s.hi[i]);
            if constexpr (compensated)
                v += static_cast<double>(s.lo[i]);
            as_global(a.slot)[lo_entry >> 16] = v;
        }

        // 4. staging of the pending cells, entry k by thread k mod nthreads (entries in chain-position order, so
        // that a wavefront's stores cover runs of consecutive positions): inputs and the FP64 sum (neighbour
        // order) of their upstream cells that this tile solved, at their chain positions (accu: that sum plus
        // the cell's inflow, its only input)
#pragma unroll
        for (int j = 0; j < per_thread; ++j)
        {
            std::uint32_t const rank = static_cast<std::uint32_t>(tid + j * nthreads);
            if (rank >= npending)
                break;
            int const i = static_cast<int>(cell_of[j] & (G::n - 1));
            double const side = tile_neighbour_sum<OP, T, TW, TH>(s, i, (cell_of[j] >> 16) & 0xFFu);
            if constexpr (OP == op_accu)
            {
                // accu adds the cell's own inflow to the side sum (a NaN inflow makes it NaN): one FP64 value
                // per pending cell, nothing else staged
                as_global(a.stage_side)[pos_of[j]] = side + static_cast<double>(s.in[i].x);
            }
            else
            {
                store_staged<OP, T>(a, pos_of[j], s.in[i]);
                as_global(a.stage_side)[pos_of[j]] = side;
            }
        }
    }

    // One workgroup (sweep_threads threads) per tile of the batch; grid = number of tiles.
    template<int OP, typename T, int TW, int TH>
    __global__ void __launch_bounds__(sweep_threads, (tile_pass_waves_per_simd<OP, T, TW, TH>()))
        tile_pass_kernel(SweepArgs<T> a)
    {
        __shared__ TilePassLds<OP, T, TW, TH> s;
        tile_pass_tile<OP, T, TW, TH>(a, s, blockIdx.x);
    }

    // The same as a persistent launch fed by a tile work queue: as many workgroups as the device holds at once,
    // each taking the next tile from `next_tile` (zero at the launch; one atomic per tile) until none is left.
    // Every workgroup reaches the end: a tile index at or past ntiles ends its loop.
    template<int OP, typename T, int TW, int TH>
    __global__ void __launch_bounds__(sweep_threads, (tile_pass_waves_per_simd<OP, T, TW, TH>()))
        tile_pass_queue_kernel(SweepArgs<T> a, std::uint32_t* next_tile)
    {
        __shared__ TilePassLds<OP, T, TW, TH> s;
        __shared__ std::uint32_t taken;
        while (true)
        {
            if (threadIdx.x == 0)
                taken = atomicAdd(next_tile, 1u);
            __syncthreads();
            std::int64_t const tile = taken;
            if (tile >= a.ntiles)
                break;
            tile_pass_tile<OP, T, TW, TH>(a, s, tile);
            __syncthreads();  // the LDS of this tile is read to the end before the next one is staged
        }
    }

}  // namespace lue_amd
