// Tiled topological sweep of the accumulating router on gfx950, over a batch of catchments.
//
// What is computed: LUE's accumulating router (reference: source/framework/algorithm/include/lue/
// framework/algorithm/definition/accumulating_router.hpp:525-1209) with one of four cell functors
// (accu.hpp, accu_threshold.hpp, accu_fraction.hpp, kinematic_wave.hpp). A cell is solved once all of
// its upstream neighbours are solved; its value is a function of the sum of their outflows and of its
// own inputs.
//
// How. Every catchment (an independent raster) is cut into TW x TH tiles; the tiles of all catchments
// of a batch are numbered consecutively. One launch ("round") processes a set of tiles; a workgroup
// owns one tile at a time and
//   1. stages the tile's ldd with a one-cell halo, derives per cell the mask of upstream neighbours
//      and stages the inputs of the tile's pending cells in LDS (coalesced loads);
//   2. counts, per pending cell, the upstream neighbours that are not solved yet (neighbours in other
//      tiles count as solved when their border stamp is from an earlier round);
//   3. walks: every lane starts at each of its initially ready cells and continues downstream while
//      it is the last upstream contributor of the next cell (LDS atomics, workgroup scope). The cell
//      value gathers the upstream outflows in fixed neighbour order (N, NE, E, SE, S, SW, W, NW), so
//      results do not depend on which lane arrived last;
//   4. writes the solved cells' outputs coalesced, a pending bitmap and border stamps, and, when the
//      schedule is being recorded, queues the neighbouring tiles that received flow for the next round.
// Which tiles take part in which round depends on the ldd only, so a recorded schedule (per-round tile
// lists) is replayed by later operations on the same network without queueing. Tiles only exchange
// data across launches (kernel boundaries), so no device-scope fence runs inside a kernel.
//
// Precision. T = double: plain FP64. T = float: outflows carry a compensation term (out + lo is the
// FP64 sum rounded to FP32 pairs), sums are formed in FP64, so FP32 rounding does not accumulate along
// long flow paths. Kinematic wave: FP32 Newton + one FP64 correction (kinematic_wave.hpp).
#pragma once

#include "block_solve.hpp"
#include "cells.hpp"
#include "kinematic_wave.hpp"
#include "routing_common.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <limits>
#include <type_traits>

namespace lue_amd {

    // LDS of the schedule inside a tile: the walks' counters, links and ready queue.
    template<int N>
    struct TileScheduleLds
    {
        std::uint32_t count[N / 4];  // four 8-bit counters per word
        std::int16_t next[N];        // downstream cell inside the tile, -1 if none
        std::uint16_t queue[N];      // initially ready cells
    };

    template<int OP, typename T, int TW, int TH>
    struct SweepLds
    {
        using G = Tile<TW, TH>;
        static constexpr bool compensated = std::is_same_v<T, float>;
        static constexpr bool two_outputs = is_split_op(OP);

        CellInput<OP, T> in[G::n];
        T hi[G::n];
        float lo[compensated ? G::n : 1];
        T rem[two_outputs ? G::n : 1];
        TileScheduleLds<G::n> sched;
        std::uint8_t ldd[G::hn];
        alignas(4) std::uint8_t upmask[G::n];  // bit k: neighbour k drains into the cell (set by word-wide ORs)
        std::uint8_t state[G::n];  // see state_* below
        unsigned int solved;
        unsigned int no_data;
        unsigned int queued;
        unsigned int taken;
    };


    // The cell at tile coordinates (lr, lc) (the halo ring included, -1 .. TW / TH) with ldd `code` drains into
    // its downstream neighbour t: if t is a cell of the tile and not no-data, set bit opposite(direction) of
    // t's upstream mask (an LDS OR on the word holding four masks).
    template<int TW, int TH>
    __device__ inline void mark_upstream(std::uint8_t const* ldd, std::uint8_t* upmask, int lr, int lc,
                                         std::uint8_t code)
    {
        constexpr int hw = TW + 2;
        int const dir = ldd_direction(code);
        if (dir < 0)
            return;
        int const tr = lr + neighbour_dr(dir);
        int const tc = lc + neighbour_dc(dir);
        if (tr < 0 || tr >= TH || tc < 0 || tc >= TW || ldd[(tr + 1) * hw + (tc + 1)] == no_data_ldd)
            return;
        int const t = tr * TW + tc;
        atomicOr(reinterpret_cast<unsigned int*>(upmask) + (t >> 2), (1u << opposite(dir)) << (8 * (t & 3)));
    }

    // Offset of element (lr, lc) of a tile from the tile's first element, in a row-major raster of `cols`
    // columns: one 32 x 32 -> 64-bit multiply-add per lane.
    __device__ inline std::int64_t tile_offset(std::uint32_t lr, std::uint32_t cols, std::uint32_t lc)
    {
        return static_cast<std::int64_t>(static_cast<std::uint64_t>(lr) * cols + lc);
    }

    // Cell states inside a round.
    constexpr std::uint8_t state_absent = 0;   // outside the array or ldd no-data
    constexpr std::uint8_t state_pending = 1;  // not solved yet
    constexpr std::uint8_t state_solved = 2;   // solved in an earlier round (value in HBM)
    constexpr std::uint8_t state_new = 3;      // solved in this round (value in LDS)

    // One round over one tile. All threads of the workgroup call it.
    template<int OP, typename T, int TW, int TH>
    __device__ void sweep_tile(SweepArgs<T> const& a, SweepLds<OP, T, TW, TH>& s, std::int64_t tile,
                               std::uint32_t round)
    {
        using G = Tile<TW, TH>;
        constexpr bool compensated = SweepLds<OP, T, TW, TH>::compensated;
        constexpr bool two_outputs = SweepLds<OP, T, TW, TH>::two_outputs;
        constexpr int cells_per_thread = (G::n + sweep_threads - 1) / sweep_threads;

        int const tid = threadIdx.x;
        constexpr int nthreads = sweep_threads;  // every tile kernel is launched with sweep_threads threads
        Catchment cat;
        CatchmentIO<T> io;
        tile_catchment(a, tile, cat, io);
        auto const g_ldd = as_global(cat.ldd);
        auto const g_out = as_global(io.out);
        auto const g_rem = as_global(io.rem);
        // tile coordinates in the catchment (32-bit division: a catchment has fewer than 2^32 tiles)
        std::uint32_t const local = static_cast<std::uint32_t>(tile - cat.tile_base);
        std::uint32_t const tiles_x = static_cast<std::uint32_t>(cat.tiles_x);
        int const ty = static_cast<int>(local / tiles_x);
        int const tx = static_cast<int>(local - static_cast<std::uint32_t>(ty) * tiles_x);
        std::int64_t const rows = cat.rows;
        std::int64_t const cols = cat.cols;
        std::int64_t const r0 = static_cast<std::int64_t>(ty) * TH;
        std::int64_t const c0 = static_cast<std::int64_t>(tx) * TW;
        // 32-bit coordinates (a network's rasters have at most 2^31 rows and columns and 2^32 cells): a row
        // above the array wraps to a large unsigned value, so one unsigned compare per axis tests bounds; lanes
        // address the tile's elements from its first one (`origin`) with tile_offset
        std::uint32_t const rows32 = static_cast<std::uint32_t>(rows);
        std::uint32_t const cols32 = static_cast<std::uint32_t>(cols);
        std::uint32_t const r0u = static_cast<std::uint32_t>(ty) * TH;
        std::uint32_t const c0u = static_cast<std::uint32_t>(tx) * TW;
        std::int64_t const origin = r0 * cols + c0;
        auto const pbits = as_global(a.pending_bits + tile * G::words);
        auto const stamps = as_global(a.border_stamp + tile * G::perimeter);
        auto const out_lo = as_global(compensated ? a.out_lo + cat.cell_base : nullptr);

        if (tid == 0)
        {
            s.solved = 0;
            s.no_data = 0;
            s.queued = 0;
            s.taken = 0;
        }

        // 1. ldd with halo (cells outside the array read as no-data); in round 0 also the raw inputs of every
        // cell. All loads of a thread are issued before the first one is used.
        constexpr int halo_per_thread = (G::hn + sweep_threads - 1) / sweep_threads;
        RawInput<OP, T> raw[cells_per_thread];
        if (round == 0)
        {
#pragma unroll
            for (int j = 0; j < cells_per_thread; ++j)
            {
                int const i = tid + j * nthreads;
                std::uint32_t const lr = static_cast<std::uint32_t>(i / TW);
                std::uint32_t const lc = static_cast<std::uint32_t>(i % TW);
                if (i < G::n && r0u + lr < rows32 && c0u + lc < cols32)
                    load_raw<OP, T>(a, io, origin + tile_offset(lr, cols32, lc), raw[j]);
            }
        }
        std::uint8_t halo[halo_per_thread];
        std::int64_t const halo_origin = origin - cols - 1;  // element (-1, -1) of the tile (may lie outside)
#pragma unroll
        for (int j = 0; j < halo_per_thread; ++j)
        {
            int const i = tid + j * nthreads;
            std::uint32_t const hr = static_cast<std::uint32_t>(i / G::hw);
            std::uint32_t const hc = static_cast<std::uint32_t>(i % G::hw);
            halo[j] = (i < G::hn && r0u + hr - 1u < rows32 && c0u + hc - 1u < cols32)
                          ? g_ldd[halo_origin + tile_offset(hr, cols32, hc)]
                          : no_data_ldd;
        }
#pragma unroll
        for (int j = 0; j < halo_per_thread; ++j)
        {
            int const i = tid + j * nthreads;
            if (i < G::hn)
                s.ldd[i] = halo[j];
        }
        for (int w = tid; w < G::n / 4; w += nthreads)
            reinterpret_cast<std::uint32_t*>(s.upmask)[w] = 0u;
        __syncthreads();

        // 2. state, upstream mask and staged inputs of pending cells
#pragma unroll
        for (int j = 0; j < cells_per_thread; ++j)
        {
            int const i = tid + j * nthreads;
            if (i >= G::n)
                continue;
            int const lr = i / TW;
            int const lc = i % TW;
            int const h = (lr + 1) * G::hw + (lc + 1);
            std::uint8_t const code = s.ldd[h];
            std::uint8_t st;
            if (code == no_data_ldd)
            {
                st = state_absent;
                if (round == 0 && r0u + static_cast<std::uint32_t>(lr) < rows32 &&
                    c0u + static_cast<std::uint32_t>(lc) < cols32)
                {
                    // mark_no_data (accumulating_router.hpp:866-872)
                    std::int64_t const g = origin + tile_offset(lr, cols32, lc);
                    g_out[g] = quiet_nan<T>();
                    if constexpr (two_outputs)
                    {
                        g_rem[g] = quiet_nan<T>();
                    }
                    atomicAdd(&s.no_data, 1u);
                }
            }
            else
            {
                bool const pending = round == 0 || ((pbits[i >> 5] >> (i & 31)) & 1u);
                st = pending ? state_pending : state_solved;
                mark_upstream<TW, TH>(s.ldd, s.upmask, lr, lc, code);
                if (round == 0)
                {
                    stage_raw<OP, T>(a, raw[j], s.in[i], io.dx == nullptr);
                }
                else if (pending)
                {
                    std::int64_t const g = (r0 + lr) * cols + (c0 + lc);
                    stage_input<OP, T>(a, io, g, s.in[i]);
                }
            }
            s.state[i] = st;
            int next = -1;
            if (st != state_absent)
            {
                int const dir = ldd_direction(code);
                if (dir >= 0)
                {
                    int const dr = lr + neighbour_dr(dir);
                    int const dc = lc + neighbour_dc(dir);
                    if (dr >= 0 && dr < TH && dc >= 0 && dc < TW &&
                        s.ldd[(dr + 1) * G::hw + (dc + 1)] != no_data_ldd)
                        next = dr * TW + dc;
                }
            }
            s.sched.next[i] = static_cast<std::int16_t>(next);
        }
        // cells of the halo ring (neighbouring tiles) that drain into this tile
        constexpr int ring = 2 * (TW + 2) + 2 * TH;
        for (int q = tid; q < ring; q += nthreads)
        {
            int hr;
            int hc;
            if (q < TW + 2)
            {
                hr = 0;
                hc = q;
            }
            else if (q < 2 * (TW + 2))
            {
                hr = TH + 1;
                hc = q - (TW + 2);
            }
            else
            {
                int const side = q - 2 * (TW + 2);
                hr = 1 + (side >> 1);
                hc = (side & 1) ? TW + 1 : 0;
            }
            mark_upstream<TW, TH>(s.ldd, s.upmask, hr - 1, hc - 1, s.ldd[hr * G::hw + hc]);
        }
        __syncthreads();

        // 3. counts of unsolved upstream neighbours; the initially ready cells are compacted into an LDS
        // queue (wavefront ballot + prefix count, one LDS atomic per wavefront)
        int const lane = tid & 63;
        std::uint64_t const lanes_below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        for (int j = 0; j < cells_per_thread; ++j)
        {
            int const i = tid + j * nthreads;
            bool is_ready = false;
            if (i < G::n)
            {
                std::uint32_t n = 0;
                if (s.state[i] == state_pending)
                {
                    int const lr = i / TW;
                    int const lc = i % TW;
                    std::uint8_t const mask = s.upmask[i];
                    for (int k = 0; k < 8; ++k)
                    {
                        if (!((mask >> k) & 1u))
                            continue;
                        int const ur = lr + neighbour_dr(k);
                        int const uc = lc + neighbour_dc(k);
                        if (ur >= 0 && ur < TH && uc >= 0 && uc < TW)
                        {
                            n += s.state[ur * TW + uc] == state_pending ? 1u : 0u;
                        }
                        else if (round == 0)
                        {
                            ++n;
                        }
                        else
                        {
                            // upstream neighbour in another tile: solved when stamped in an earlier round
                            int const nty = ty + (ur < 0 ? -1 : (ur >= TH ? 1 : 0));
                            int const ntx = tx + (uc < 0 ? -1 : (uc >= TW ? 1 : 0));
                            int const nlr = ur < 0 ? TH - 1 : (ur >= TH ? 0 : ur);
                            int const nlc = uc < 0 ? TW - 1 : (uc >= TW ? 0 : uc);
                            std::int64_t const ntile =
                                cat.tile_base + static_cast<std::int64_t>(nty) * cat.tiles_x + ntx;
                            std::uint32_t const st =
                                as_global(a.border_stamp)[ntile * G::perimeter + G::perimeter_index(nlr, nlc)];
                            n += st < round ? 0u : 1u;
                        }
                    }
                    is_ready = n == 0;
                }
                // 8-bit counter, written as a byte of the packed word
                reinterpret_cast<std::uint8_t*>(s.sched.count)[i] = static_cast<std::uint8_t>(n);
            }
            std::uint64_t const ballot = __ballot(is_ready);
            if (ballot)
            {
                unsigned int base = 0;
                if (lane == 0)
                    base = atomicAdd(&s.queued, static_cast<unsigned int>(__popcll(ballot)));
                base = __shfl(base, 0, 64);
                if (is_ready)
                    s.sched.queue[base + __popcll(ballot & lanes_below)] = static_cast<std::uint16_t>(i);
            }
        }
        __syncthreads();

        // Sum (FP64, neighbour order) of the outflow of the upstream neighbours of tile cell i: solved in
        // this round (LDS) or earlier (HBM).
        auto gather = [&](int i, double& upstream, bool& poisoned)
        {
            int const lr = i / TW;
            int const lc = i % TW;
            std::uint8_t const mask = s.upmask[i];
            upstream = 0.0;
            poisoned = false;
            for (int k = 0; k < 8; ++k)
            {
                if (!((mask >> k) & 1u))
                    continue;
                int const ur = lr + neighbour_dr(k);
                int const uc = lc + neighbour_dc(k);
                bool const inside = ur >= 0 && ur < TH && uc >= 0 && uc < TW;
                double v;
                // in round 0 every upstream cell of a ready cell lies in the tile and was solved now
                if (inside && (round == 0 || s.state[ur * TW + uc] == state_new))
                {
                    int const u = ur * TW + uc;
                    v = static_cast<double>(s.hi[u]);
                    if constexpr (compensated)
                        v += static_cast<double>(s.lo[u]);
                }
                else
                {
                    std::int64_t const ug = (r0 + ur) * cols + (c0 + uc);
                    v = static_cast<double>(g_out[ug]);
                    if constexpr (compensated)
                        v += static_cast<double>(out_lo[ug]);
                }
                if (isnan(v))
                    poisoned = true;
                else
                    upstream += v;
            }
        };

        // 4. walks: lanes take initially ready cells from the queue and walk downstream while they are the
        // last upstream contributor of the next cell
        unsigned int const queued = s.queued;
        while (true)
        {
            unsigned int const item = atomicAdd(&s.taken, 1u);
            if (item >= queued)
                break;
            int i = s.sched.queue[item];
            while (true)
            {
                int const next = s.sched.next[i];
                double upstream;
                bool poisoned;
                gather(i, upstream, poisoned);
                double out;
                T rem{};
                solve_cell<OP, T>(a, s.in[i], upstream, poisoned, out, rem);
                T const hi = static_cast<T>(out);
                s.hi[i] = hi;
                if constexpr (compensated)
                    s.lo[i] = isnan(out) ? 0.0f : static_cast<float>(out - static_cast<double>(hi));
                if constexpr (two_outputs)
                    s.rem[i] = rem;

                // publish: the state write is ordered before the counter update (release)
                __hip_atomic_store(&s.state[i], state_new, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (next < 0)
                    break;
                std::uint32_t const shift = 8u * static_cast<std::uint32_t>(next & 3);
                std::uint32_t const old = __hip_atomic_fetch_sub(
                    &s.sched.count[next >> 2], 1u << shift, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (((old >> shift) & 0xFFu) != 1u)
                    break;  // confluence: another stream still has to arrive
                i = next;
            }
        }
        __syncthreads();


        // 5. write back: outputs of cells solved in this round (coalesced), compensation terms that later
        // rounds read, pending bitmap, border stamps, and queue the tiles that receive flow
        std::uint32_t const next_slot = (round + 1) % 3;
        unsigned int solved = 0;
#pragma unroll
        for (int i = tid; i < G::n; i += nthreads)
        {
            std::uint8_t const st = s.state[i];
            int const lr = i / TW;
            int const lc = i % TW;
            // border stamps serve the later rounds of the rounds schedule only
            int const p = G::perimeter_index(lr, lc);
            if (st == state_new)
            {
                ++solved;
                std::int64_t const g = origin + tile_offset(lr, cols32, lc);
                g_out[g] = s.hi[i];
                if constexpr (two_outputs)
                    g_rem[g] = s.rem[i];
                int const dir = ldd_direction(s.ldd[(lr + 1) * G::hw + (lc + 1)]);
                if (dir >= 0)
                {
                    int const dr = lr + neighbour_dr(dir);
                    int const dc = lc + neighbour_dc(dir);
                    bool const inside = dr >= 0 && dr < TH && dc >= 0 && dc < TW;
                    bool const target_ok = s.ldd[(dr + 1) * G::hw + (dc + 1)] != no_data_ldd;
                    if (target_ok && (!inside || s.state[dr * TW + dc] == state_pending))
                    {
                        // read by a later round
                        if constexpr (compensated)
                            out_lo[g] = s.lo[i];
                        if (!inside && a.record)
                        {
                            int const nty = ty + (dr < 0 ? -1 : (dr >= TH ? 1 : 0));
                            int const ntx = tx + (dc < 0 ? -1 : (dc >= TW ? 1 : 0));
                            std::int64_t const ntile = cat.tile_base + static_cast<std::int64_t>(nty) * cat.tiles_x + ntx;
                            std::uint32_t* flag = a.tile_flag + next_slot * a.ntiles + ntile;
                            if (atomicExch(flag, 1u) == 0u)
                            {
                                std::uint32_t const slot = atomicAdd(&a.work_count[next_slot], 1u);
                                a.worklist[next_slot * a.ntiles + slot] = static_cast<std::uint32_t>(ntile);
                            }
                        }
                    }
                }
                if (p >= 0)
                    stamps[p] = round;
            }
            else if (round == 0 && p >= 0)
            {
                stamps[p] = st == state_pending ? stamp_pending : 0u;
            }
        }
        // pending bitmap: one word per 32 cells; a wavefront's 64 consecutive cells give two words (ballot)
        static_assert(G::n % sweep_threads == 0, "tiles hold a multiple of sweep_threads cells");
#pragma unroll
        for (int j = 0; j < cells_per_thread; ++j)
        {
            int const i = tid + j * nthreads;
            std::uint64_t const bits = __ballot(s.state[i] == state_pending);
            if ((tid & 63) == 0)
            {
                int const w = i >> 5;
                pbits[w] = static_cast<std::uint32_t>(bits);
                pbits[w + 1] = static_cast<std::uint32_t>(bits >> 32);
            }
        }
        if (a.record)
        {
            atomicAdd(&s.solved, solved);
            __syncthreads();
            if (tid == 0)
            {
                atomicAdd(&a.counters[0], static_cast<unsigned long long>(s.solved));
                if (round == 0)
                    atomicAdd(&a.counters[1], static_cast<unsigned long long>(s.no_data));
            }
        }
        __syncthreads();
    }

    // Workgroups of a tile kernel that fit one CU (160 KB of LDS, 32 wavefront slots), and the wavefronts per
    // SIMD that follow: the register budget the kernels are compiled for.
    template<int OP, typename T, int TW, int TH>
    constexpr int sweep_workgroups_per_cu()
    {
        int const by_lds = static_cast<int>((160 * 1024) / sizeof(SweepLds<OP, T, TW, TH>));
        int const by_waves = 32 / sweep_waves;
        int const n = by_lds < by_waves ? by_lds : by_waves;
        return n < 1 ? 1 : n;
    }

    template<int OP, typename T, int TW, int TH>
    constexpr int sweep_waves_per_simd()
    {
        return sweep_workgroups_per_cu<OP, T, TW, TH>() * sweep_waves / 4;
    }

    // Round 0 with walks: every tile of the batch, one workgroup per tile (grid = number of tiles). It records
    // the schedule, and runs round 0 of the rounds schedule; replays of the chain schedule run the recorded
    // tile programs instead (tile_pass.hpp).
    template<int OP, typename T, int TW, int TH>
    __global__ void __launch_bounds__(sweep_threads, (sweep_waves_per_simd<OP, T, TW, TH>()))
        sweep_first_round(SweepArgs<T> a)
    {
        __shared__ SweepLds<OP, T, TW, TH> s;
        sweep_tile<OP, T, TW, TH>(a, s, blockIdx.x, 0u);
    }

    // Round r >= 1 while recording: the tiles queued by round r - 1. Workgroup b takes items b, b + grid, ...
    template<int OP, typename T, int TW, int TH>
    __global__ void __launch_bounds__(sweep_threads, (sweep_waves_per_simd<OP, T, TW, TH>())) sweep_next_round(SweepArgs<T> a, std::uint32_t round)
    {
        __shared__ SweepLds<OP, T, TW, TH> s;
        std::uint32_t const slot = round % 3;
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            // the slot that round + 1 fills was last read by round - 1
            a.work_count[(round + 2) % 3] = 0u;
        }
        std::uint32_t const count = a.work_count[slot];
        for (std::uint32_t item = blockIdx.x; item < count; item += gridDim.x)
        {
            std::int64_t const tile = a.worklist[slot * a.ntiles + item];
            if (threadIdx.x == 0)
                a.tile_flag[slot * a.ntiles + tile] = 0u;
            sweep_tile<OP, T, TW, TH>(a, s, tile, round);
        }
    }

    // Round r >= 1 replayed from a recorded tile list.
    template<int OP, typename T, int TW, int TH>
    __global__ void __launch_bounds__(sweep_threads, (sweep_waves_per_simd<OP, T, TW, TH>())) sweep_list_round(SweepArgs<T> a, std::uint32_t const* tiles,
                                                            std::uint32_t count, std::uint32_t round)
    {
        __shared__ SweepLds<OP, T, TW, TH> s;
        for (std::uint32_t item = blockIdx.x; item < count; item += gridDim.x)
        {
            sweep_tile<OP, T, TW, TH>(a, s, tiles[item], round);
        }
    }

}  // namespace lue_amd
