// In-tile chain layout of the first round of the tiled sweep (chain schedule).
//
// Round 0 solves, in every tile, the cells whose whole upstream area lies in the tile ("local" cells;
// the others are left to the chain phase, chains.hpp). Walking them cell by cell costs one dependent
// LDS round trip chain per step. Instead the local cells of a tile are laid out once per flow network as
// heavy-path chains (the same decomposition as the chain phase, restricted to the tile): the heavy child
// of a cell is its local upstream neighbour with the most local cells upstream, a chain follows heavy
// children, and the level of a chain is one more than the highest level of the chains that join it from
// the side (at most log2(TW * TH) + 1 levels). A run of round 0 then solves level after level, each
// wavefront taking a contiguous part of the level in blocks of 64 positions (solve_block, block_solve.hpp).
//
// Recorded per tile: order[p] = local cell index at position p (chains level after level, each in flow
// order) | order_start_bit at the first (most upstream) cell of a chain, flags[p] = the light children of
// that cell (bit k: neighbour k is a local upstream cell other than the chain predecessor), and per level the
// positions of the sweep_waves wavefronts: split[l][w], w = 0..sweep_waves (parts made of whole chains,
// packed so that the level needs few 64-position blocks). tile_program_kernel records the rest of what a
// replay's tile pass needs (tile_pass.hpp), so that a replay reads no ldd.
#pragma once

#include "cells.hpp"
#include "routing_common.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lue_amd {

    constexpr int tile_max_levels = 16;
    constexpr std::uint16_t order_start_bit = 0x8000;  // tiles hold at most 4096 cells (12 index bits)
    constexpr int tile_split_entries = tile_max_levels * (sweep_waves + 1);

    template<int TW, int TH>
    struct TileLayoutLds
    {
        using G = Tile<TW, TH>;
        std::uint8_t ldd[G::hn];
        std::uint8_t local[G::n];        // 1: solved by round 0
        std::int16_t parent[G::n];       // local downstream cell in the tile, -1 if none
        std::int16_t heavy[G::n];        // heavy child, -1 if none
        std::uint8_t heavy_dir[G::n];
        alignas(16) std::uint32_t count[G::n];  // children still to be sized; then the packing's head info
        std::uint32_t size[G::n];        // local cells upstream, including the cell
        alignas(4) std::uint8_t children[G::n];  // bit k: neighbour k is a local child (set by word-wide ORs)
        alignas(16) std::uint16_t heads[G::n];  // chain heads in cell order
        std::uint32_t wave_heads[G::n / 64];  // chain heads per 64 cells (prefix: offsets into heads)
        std::uint16_t queue[G::n];       // leaves to size
        std::int16_t head_of[G::n];
        std::uint16_t depth[G::n];       // distance from the chain head (its most downstream cell)
        std::uint16_t length[G::n];      // per head: cells in the chain
        std::uint32_t level[G::n];       // per head: chain level
        std::uint16_t base[G::n];        // per head: first position of the chain
        std::uint32_t level_cells[tile_max_levels];
        std::uint16_t split[tile_split_entries];
        unsigned int queued;
        unsigned int taken;
        unsigned int changed;
        unsigned int levels;
        unsigned int nheads;
    };

    // One workgroup per tile. `pending_bits`: the cells round 0 leaves pending. On a tile whose levels
    // exceed tile_max_levels (impossible for heavy paths of at most 64 x 64 cells) `error` is set.
    template<int TW, int TH>
    __global__ void __launch_bounds__(256)
        tile_layout_kernel(Catchment const* cats, std::uint32_t const* tile_cat, std::uint32_t const* pending_bits,
                           std::uint16_t* order, std::uint8_t* flags, std::uint16_t* split, std::uint8_t* nlevels,
                           std::uint32_t* error)
    {
        using G = Tile<TW, TH>;
        __shared__ TileLayoutLds<TW, TH> s;
        std::int64_t const tile = blockIdx.x;
        int const tid = threadIdx.x;
        int const nthreads = blockDim.x;
        Catchment const cat = cats[tile_cat[tile]];
        std::int64_t const local_tile = tile - cat.tile_base;
        std::int64_t const r0 = (local_tile / cat.tiles_x) * TH;
        std::int64_t const c0 = (local_tile % cat.tiles_x) * TW;
        auto const ldd = as_global(cat.ldd);
        auto const bits = as_global(pending_bits + tile * G::words);

        if (tid == 0)
        {
            s.queued = 0;
            s.taken = 0;
            s.levels = 0;
        }
        for (int i = tid; i < G::hn; i += nthreads)
        {
            std::int64_t const r = r0 + i / G::hw - 1;
            std::int64_t const c = c0 + i % G::hw - 1;
            s.ldd[i] = (r >= 0 && r < cat.rows && c >= 0 && c < cat.cols) ? ldd[r * cat.cols + c] : no_data_ldd;
        }
        for (int l = tid; l < tile_max_levels; l += nthreads)
            s.level_cells[l] = 0;
        __syncthreads();

        // local cells, their local parent, child counts
        for (int i = tid; i < G::n; i += nthreads)
        {
            int const lr = i / TW;
            int const lc = i % TW;
            std::uint8_t const code = s.ldd[(lr + 1) * G::hw + (lc + 1)];
            bool const pending = (bits[i >> 5] >> (i & 31)) & 1u;
            s.local[i] = (code != no_data_ldd && !pending) ? 1 : 0;
            s.count[i] = 0;
            s.heavy[i] = -1;
            s.head_of[i] = -1;
        }
        for (int w = tid; w < G::n / 4; w += nthreads)
            reinterpret_cast<std::uint32_t*>(s.children)[w] = 0u;
        __syncthreads();
        for (int i = tid; i < G::n; i += nthreads)
        {
            std::int16_t p = -1;
            if (s.local[i])
            {
                int const lr = i / TW;
                int const lc = i % TW;
                int const dir = ldd_direction(s.ldd[(lr + 1) * G::hw + (lc + 1)]);
                if (dir >= 0)
                {
                    int const dr = lr + neighbour_dr(dir);
                    int const dc = lc + neighbour_dc(dir);
                    if (dr >= 0 && dr < TH && dc >= 0 && dc < TW && s.local[dr * TW + dc])
                    {
                        p = static_cast<std::int16_t>(dr * TW + dc);
                        atomicAdd(&s.count[p], 1u);
                        // the parent sees this cell as its neighbour opposite(dir)
                        atomicOr(reinterpret_cast<unsigned int*>(s.children) + (p >> 2),
                                 (1u << opposite(dir)) << (8 * (p & 3)));
                    }
                }
            }
            s.parent[i] = p;
        }
        __syncthreads();

        // subtree sizes: walks from the leaves; the last child to arrive continues downstream
        for (int i = tid; i < G::n; i += nthreads)
        {
            if (s.local[i] && s.count[i] == 0)
                s.queue[atomicAdd(&s.queued, 1u)] = static_cast<std::uint16_t>(i);
        }
        __syncthreads();
        unsigned int const queued = s.queued;
        while (true)
        {
            unsigned int const item = atomicAdd(&s.taken, 1u);
            if (item >= queued)
                break;
            int i = s.queue[item];
            while (true)
            {
                int const lr = i / TW;
                int const lc = i % TW;
                std::uint32_t size = 1;
                std::uint32_t best_size = 0;
                int best = -1;
                int best_dir = 0;
                unsigned m = s.children[i];
                while (m != 0u)  // the local children only, in neighbour order
                {
                    int const k = __builtin_ctz(m);
                    m &= m - 1u;
                    int const u = (lr + neighbour_dr(k)) * TW + (lc + neighbour_dc(k));
                    std::uint32_t const su = __hip_atomic_load(&s.size[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    size += su;
                    if (su > best_size)
                    {
                        best_size = su;
                        best = u;
                        best_dir = k;
                    }
                }
                s.heavy[i] = static_cast<std::int16_t>(best);
                s.heavy_dir[i] = static_cast<std::uint8_t>(best_dir);
                __hip_atomic_store(&s.size[i], size, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                int const p = s.parent[i];
                if (p < 0)
                    break;
                std::uint32_t const old =
                    __hip_atomic_fetch_sub(&s.count[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (old != 1u)
                    break;
                i = p;
            }
        }
        __syncthreads();

        // chains: a head is a local cell that is not the heavy child of its parent; its chain follows heavy
        // children upstream
        for (int i = tid; i < G::n; i += nthreads)
        {
            if (!s.local[i])
                continue;
            int const p = s.parent[i];
            if (p >= 0 && s.heavy[p] == i)
                continue;
            int j = i;
            std::uint16_t d = 0;
            while (j >= 0)
            {
                s.head_of[j] = static_cast<std::int16_t>(i);
                s.depth[j] = d++;
                j = s.heavy[j];
            }
            s.length[i] = d;
            s.level[i] = 0;
            s.base[i] = 0xffff;  // no position unless its level is packed
        }
        __syncthreads();

        // levels: relax level(chain of parent) >= level(chain) + 1 for every head with a parent
        for (int iteration = 0; iteration <= tile_max_levels; ++iteration)
        {
            if (tid == 0)
                s.changed = 0;
            __syncthreads();
            for (int i = tid; i < G::n; i += nthreads)
            {
                if (!s.local[i] || s.head_of[i] != i)
                    continue;
                int const p = s.parent[i];
                if (p < 0)
                    continue;
                int const hp = s.head_of[p];
                std::uint32_t const want = s.level[i] + 1;
                if (s.level[hp] < want && atomicMax(&s.level[hp], want) < want)
                    s.changed = 1;
            }
            __syncthreads();
            if (s.changed == 0)
                break;
            __syncthreads();
        }
        for (int i = tid; i < G::n; i += nthreads)
        {
            if (!s.local[i] || s.head_of[i] != i)
                continue;
            std::uint32_t const l = s.level[i];
            if (l >= static_cast<std::uint32_t>(tile_max_levels))
                atomicAdd(error, 1u);
            else
            {
                atomicAdd(&s.level_cells[l], static_cast<std::uint32_t>(s.length[i]));
                atomicMax(&s.levels, l + 1);
            }
        }
        __syncthreads();

        // positions: levels one after the other. Inside a level, each of the sweep_waves wavefronts of the tile
        // kernel takes a contiguous part made of whole chains and solves it in blocks of 64 positions, so a
        // level costs the sum over the parts of ceil(part / 64) blocks. The level's ceil(cells / 64) blocks
        // are dealt out to the wavefronts as capacities (multiples of 64) and the chains, in cell order, go
        // first-fit into the parts (a chain that fits nowhere goes to the part with the most room left).
        // The heads are first listed in cell order (ballots); levels are independent, so lane l of the first
        // wavefront packs level l, its four parts in registers.
        static_assert(G::n % 64 == 0 && sweep_waves == 4, "tile layout packing assumes 64-cell groups, 4 parts");
        int const lane = tid & 63;
        for (int g = tid >> 6; g < G::n / 64; g += nthreads >> 6)
        {
            int const i = g * 64 + lane;
            bool const head = s.local[i] && s.head_of[i] == i;
            std::uint64_t const ballot = __ballot(head);
            if (lane == 0)
                s.wave_heads[g] = static_cast<std::uint32_t>(__popcll(ballot));
        }
        __syncthreads();
        if (tid < 64)
        {
            // offsets of the 64-cell groups (at most G::n / 64 = 64 groups: one per lane)
            std::uint32_t count = lane < G::n / 64 ? s.wave_heads[lane] : 0u;
            std::uint32_t incl = count;
            for (int d = 1; d < 64; d <<= 1)
            {
                std::uint32_t const up = __shfl_up(incl, d, 64);
                if (lane >= d)
                    incl += up;
            }
            if (lane < G::n / 64)
                s.wave_heads[lane] = incl - count;
            if (lane == G::n / 64 - 1)
                s.nheads = incl;
        }
        __syncthreads();
        for (int g = tid >> 6; g < G::n / 64; g += nthreads >> 6)
        {
            int const i = g * 64 + lane;
            bool const head = s.local[i] && s.head_of[i] == i;
            std::uint64_t const ballot = __ballot(head);
            if (head)
            {
                std::uint64_t const below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
                s.heads[s.wave_heads[g] + __popcll(ballot & below)] = static_cast<std::uint16_t>(i);
            }
        }
        __syncthreads();
        // per head (cell order): its level << 16 | its length, padded to a multiple of 8 entries with a level no
        // lane packs, in `count` (free since the subtree sizes), so that the packing below reads four heads per
        // LDS load, two loads at a time
        std::uint32_t* const hinfo = s.count;
        {
            std::uint32_t const nheads = s.nheads;
            std::uint32_t const padded = (nheads + 7u) & ~7u;
            for (std::uint32_t h = static_cast<std::uint32_t>(tid); h < padded; h += static_cast<std::uint32_t>(nthreads))
            {
                std::uint32_t v = 0xFFFF0000u;
                if (h < nheads)
                {
                    int const i = s.heads[h];
                    v = (s.level[i] << 16) | s.length[i];
                }
                hinfo[h] = v;
            }
        }
        __syncthreads();
        if (tid < 64)
        {
            unsigned int const levels = s.levels;
            std::uint32_t const nheads = s.nheads;
            unsigned int const l = static_cast<unsigned int>(lane);
            if (l < levels)
            {
                std::uint32_t const blocks = (s.level_cells[l] + 63) / 64;
                std::uint32_t cap[sweep_waves];
                std::uint32_t part[sweep_waves];
#pragma unroll
                for (int w = 0; w < sweep_waves; ++w)
                {
                    cap[w] = 64 * ((blocks + sweep_waves - 1 - w) / sweep_waves);
                    part[w] = 0;
                }
                // the part of every chain of level l and its offset there (first fit, in cell order), kept in
                // base[head] as part << 14 | offset until the parts' starts are known
                for (std::uint32_t h8 = 0; h8 < nheads; h8 += 8)
                {
                    uint4 const info0 = *reinterpret_cast<uint4 const*>(hinfo + h8);
                    uint4 const info1 = *reinterpret_cast<uint4 const*>(hinfo + h8 + 4);
                    std::uint32_t const quad[8] = {info0.x, info0.y, info0.z, info0.w, info1.x, info1.y, info1.z, info1.w};
                    uint4 const head_words = *reinterpret_cast<uint4 const*>(s.heads + h8);
                    std::uint32_t const head4[8] = {head_words.x & 0xFFFFu, head_words.x >> 16, head_words.y & 0xFFFFu,
                                                    head_words.y >> 16,     head_words.z & 0xFFFFu, head_words.z >> 16,
                                                    head_words.w & 0xFFFFu, head_words.w >> 16};
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                    {
                    if ((quad[k] >> 16) != l)
                        continue;
                    std::uint32_t const len = quad[k] & 0xFFFFu;
                    int chosen = -1;
                    int roomiest = 0;
                    // rooms fit 32 bits: capacities and fills are at most the tile's cells
                    int room_best = static_cast<int>(cap[0]) - static_cast<int>(part[0]);
#pragma unroll
                    for (int w = 0; w < sweep_waves; ++w)
                    {
                        int const room = static_cast<int>(cap[w]) - static_cast<int>(part[w]);
                        if (chosen < 0 && room >= static_cast<int>(len))
                            chosen = w;
                        if (room > room_best)
                        {
                            roomiest = w;
                            room_best = room;
                        }
                    }
                    if (chosen < 0)
                        chosen = roomiest;
                    std::uint32_t offset = 0;
#pragma unroll
                    for (int w = 0; w < sweep_waves; ++w)
                    {
                        offset = w == chosen ? part[w] : offset;
                        part[w] += w == chosen ? len : 0u;
                    }
                    s.base[head4[k]] = static_cast<std::uint16_t>((static_cast<std::uint32_t>(chosen) << 14) | offset);
                    }
                }
                // part boundaries: level l starts after the cells of the levels below it
                std::uint32_t at = 0;
                for (unsigned int k = 0; k < l; ++k)
                    at += s.level_cells[k];
#pragma unroll
                for (int w = 0; w < sweep_waves; ++w)
                {
                    s.split[l * (sweep_waves + 1) + w] = static_cast<std::uint16_t>(at);
                    at += part[w];
                }
                s.split[l * (sweep_waves + 1) + sweep_waves] = static_cast<std::uint16_t>(at);
            }
            else if (l < static_cast<unsigned int>(tile_max_levels))
            {
                std::uint32_t total = 0;
                for (unsigned int k = 0; k < levels; ++k)
                    total += s.level_cells[k];
                for (int w = 0; w <= sweep_waves; ++w)
                    s.split[l * (sweep_waves + 1) + w] = static_cast<std::uint16_t>(total);
            }
        }
        __syncthreads();
        // positions: a packed chain starts at its part's start plus its offset there
        {
            std::uint32_t const nheads = s.nheads;
            unsigned int const levels = s.levels;
            for (std::uint32_t h = static_cast<std::uint32_t>(tid); h < nheads; h += static_cast<std::uint32_t>(nthreads))
            {
                int const i = s.heads[h];
                std::uint32_t const l = s.level[i];
                if (l >= levels)
                    continue;  // not packed (a tile with too many levels; `error` is set)
                std::uint32_t const v = s.base[i];
                s.base[i] = static_cast<std::uint16_t>(s.split[l * (sweep_waves + 1) + (v >> 14)] + (v & 0x3FFFu));
            }
        }
        __syncthreads();

        std::uint16_t* const tile_order = order + tile * G::n;
        std::uint8_t* const tile_flags = flags + tile * G::n;
        for (int i = tid; i < G::n; i += nthreads)
        {
            if (!s.local[i])
                continue;
            int const h = s.head_of[i];
            int const len = s.length[h];
            int const d = s.depth[i];
            int const pos = s.base[h] + len - 1 - d;
            if (pos >= G::n)
                continue;
            // the chain start (its most upstream cell) in bit 15 of the order entry; the flags byte holds the
            // light children: the local upstream neighbours other than the chain predecessor (heavy child)
            tile_order[pos] = static_cast<std::uint16_t>(i | ((d == len - 1) ? order_start_bit : 0));
            unsigned light = s.children[i];
            if (s.heavy[i] >= 0)
                light &= ~(1u << s.heavy_dir[i]);
            tile_flags[pos] = static_cast<std::uint8_t>(light);
        }
        for (int e = tid; e < tile_split_entries; e += nthreads)
            split[tile * tile_split_entries + e] = s.split[e];
        if (tid == 0)
            nlevels[tile] = static_cast<std::uint8_t>(s.levels);
    }

    // The rest of a tile's replay program, recorded once per flow network after the chain layout (tile_off: the
    // first compact index of each tile's pending cells, cell order):
    //   lo_bits[tile][w]: bit set for a local cell whose downstream cell is a data cell in another tile (that
    //     cell is pending: such a local cell feeds one of its side slots); lo_count[tile]: their number (the
    //     list of them, with their slots, is tile_lo_list_kernel's);
    //   stage_cell[tile_off[tile] + rank]: the rank-th pending cell of the tile (cell order) as its local index
    //     | (mask of its upstream neighbours that are local cells of the tile, bit k = neighbour k) << 16: the
    //     tile pass adds their outflows (neighbour order) to the cell's side sum (tile_sort_kernel then lists
    //     the same entries in chain-position order).
    template<int TW, int TH>
    __global__ void __launch_bounds__(256)
        tile_program_kernel(Catchment const* cats, std::uint32_t const* tile_cat, std::uint32_t const* pending_bits,
                            std::uint32_t const* tile_off, std::uint32_t* lo_bits, std::uint32_t* lo_count,
                            std::uint32_t* stage_cell)
    {
        using G = Tile<TW, TH>;
        static_assert(G::n % 256 == 0 && G::n <= 4096, "tile programs assume 256-cell groups of at most 4096 cells");
        __shared__ std::uint8_t ldd[G::hn];
        __shared__ std::uint32_t words[G::words];
        __shared__ std::uint32_t prefix[G::words];
        __shared__ alignas(4) std::uint8_t side[G::n];
        __shared__ unsigned int nlo;
        std::int64_t const tile = blockIdx.x;
        int const tid = threadIdx.x;
        Catchment const cat = cats[tile_cat[tile]];
        std::int64_t const local_tile = tile - cat.tile_base;
        std::int64_t const r0 = (local_tile / cat.tiles_x) * TH;
        std::int64_t const c0 = (local_tile % cat.tiles_x) * TW;
        auto const g_ldd = as_global(cat.ldd);
        for (int i = tid; i < G::hn; i += 256)
        {
            std::int64_t const r = r0 + i / G::hw - 1;
            std::int64_t const c = c0 + i % G::hw - 1;
            ldd[i] = (r >= 0 && r < cat.rows && c >= 0 && c < cat.cols) ? g_ldd[r * cat.cols + c] : no_data_ldd;
        }
        for (int w = tid; w < G::words; w += 256)
            words[w] = as_global(pending_bits)[tile * G::words + w];
        for (int w = tid; w < G::n / 4; w += 256)
            reinterpret_cast<std::uint32_t*>(side)[w] = 0u;
        if (tid == 0)
            nlo = 0;
        __syncthreads();
        auto pending = [&](int i) { return ((words[i >> 5] >> (i & 31)) & 1u) != 0u; };
        for (int i = tid; i < G::n; i += 256)
        {
            int const lr = i / TW;
            int const lc = i % TW;
            std::uint8_t const code = ldd[(lr + 1) * G::hw + (lc + 1)];
            bool lo = false;
            int const dir = ldd_direction(code);
            if (code != no_data_ldd && !pending(i) && dir >= 0)
            {
                int const dr = lr + neighbour_dr(dir);
                int const dc = lc + neighbour_dc(dir);
                if (ldd[(dr + 1) * G::hw + (dc + 1)] != no_data_ldd)
                {
                    if (dr < 0 || dr >= TH || dc < 0 || dc >= TW)
                        lo = true;
                    else if (pending(dr * TW + dc))
                    {
                        int const t = dr * TW + dc;
                        atomicOr(reinterpret_cast<unsigned int*>(side) + (t >> 2),
                                 (1u << opposite(dir)) << (8 * (t & 3)));
                    }
                }
            }
            // a wavefront's 64 consecutive cells are two bitmap words
            std::uint64_t const bits = __ballot(lo);
            if ((tid & 63) == 0)
            {
                as_global(lo_bits)[tile * G::words + (i >> 5)] = static_cast<std::uint32_t>(bits);
                as_global(lo_bits)[tile * G::words + (i >> 5) + 1] = static_cast<std::uint32_t>(bits >> 32);
                atomicAdd(&nlo, static_cast<unsigned int>(__popcll(bits)));
            }
        }
        if (tid == 0)
        {
            std::uint32_t sum = 0;
            for (int w = 0; w < G::words; ++w)
            {
                prefix[w] = sum;
                sum += __popc(words[w]);
            }
        }
        __syncthreads();
        if (tid == 0)
            as_global(lo_count)[tile] = nlo;
        std::uint32_t const begin = as_global(tile_off)[tile];
        for (int i = tid; i < G::n; i += 256)
        {
            if (!pending(i))
                continue;
            std::uint32_t const word = words[i >> 5];
            std::uint32_t const b = static_cast<std::uint32_t>(i & 31);
            std::uint32_t const rank = prefix[i >> 5] + __popc(word & ((1u << b) - 1u));
            as_global(stage_cell)[begin + rank] = static_cast<std::uint32_t>(i) | (static_cast<std::uint32_t>(side[i]) << 16);
        }
    }

    // The pending cells of each tile in chain-position order (a chain crossing the tile is a run of consecutive
    // positions there), so that the tile pass stages and the scatter reads whole runs of positions: from
    // tile_off[tile], sorted_pos[k] = the k-th smallest chain position of the tile's pending cells,
    // sorted_cell[k] = that cell's staging entry (stage_cell), and scatter_entry[rank] = the local cell of that
    // rank (cell order) | its place k in the sorted order << 16. A bitonic sort in LDS, one workgroup per tile.
    template<int TW, int TH>
    __global__ void __launch_bounds__(256)
        tile_sort_kernel(std::uint32_t const* tile_off, std::uint32_t const* pos_of, std::uint32_t const* stage_cell,
                         std::uint32_t* sorted_pos, std::uint32_t* sorted_cell, std::uint32_t* scatter_entry)
    {
        using G = Tile<TW, TH>;
        static_assert(G::n <= 65536, "ranks fit 16 bits");
        __shared__ std::uint32_t key[G::n];
        __shared__ std::uint16_t val[G::n];
        std::int64_t const tile = blockIdx.x;
        int const tid = threadIdx.x;
        std::uint32_t const begin = as_global(tile_off)[tile];
        int const npending = static_cast<int>(as_global(tile_off)[tile + 1] - begin);
        if (npending == 0)
            return;
        int width = 1;  // the sort's width: a power of two at least npending (at most G::n)
        while (width < npending)
            width <<= 1;
        for (int i = tid; i < width; i += 256)
        {
            key[i] = i < npending ? as_global(pos_of)[begin + i] : 0xFFFFFFFFu;
            val[i] = static_cast<std::uint16_t>(i);
        }
        __syncthreads();
        for (int k = 2; k <= width; k <<= 1)
        {
            for (int j = k >> 1; j > 0; j >>= 1)
            {
                for (int i = tid; i < width; i += 256)
                {
                    int const partner = i ^ j;
                    if (partner > i)
                    {
                        bool const ascending = (i & k) == 0;
                        std::uint32_t const a = key[i];
                        std::uint32_t const b = key[partner];
                        if ((a > b) == ascending)
                        {
                            key[i] = b;
                            key[partner] = a;
                            std::uint16_t const t = val[i];
                            val[i] = val[partner];
                            val[partner] = t;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int i = tid; i < npending; i += 256)
        {
            std::uint32_t const rank = val[i];
            std::uint32_t const entry = as_global(stage_cell)[begin + rank];
            as_global(sorted_pos)[begin + i] = key[i];
            as_global(sorted_cell)[begin + i] = entry;
            as_global(scatter_entry)[begin + rank] = (entry & 0xFFFFu) | (static_cast<std::uint32_t>(i) << 16);
        }
    }

    // The local cells of each tile that feed side slots (lo_bits), in cell order from lo_off[tile]: their side
    // slot (lo_slot, by cell index in the batch, assigned with the chain layout's side slots) << 16 | local cell.
    // A cell without a slot below nslots (none when the record is consistent) is counted in `unassigned`.
    template<int TW, int TH>
    __global__ void __launch_bounds__(256)
        tile_lo_list_kernel(Catchment const* cats, std::uint32_t const* tile_cat, std::uint32_t const* lo_bits,
                            std::uint32_t const* lo_off, std::uint32_t const* lo_slot, std::uint32_t nslots,
                            std::uint64_t* lo_entry, std::uint32_t* unassigned)
    {
        using G = Tile<TW, TH>;
        __shared__ std::uint32_t words[G::words];
        __shared__ std::uint32_t prefix[G::words];
        std::int64_t const tile = blockIdx.x;
        int const tid = threadIdx.x;
        std::uint32_t const begin = as_global(lo_off)[tile];
        if (as_global(lo_off)[tile + 1] == begin)
            return;
        Catchment const cat = cats[tile_cat[tile]];
        std::int64_t const local_tile = tile - cat.tile_base;
        std::int64_t const r0 = (local_tile / cat.tiles_x) * TH;
        std::int64_t const c0 = (local_tile % cat.tiles_x) * TW;
        for (int w = tid; w < G::words; w += 256)
            words[w] = as_global(lo_bits)[tile * G::words + w];
        __syncthreads();
        if (tid == 0)
        {
            std::uint32_t sum = 0;
            for (int w = 0; w < G::words; ++w)
            {
                prefix[w] = sum;
                sum += __popc(words[w]);
            }
        }
        __syncthreads();
        for (int i = tid; i < G::n; i += 256)
        {
            std::uint32_t const word = words[i >> 5];
            std::uint32_t const b = static_cast<std::uint32_t>(i & 31);
            if (((word >> b) & 1u) == 0u)
                continue;
            std::uint32_t const rank = prefix[i >> 5] + __popc(word & ((1u << b) - 1u));
            std::int64_t const cell = cat.cell_base + (r0 + i / TW) * cat.cols + (c0 + i % TW);
            std::uint32_t slot = as_global(lo_slot)[cell];
            if (slot >= nslots)
            {
                // recorded for every source; the replay's stores stay in bounds whatever the record
                atomicAdd(unassigned, 1u);
                slot = 0;
            }
            as_global(lo_entry)[begin + rank] = (static_cast<std::uint64_t>(slot) << 16) | static_cast<std::uint64_t>(i);
        }
    }

    // Block statistics of the in-tile layouts (one thread per tile): counts[0] += the 64-position blocks the
    // tile pass runs (over levels and wavefronts), counts[1] += the blocks on the tile's critical path (per
    // level the most of any wavefront).
    __global__ inline void tile_block_stats_kernel(std::uint16_t const* split, std::uint8_t const* nlevels,
                                                   std::int64_t ntiles, unsigned long long* counts)
    {
        std::int64_t const tile = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        unsigned long long blocks = 0;
        unsigned long long critical = 0;
        if (tile < ntiles)
        {
            int const levels = min(static_cast<int>(nlevels[tile]), tile_max_levels);
            std::uint16_t const* t = split + tile * tile_split_entries;
            for (int l = 0; l < levels; ++l)
            {
                unsigned most = 0;
                for (int w = 0; w < sweep_waves; ++w)
                {
                    unsigned const part = static_cast<unsigned>(t[l * (sweep_waves + 1) + w + 1]) -
                                          static_cast<unsigned>(t[l * (sweep_waves + 1) + w]);
                    unsigned const b = (part + 63u) / 64u;
                    blocks += b;
                    most = b > most ? b : most;
                }
                critical += most;
            }
        }
        for (int offset = 32; offset > 0; offset >>= 1)
        {
            blocks += __shfl_down(blocks, offset, 64);
            critical += __shfl_down(critical, offset, 64);
        }
        if ((threadIdx.x & 63) == 0)
        {
            atomicAdd(&counts[0], blocks);
            atomicAdd(&counts[1], critical);
        }
    }

}  // namespace lue_amd
