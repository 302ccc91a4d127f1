// Chain phase of the routing sweep: the cells that round 0 of the tiled sweep leaves pending.
//
// Round 0 (sweep.hpp) solves every cell whose upstream area lies inside its own tile. What is left,
// the pending cells, are the cells downstream of a tile border crossing: a forest (each pending cell
// has at most one pending downstream cell) whose long paths are the main streams. They are solved by
// a heavy-path decomposition of that forest, recorded once per flow network:
//   1. compact the pending cells (tile after tile), link each to its pending downstream cell;
//   2. Euler tour of the forest (children in neighbour order) and list ranking (a sparse ruling set: rulers
//      walk their sublists, the rulers are ranked by pointer jumping) give every pending cell the number of
//      pending cells upstream of it (its subtree);
//   3. the heavy child of a cell is its pending upstream neighbour with the largest subtree; a second
//      Euler tour with heavy children first numbers the cells in preorder, in which every heavy path
//      (a chain) is one contiguous run;
//   4. the level of a chain is one more than the highest level of the chains that join it from the
//      side (light children); 0 when none do. Any path crosses at most log2(n) light edges, so there
//      are few levels. Chains are laid out level after level, each in flow order (upstream first);
//   5. per chain cell, the slots of the values it receives from the side, in neighbour order: light
//      children (the last position of a chain of a lower level) and upstream cells in other tiles that
//      round 0 solves (raster cells). Every source feeds exactly one slot, which it writes itself once its
//      value is final (push): round 0 for raster cells, the chain run for chain ends.
// Running an operation: round 0 solves the non-pending cells and stages the inputs of the pending ones
// at their chain positions, together with the FP64 sum of their upstream cells in the same tile that
// round 0 solved. Then per level a gather (each position with side slots adds them, contiguous, to its
// staged sum) and a run: a wavefront takes a work item (one long chain, or several
// short consecutive chains of at most 64 cells), 64 positions at a time, and runs the cell function
// along the chains (scans where the function allows it, serially with lane broadcasts where it does
// not). A final pass writes the chain results to the output rasters. Levels only exchange data across
// launches, so no kernel synchronises workgroups.
#pragma once

#include "kinematic_wave.hpp"
#include "routing_common.hpp"
#include "sweep.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace lue_amd {

    // No side slot (a position that is not the last of a light chain).
    constexpr std::uint32_t no_slot = 0xFFFFFFFFu;

    // Default lane limit of a network: work items of a chain level with at most this many positions are walked
    // one per lane (run_items_lane_walk), longer ones one per wavefront, 64 positions at a time. 0: the lane
    // walk is off by default; at 16384^2 it measured slower than the wavefront blocks for every operation but
    // the trigger, so the autotuner chooses it per network.
    constexpr std::uint32_t lane_max_positions = 0;

    template<int TW, int TH>
    __device__ inline bool cell_pending(std::uint32_t const* pending_bits, Catchment const& cat, std::int64_t r,
                                        std::int64_t c)
    {
        using G = Tile<TW, TH>;
        std::int64_t const tile = cat.tile_base + (r / TH) * cat.tiles_x + (c / TW);
        int const i = static_cast<int>((r % TH) * TW + (c % TW));
        return (as_global(pending_bits)[tile * G::words + (i >> 5)] >> (i & 31)) & 1u;
    }

    // Catchment of a global cell id (cell_base is increasing).
    __device__ inline int find_catchment(Catchment const* cats, int ncats, std::uint64_t cell)
    {
        int lo = 0;
        int hi = ncats - 1;
        while (lo < hi)
        {
            int const mid = (lo + hi + 1) / 2;
            if (static_cast<std::uint64_t>(cats[mid].cell_base) <= cell)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    }

    // Pending upstream neighbours (local cell indices) of local cell `local`, in neighbour order.
    template<int TW, int TH>
    __device__ inline int pending_upstream(Catchment const& cat, std::uint32_t const* pending_bits,
                                           std::int64_t local, std::int64_t (&up)[8], int (&dir)[8])
    {
        std::int64_t const r = local / cat.cols;
        std::int64_t const c = local % cat.cols;
        auto const ldd = as_global(cat.ldd);
        int n = 0;
        for (int k = 0; k < 8; ++k)
        {
            std::int64_t const r2 = r + neighbour_dr(k);
            std::int64_t const c2 = c + neighbour_dc(k);
            if (r2 < 0 || r2 >= cat.rows || c2 < 0 || c2 >= cat.cols)
                continue;
            std::uint8_t const code = ldd[r2 * cat.cols + c2];
            if (code == no_data_ldd || ldd_direction(code) != opposite(k))
                continue;
            if (!cell_pending<TW, TH>(pending_bits, cat, r2, c2))
                continue;
            up[n] = r2 * cat.cols + c2;
            dir[n] = k;
            ++n;
        }
        return n;
    }

    // ---------------------------------------------------------------------------------------------------
    // Recording

    // Pending cells per tile.
    template<int TW, int TH>
    __global__ void chains_count_kernel(std::uint32_t const* pending_bits, std::int64_t ntiles, std::uint32_t* counts)
    {
        using G = Tile<TW, TH>;
        std::int64_t const tile = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (tile >= ntiles)
            return;
        std::uint32_t n = 0;
        for (int w = 0; w < G::words; ++w)
            n += __popc(pending_bits[tile * G::words + w]);
        counts[tile] = n;
    }

    // List the pending cells of each tile at its offset (cell order inside the tile); record global cell
    // id -> list index.
    template<int TW, int TH>
    __global__ void chains_compact_kernel(std::uint32_t const* pending_bits, Catchment const* cats,
                                          std::uint32_t const* tile_cat, std::uint32_t const* offsets,
                                          std::uint32_t* cells, std::uint32_t* index_of)
    {
        using G = Tile<TW, TH>;
        std::int64_t const tile = blockIdx.x;
        Catchment const cat = cats[tile_cat[tile]];
        std::int64_t const local = tile - cat.tile_base;
        int const ty = static_cast<int>(local / cat.tiles_x);
        int const tx = static_cast<int>(local % cat.tiles_x);
        std::uint32_t const* bits = pending_bits + tile * G::words;
        for (int w = threadIdx.x; w < G::words; w += blockDim.x)
        {
            std::uint32_t word = bits[w];
            if (!word)
                continue;
            std::uint32_t pos = offsets[tile];
            for (int v = 0; v < w; ++v)
                pos += __popc(bits[v]);
            while (word)
            {
                int const b = __ffs(word) - 1;
                word &= word - 1;
                int const i = w * 32 + b;
                std::int64_t const r = static_cast<std::int64_t>(ty) * TH + i / TW;
                std::int64_t const c = static_cast<std::int64_t>(tx) * TW + i % TW;
                std::uint32_t const cell = static_cast<std::uint32_t>(cat.cell_base + r * cat.cols + c);
                cells[pos] = cell;
                index_of[cell] = pos;
                ++pos;
            }
        }
    }

    // Parent (pending downstream cell) of every pending cell, -1 at the end of a pending path.
    template<int TW, int TH>
    __global__ void chains_parent_kernel(Catchment const* cats, int ncats, std::uint32_t const* pending_bits,
                                         std::uint32_t const* cells, std::uint32_t const* index_of, std::uint32_t n,
                                         std::int32_t* parent)
    {
        std::uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q >= n)
            return;
        std::uint32_t const cell = cells[q];
        Catchment const cat = cats[find_catchment(cats, ncats, cell)];
        std::int64_t const local = cell - cat.cell_base;
        std::int64_t const r = local / cat.cols;
        std::int64_t const c = local % cat.cols;
        int const dir = ldd_direction(cat.ldd[local]);
        std::int32_t p = -1;
        if (dir >= 0)
        {
            std::int64_t const r2 = r + neighbour_dr(dir);
            std::int64_t const c2 = c + neighbour_dc(dir);
            if (r2 >= 0 && r2 < cat.rows && c2 >= 0 && c2 < cat.cols && cat.ldd[r2 * cat.cols + c2] != no_data_ldd &&
                cell_pending<TW, TH>(pending_bits, cat, r2, c2))
            {
                p = static_cast<std::int32_t>(index_of[cat.cell_base + r2 * cat.cols + c2]);
            }
        }
        parent[q] = p;
    }

    // Children of pending cell q in tour order: the heavy child first (when given), then the others in
    // neighbour order. Returns the count.
    template<int TW, int TH>
    __device__ inline int tour_children(Catchment const& cat, std::uint32_t const* pending_bits,
                                        std::uint32_t const* index_of, std::int64_t local, std::int32_t heavy,
                                        std::int32_t (&child)[8])
    {
        std::int64_t up[8];
        int dir[8];
        int const m = pending_upstream<TW, TH>(cat, pending_bits, local, up, dir);
        int n = 0;
        if (heavy >= 0)
            child[n++] = heavy;
        for (int j = 0; j < m; ++j)
        {
            std::int32_t const q = static_cast<std::int32_t>(as_global(index_of)[cat.cell_base + up[j]]);
            if (q != heavy)
                child[n++] = q;
        }
        return n;
    }

    // Euler tour successor of every tour element: enter(q) = 2q, exit(q) = 2q + 1.
    // enter(q) -> enter(first child) or exit(q); exit(q) -> enter(next sibling) or exit(parent) or end (-1).
    template<int TW, int TH>
    __global__ void chains_tour_kernel(Catchment const* cats, int ncats, std::uint32_t const* pending_bits,
                                       std::uint32_t const* cells, std::uint32_t const* index_of,
                                       std::int32_t const* parent, std::int32_t const* heavy, std::uint32_t n,
                                       std::int32_t* succ)
    {
        std::uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q >= n)
            return;
        std::uint32_t const cell = cells[q];
        Catchment const cat = cats[find_catchment(cats, ncats, cell)];
        std::int32_t child[8];
        int const nc = tour_children<TW, TH>(cat, pending_bits, index_of, cell - cat.cell_base,
                                             heavy ? heavy[q] : -1, child);
        succ[2 * q] = nc > 0 ? 2 * child[0] : static_cast<std::int32_t>(2 * q + 1);
        std::int32_t const d = parent[q];
        std::int32_t s = -1;
        if (d >= 0)
        {
            std::int32_t sib[8];
            std::uint32_t const dcell = cells[d];
            int const ns = tour_children<TW, TH>(cat, pending_bits, index_of, dcell - cat.cell_base,
                                                 heavy ? heavy[d] : -1, sib);
            s = 2 * d + 1;
            for (int j = 0; j + 1 < ns; ++j)
            {
                if (sib[j] == static_cast<std::int32_t>(q))
                {
                    s = 2 * sib[j + 1];
                    break;
                }
            }
        }
        succ[2 * q + 1] = s;
    }

    // One pointer-jumping step (double buffered) over linked elements with values, next (-1 at the end) and
    // the list's last element (known at the end, carried back): value becomes the sum from the element to the
    // list's end, last the list's last element. Counts the elements still linked.
    __global__ inline void chains_rank_step_kernel(std::int32_t const* next, std::int32_t const* value,
                                                   std::int32_t const* last, std::uint32_t n2, std::int32_t* next2,
                                                   std::int32_t* value2, std::int32_t* last2, std::uint32_t* linked)
    {
        std::uint32_t const e = blockIdx.x * blockDim.x + threadIdx.x;
        bool still = false;
        if (e < n2)
        {
            std::int32_t const s = next[e];
            if (s >= 0)
            {
                std::int32_t const s2 = next[s];
                next2[e] = s2;
                value2[e] = value[e] + value[s];
                last2[e] = last[s];
                still = s2 >= 0;
            }
            else
            {
                next2[e] = -1;
                value2[e] = value[e];
                last2[e] = last[e];
            }
        }
        std::uint64_t const ballot = __ballot(still);
        if ((threadIdx.x & 63) == 0 && ballot)
            atomicAdd(linked, static_cast<std::uint32_t>(__popcll(ballot)));
    }

    // List ranking by a sparse ruling set (work O(n) instead of pointer jumping's O(n log n)). Rulers are the
    // list heads (elements no other element links to) and a pseudo-random 1/64 of the others; each ruler
    // walks its sublist up to the next ruler, recording for every element its ruler and the inclusive prefix
    // of the values from the ruler; the rulers, linked to their next rulers with their sublists' sums, are
    // ranked by pointer jumping (few: about n / 64 plus the heads); every element's rank follows from its
    // ruler's. value(e) = 1 for enter elements (even e), 0 for exit elements.
    constexpr std::uint32_t ruler_sample = 64;
    constexpr std::uint32_t ruler_walk_limit = 1u << 20;  // a longer walk can only circle on a cycle

    __device__ inline bool sampled_ruler(std::uint32_t e)
    {
        std::uint64_t z = static_cast<std::uint64_t>(e) + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return ((z ^ (z >> 31)) & (ruler_sample - 1)) == 0;
    }

    // has_pred[succ[e]] = 1 (has_pred zeroed before).
    __global__ inline void rank_mark_successors_kernel(std::int32_t const* succ, std::uint32_t n2,
                                                       std::uint32_t* has_pred)
    {
        std::uint32_t const e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= n2)
            return;
        std::int32_t const s = succ[e];
        if (s >= 0 && static_cast<std::uint32_t>(s) < n2)
            has_pred[s] = 1u;
    }

    // is_ruler[e] (0/1) for the scan that numbers the rulers; owner[e] = -1 (no ruler reached it yet).
    __global__ inline void rank_rulers_kernel(std::uint32_t const* has_pred, std::uint32_t n2, std::uint32_t* is_ruler,
                                              std::int32_t* owner)
    {
        std::uint32_t const e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= n2)
            return;
        is_ruler[e] = (has_pred[e] == 0u || sampled_ruler(e)) ? 1u : 0u;
        owner[e] = -1;
    }

    // An element's ruler and its inclusive prefix from the ruler, in one word (the walk writes both with one
    // store): owner in the low half (-1: no ruler reached it), prefix in the high half.
    __device__ inline std::uint64_t owner_prefix(std::int32_t owner, std::int32_t prefix)
    {
        return (static_cast<std::uint64_t>(static_cast<std::uint32_t>(prefix)) << 32) | static_cast<std::uint32_t>(owner);
    }

    // is_ruler[e] (0/1) for the scan that numbers the rulers; ownpre[e] = no ruler yet.
    __global__ inline void rank_rulers_packed_kernel(std::uint32_t const* has_pred, std::uint32_t n2,
                                                     std::uint32_t* is_ruler, std::uint64_t* ownpre)
    {
        std::uint32_t const e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= n2)
            return;
        is_ruler[e] = (has_pred[e] == 0u || sampled_ruler(e)) ? 1u : 0u;
        ownpre[e] = owner_prefix(-1, 0);
    }

    // Walk of ruler r (element ruler_elem[r]) to the next ruler or the list's end. An element a walk steps to
    // has a predecessor, so it is a ruler exactly when it is sampled (sampled_ruler, computed here instead of
    // read: the walk's random accesses are one load and one store per element). `bad` counts walks that
    // reached the limit (cycles).
    __global__ inline void rank_walk_kernel(std::int32_t const* succ, std::uint32_t const* ruler_rank,
                                            std::uint32_t const* ruler_elem, std::uint32_t nrulers,
                                            std::uint64_t* ownpre, std::int32_t* sub_sum, std::int32_t* next_ruler,
                                            std::int32_t* last_elem, std::uint32_t* bad)
    {
        std::uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
        if (r >= nrulers)
            return;
        std::uint32_t x = ruler_elem[r];
        std::int32_t acc = 0;
        std::int32_t next = -1;
        std::int32_t last = -1;
        for (std::uint32_t step = 0;; ++step)
        {
            if (step == ruler_walk_limit)
            {
                atomicAdd(bad, 1u);
                break;
            }
            acc += (x & 1u) == 0u ? 1 : 0;
            ownpre[x] = owner_prefix(static_cast<std::int32_t>(r), acc);
            std::int32_t const s = succ[x];
            if (s < 0)
            {
                last = static_cast<std::int32_t>(x);
                break;
            }
            if (sampled_ruler(static_cast<std::uint32_t>(s)))
            {
                next = static_cast<std::int32_t>(ruler_rank[s]);
                break;
            }
            x = static_cast<std::uint32_t>(s);
        }
        sub_sum[r] = acc;
        next_ruler[r] = next;
        last_elem[r] = last;
    }

    // ruler_elem[ruler_rank[e]] = e for the rulers.
    __global__ inline void rank_ruler_elements_kernel(std::uint32_t const* is_ruler, std::uint32_t const* ruler_rank,
                                                      std::uint32_t n2, std::uint32_t* ruler_elem)
    {
        std::uint32_t const e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e < n2 && is_ruler[e])
            ruler_elem[ruler_rank[e]] = e;
    }

    // Element ranks from their rulers' (rv: sum from the ruler to the list's end, rlast: the list's last
    // element); elements no ruler reached lie on a cycle (`bad`).
    __global__ inline void rank_elements_kernel(std::uint64_t const* ownpre, std::int32_t const* rv,
                                                std::int32_t const* rlast, std::uint32_t n2, std::int32_t* value,
                                                std::int32_t* last, std::uint32_t* bad)
    {
        std::uint32_t const e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= n2)
            return;
        std::uint64_t const op = ownpre[e];
        std::int32_t const r = static_cast<std::int32_t>(static_cast<std::uint32_t>(op));
        std::int32_t const pre = static_cast<std::int32_t>(static_cast<std::uint32_t>(op >> 32));
        if (r < 0)
        {
            atomicAdd(bad, 1u);
            value[e] = 0;
            last[e] = static_cast<std::int32_t>(e);
            return;
        }
        value[e] = rv[r] - pre + ((e & 1u) == 0u ? 1 : 0);
        last[e] = rlast[r];
    }

    // Second level of the ruling set, over the rulers' own list (rnext: the next ruler or -1, rsum: the sum of
    // the ruler's sublist, rlast: the list's last element where the sublist ends the list): walk of
    // super-ruler q (ruler super_elem[q]) to the next super-ruler or the list's end, recording for every ruler
    // its super-ruler and the inclusive prefix of the sums from it.
    __global__ inline void rank2_walk_kernel(std::int32_t const* rnext, std::int32_t const* rsum,
                                             std::int32_t const* rlast, std::uint32_t const* is_super,
                                             std::uint32_t const* super_rank, std::uint32_t const* super_elem,
                                             std::uint32_t nsupers, std::int32_t* owner, std::int32_t* prefix,
                                             std::int32_t* sub_sum, std::int32_t* next_super, std::int32_t* last_elem,
                                             std::uint32_t* bad)
    {
        std::uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q >= nsupers)
            return;
        std::uint32_t x = super_elem[q];
        std::int32_t acc = 0;
        std::int32_t next = -1;
        std::int32_t last = -1;
        for (std::uint32_t step = 0;; ++step)
        {
            if (step == ruler_walk_limit)
            {
                atomicAdd(bad, 1u);
                break;
            }
            acc += rsum[x];
            owner[x] = static_cast<std::int32_t>(q);
            prefix[x] = acc;
            std::int32_t const s = rnext[x];
            if (s < 0)
            {
                last = rlast[x];
                break;
            }
            if (is_super[s])
            {
                next = static_cast<std::int32_t>(super_rank[s]);
                break;
            }
            x = static_cast<std::uint32_t>(s);
        }
        sub_sum[q] = acc;
        next_super[q] = next;
        last_elem[q] = last;
    }

    // The rulers' sums to their list's end and the list's last element from their super-rulers' (qv, qlast);
    // rulers no super-ruler reached lie on a cycle (`bad`).
    __global__ inline void rank2_rulers_kernel(std::int32_t const* owner, std::int32_t const* prefix,
                                               std::int32_t const* rsum, std::int32_t const* qv,
                                               std::int32_t const* qlast, std::uint32_t nrulers, std::int32_t* rv,
                                               std::int32_t* rlast_out, std::uint32_t* bad)
    {
        std::uint32_t const r = blockIdx.x * blockDim.x + threadIdx.x;
        if (r >= nrulers)
            return;
        std::int32_t const q = owner[r];
        if (q < 0)
        {
            atomicAdd(bad, 1u);
            rv[r] = 0;
            rlast_out[r] = -1;
            return;
        }
        rv[r] = qv[q] - prefix[r] + rsum[r];
        rlast_out[r] = qlast[q];
    }

    // Counts entries of `perm` that are out of range or repeated (perm must be a permutation of [0, n)).
    __global__ inline void chains_check_permutation_kernel(std::uint32_t const* perm, std::uint32_t n,
                                                           std::uint32_t* seen, std::uint32_t* bad)
    {
        std::uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q >= n)
            return;
        std::uint32_t const p = perm[q];
        if (p >= n || atomicAdd(&seen[p], 1u) != 0u)
            atomicAdd(bad, 1u);
    }

    // Heavy child: the child with the largest subtree (subtree = value(enter) - value(exit)); ties go to
    // the first in neighbour order.
    template<int TW, int TH>
    __global__ void chains_heavy_kernel(Catchment const* cats, int ncats, std::uint32_t const* pending_bits,
                                        std::uint32_t const* cells, std::uint32_t const* index_of,
                                        std::int32_t const* value, std::uint32_t n, std::int32_t* heavy)
    {
        std::uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q >= n)
            return;
        std::uint32_t const cell = cells[q];
        Catchment const cat = cats[find_catchment(cats, ncats, cell)];
        std::int32_t child[8];
        int const nc = tour_children<TW, TH>(cat, pending_bits, index_of, cell - cat.cell_base, -1, child);
        std::int32_t best = -1;
        std::int32_t best_size = -1;
        for (int j = 0; j < nc; ++j)
        {
            std::int32_t const c = child[j];
            std::int32_t const size = value[2 * c] - value[2 * c + 1];
            if (size > best_size)
            {
                best = c;
                best_size = size;
            }
        }
        heavy[q] = best;
    }

    __global__ inline void chains_root_flag_kernel(std::int32_t const* parent, std::uint32_t n, std::uint32_t* flag)
    {
        std::uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q < n)
            flag[q] = parent[q] < 0 ? 1u : 0u;
    }

    __global__ inline void chains_tree_size_kernel(std::int32_t const* parent, std::int32_t const* value,
                                                   std::uint32_t const* root_index, std::uint32_t n,
                                                   std::uint32_t* tree_size)
    {
        std::uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q < n && parent[q] < 0)
            tree_size[root_index[q]] = static_cast<std::uint32_t>(value[2 * q]);
    }

    // Preorder number (heavy children first) of every pending cell, and head flags by preorder.
    __global__ inline void chains_preorder_kernel(std::int32_t const* parent, std::int32_t const* heavy,
                                                  std::int32_t const* value, std::int32_t const* last,
                                                  std::uint32_t const* root_index, std::uint32_t const* tree_base,
                                                  std::uint32_t n, std::uint32_t* pre, std::uint32_t* head_flag,
                                                  std::uint32_t* at_pre)
    {
        std::uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q >= n)
            return;
        std::int32_t const end = last[2 * q];  // exit(root)
        std::uint32_t const root = static_cast<std::uint32_t>(end) >> 1;
        std::uint32_t const size = static_cast<std::uint32_t>(value[2 * root]);
        std::uint32_t const p = tree_base[root_index[root]] + size - static_cast<std::uint32_t>(value[2 * q]);
        pre[q] = p;
        std::int32_t const d = parent[q];
        if (p < n)
        {
            head_flag[p] = (d < 0 || heavy[d] != static_cast<std::int32_t>(q)) ? 1u : 0u;
            at_pre[p] = q;
        }
    }

    // Chains are numbered by preorder of their heads: chain_pre[c] = preorder of head of chain c.
    __global__ inline void chains_head_kernel(std::uint32_t const* head_flag, std::uint32_t const* chain_incl,
                                              std::uint32_t n, std::uint32_t* chain_pre)
    {
        std::uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
        if (p < n && head_flag[p])
            chain_pre[chain_incl[p] - 1] = p;
    }

    __global__ inline void chains_length_kernel(std::uint32_t const* chain_pre, std::uint32_t nchains, std::uint32_t n,
                                                std::uint32_t* length)
    {
        std::uint32_t const c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c < nchains)
            length[c] = (c + 1 < nchains ? chain_pre[c + 1] : n) - chain_pre[c];
    }

    // One relaxation of the levels over the chains: level(chain of parent(head)) >= level(chain) + 1 for every
    // chain whose head has a parent.
    // The chain that chain c joins from the side (the chain of its head's downstream cell), ~0 for none: the
    // chain levels' relaxation below reads it in every pass.
    __global__ inline void chains_parent_chain_kernel(std::int32_t const* parent, std::uint32_t const* pre,
                                                      std::uint32_t const* at_pre, std::uint32_t const* chain_pre,
                                                      std::uint32_t const* chain_incl, std::uint32_t nchains,
                                                      std::uint32_t* parent_chain)
    {
        std::uint32_t const c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= nchains)
            return;
        std::uint32_t const q = at_pre[chain_pre[c]];
        std::int32_t const d = parent[q];
        parent_chain[c] = d < 0 ? ~0u : chain_incl[pre[d]] - 1;
    }

    // One pass of the relaxation level(parent chain) >= level(chain) + 1; `changed` counts raised levels.
    __global__ inline void chains_level_by_chain_kernel(std::uint32_t const* parent_chain, std::uint32_t nchains,
                                                        std::uint32_t* level, std::uint32_t* changed)
    {
        std::uint32_t const c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= nchains)
            return;
        std::uint32_t const cd = parent_chain[c];
        if (cd == ~0u)
            return;
        std::uint32_t const want = level[c] + 1;
        if (level[cd] < want)
        {
            if (atomicMax(&level[cd], want) < want)
                atomicAdd(changed, 1u);
        }
    }

    // Chain layout (device). Chains sorted by level (stable, so in preorder of their heads inside a level) lie
    // one after the other: their positions are the exclusive scan of their lengths in that order. `order` is
    // the sorted chain indices, `level_sorted` their levels.
    constexpr int chains_max_levels = 64;

    // Lengths in sorted order, the first sorted chain of every level (level_first, set to ~0 before) and the
    // longest chain (a wavefront maximum, one atomic per wavefront).
    __global__ inline void chains_sorted_lengths_kernel(std::uint32_t const* order, std::uint32_t const* length,
                                                        std::uint32_t const* level_sorted, std::uint32_t nchains,
                                                        std::uint32_t* len_sorted, std::uint32_t* level_first,
                                                        std::uint32_t* longest)
    {
        std::uint32_t const k = blockIdx.x * blockDim.x + threadIdx.x;
        std::uint32_t len = 0;
        if (k < nchains)
        {
            len = length[order[k]];
            len_sorted[k] = len;
            std::uint32_t const l = level_sorted[k];
            if ((k == 0 || level_sorted[k - 1] != l) && l < static_cast<std::uint32_t>(chains_max_levels))
                level_first[l] = k;
        }
        for (int offset = 32; offset > 0; offset >>= 1)
        {
            std::uint32_t const other = __shfl_xor(len, offset, 64);
            len = other > len ? other : len;
        }
        if ((threadIdx.x & 63) == 0 && len > 0)
            atomicMax(longest, len);
    }

    // Per level present: its first position and its first work item (a level's first chain starts an item).
    __global__ inline void chains_level_stats_kernel(std::uint32_t const* level_first, std::uint32_t const* base_sorted,
                                                     std::uint32_t const* item_incl, std::uint32_t* stats)
    {
        int const l = static_cast<int>(threadIdx.x);
        if (l >= chains_max_levels)
            return;
        std::uint32_t const k = level_first[l];
        stats[l] = k == ~0u ? ~0u : base_sorted[k];
        stats[chains_max_levels + l] = k == ~0u ? ~0u : item_incl[k] - 1;
    }

    // Base position of every chain, and the work-item starts: a chain longer than 64 cells is an item of its
    // own; shorter ones go together while they are consecutive, of one level and start in one 64-position
    // window (an item then spans at most two 64-position blocks).
    __global__ inline void chains_items_kernel(std::uint32_t const* order, std::uint32_t const* level_sorted,
                                               std::uint32_t const* len_sorted, std::uint32_t const* base_sorted,
                                               std::uint32_t nchains, std::uint32_t* base, std::uint32_t* item_flag)
    {
        std::uint32_t const k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= nchains)
            return;
        std::uint32_t const b = base_sorted[k];
        base[order[k]] = b;
        bool start = k == 0;
        if (!start)
        {
            bool const long_here = len_sorted[k] > 64;
            bool const long_before = len_sorted[k - 1] > 64;
            start = long_here || long_before || level_sorted[k] != level_sorted[k - 1] ||
                    (b >> 6) != (base_sorted[k - 1] >> 6);
        }
        item_flag[k] = start ? 1u : 0u;
    }

    // Items (unsorted): start, end, level; `key` orders them by level, then the longest first.
    __global__ inline void chains_item_fill_kernel(std::uint32_t const* level_sorted, std::uint32_t const* len_sorted,
                                                   std::uint32_t const* base_sorted, std::uint32_t const* item_flag,
                                                   std::uint32_t const* item_incl, std::uint32_t nchains,
                                                   std::uint32_t* item_begin, std::uint32_t* item_end,
                                                   std::uint32_t* item_level)
    {
        std::uint32_t const k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= nchains)
            return;
        std::uint32_t const id = item_incl[k] - 1;
        if (item_flag[k])
        {
            item_begin[id] = base_sorted[k];
            item_level[id] = level_sorted[k];
        }
        if (k + 1 == nchains || item_flag[k + 1])
            item_end[id] = base_sorted[k] + len_sorted[k];
    }

    __global__ inline void chains_item_key_kernel(std::uint32_t const* item_begin, std::uint32_t const* item_end,
                                                  std::uint32_t const* item_level, std::uint32_t nitems,
                                                  std::uint64_t* key, std::uint32_t* index)
    {
        std::uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= nitems)
            return;
        std::uint32_t const count = item_end[i] - item_begin[i];
        key[i] = (static_cast<std::uint64_t>(item_level[i]) << 32) | (0xFFFFFFFFu - count);
        index[i] = i;
    }

    __global__ inline void chains_item_gather_kernel(std::uint32_t const* sorted_index, std::uint32_t const* item_begin,
                                                     std::uint32_t const* item_end, std::uint32_t nitems,
                                                     std::uint32_t* item_start, std::uint32_t* item_count)
    {
        std::uint32_t const j = blockIdx.x * blockDim.x + threadIdx.x;
        if (j >= nitems)
            return;
        std::uint32_t const i = sorted_index[j];
        item_start[j] = item_begin[i];
        item_count[j] = item_end[i] - item_begin[i];
    }

    __global__ inline void iota_u32_kernel(std::uint32_t* p, std::uint32_t n)
    {
        std::uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < n)
            p[i] = i;
    }

    // Position of every pending cell: chains in flow order (the most upstream cell first).
    __global__ inline void chains_position_kernel(std::uint32_t const* pre, std::uint32_t const* chain_incl,
                                                  std::uint32_t const* chain_pre, std::uint32_t const* length,
                                                  std::uint32_t const* base, std::uint32_t n, std::uint32_t* pos_of,
                                                  std::uint8_t* seg_start)
    {
        std::uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q >= n)
            return;
        std::uint32_t const p = pre[q];
        std::uint32_t const c = chain_incl[p] - 1;
        std::uint32_t const offset = p - chain_pre[c];
        std::uint32_t const at = base[c] + length[c] - 1 - offset;
        pos_of[q] = at;
        if (at < n)
            seg_start[at] = offset + 1 == length[c] ? 1 : 0;
    }

    // Side slots of every pending cell, in neighbour order: pending upstream cells other than the heavy child
    // (the last positions of light chains) and upstream cells in other tiles that round 0 solves (raster
    // cells). Pass 1 counts (fill == false); pass 2 assigns the slots [offset[p], offset[p + 1]) of position p
    // to its sources: push_slot[light child's position] and lo_slot[raster cell's index in the batch].
    template<int TW, int TH>
    __global__ void chains_refs_kernel(Catchment const* cats, int ncats, std::uint32_t const* pending_bits,
                                       std::uint32_t const* cells, std::uint32_t const* index_of,
                                       std::int32_t const* heavy, std::uint32_t const* pos_of, std::uint32_t n,
                                       bool fill, std::uint32_t* count, std::uint32_t const* offset,
                                       std::uint32_t* push_slot, std::uint32_t* lo_slot)
    {
        std::uint32_t const q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q >= n)
            return;
        std::uint32_t const cell = cells[q];
        int const ci = find_catchment(cats, ncats, cell);
        Catchment const cat = cats[ci];
        std::int64_t const local = cell - cat.cell_base;
        std::int64_t const r = local / cat.cols;
        std::int64_t const c = local % cat.cols;
        std::int32_t const h = heavy[q];
        std::uint32_t const p = pos_of[q];
        std::uint32_t k_out = fill ? offset[p] : 0u;
        std::uint32_t m = 0;
        for (int k = 0; k < 8; ++k)
        {
            std::int64_t const r2 = r + neighbour_dr(k);
            std::int64_t const c2 = c + neighbour_dc(k);
            if (r2 < 0 || r2 >= cat.rows || c2 < 0 || c2 >= cat.cols)
                continue;
            std::int64_t const u = r2 * cat.cols + c2;
            std::uint8_t const code = cat.ldd[u];
            if (code == no_data_ldd || ldd_direction(code) != opposite(k))
                continue;
            if (cell_pending<TW, TH>(pending_bits, cat, r2, c2))
            {
                std::uint32_t const uq = index_of[cat.cell_base + u];
                if (static_cast<std::int32_t>(uq) == h)
                    continue;  // the chain predecessor
                if (fill)
                    push_slot[pos_of[uq]] = k_out++;
            }
            else if (r2 / TH != r / TH || c2 / TW != c / TW)
            {
                if (fill)
                    lo_slot[cat.cell_base + u] = k_out++;
            }
            else
            {
                continue;  // same tile, solved by round 0: in the staged side sum
            }
            ++m;
        }
        if (!fill)
            count[p] = m;
    }

    // ---------------------------------------------------------------------------------------------------
    // Execution

    // Chain layout and per-call buffers of one operation.
    template<int OP, typename T>
    struct ChainRun
    {
        std::uint32_t const* item_start;  // per work item: first position
        std::uint32_t const* item_count;  // per work item: positions
        std::uint8_t const* seg_start;    // per position: 1 at the first (most upstream) cell of a chain
        std::uint32_t const* push_slot;   // per position: side slot its value feeds (no_slot: none)
        double* slot;                     // per side slot: the source's outflow (FP64), written by the source
        CellInput<OP, T> const* staged;   // per position: inputs (staged by round 0)
        double* side;                     // per position: FP64 sum of the upstream cells off the chain
        T* res;                           // per position: outflow, in the output type (read by the scatter)
        T* res_rem;                       // per position: remainder (threshold, fraction)
    };

    // has_ref[p] = 1 where position p has side references (p < n), has_ref[n] = 0.
    __global__ inline void chains_has_ref_kernel(std::uint32_t const* ref_off, std::uint32_t n, std::uint32_t* has_ref)
    {
        std::uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
        if (p < n)
            has_ref[p] = ref_off[p + 1] > ref_off[p] ? 1u : 0u;
        else if (p == n)
            has_ref[p] = 0u;
    }

    // ref_pos[rank[p]] = p for the positions with side references.
    __global__ inline void chains_ref_pos_kernel(std::uint32_t const* has_ref, std::uint32_t const* rank, std::uint32_t n,
                                                 std::uint32_t* ref_pos)
    {
        std::uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
        if (p < n && has_ref[p])
            ref_pos[rank[p]] = p;
    }

    // ref_first[i] = ref_off[ref_pos[i]] for the nref_pos positions with references, ref_first[nref_pos] =
    // ref_off[n]: the positions between two of them have none, so entry i's references end where entry i + 1's
    // begin.
    __global__ inline void chains_ref_first_kernel(std::uint32_t const* ref_pos, std::uint32_t const* ref_off,
                                                   std::uint32_t nref_pos, std::uint32_t n, std::uint32_t* ref_first)
    {
        std::uint32_t const i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < nref_pos)
            ref_first[i] = ref_off[ref_pos[i]];
        else if (i == nref_pos)
            ref_first[i] = ref_off[n];
    }

    // Add the side slots of the positions ref_pos[i_begin, i_end) (those of one chain level that have any) to
    // their staged side sums, in slot order (neighbour order). The slots of a position are contiguous and
    // final: written by round 0 (raster cells) and by the runs of lower levels (chain ends). Thread t of a
    // workgroup takes entries base + t and base + t + 256, all loads in flight together.
    constexpr std::uint32_t gather_entries_per_thread = 2;

    template<int OP, typename T>
    __global__ void __launch_bounds__(256)
        chains_gather_kernel(ChainRun<OP, T> c, std::uint32_t const* ref_pos, std::uint32_t const* ref_first,
                             std::uint32_t i_begin, std::uint32_t i_end)
    {
        constexpr int per_thread = static_cast<int>(gather_entries_per_thread);
        constexpr int most = 8;  // a cell has at most 8 upstream neighbours
        std::uint32_t const base = i_begin + blockIdx.x * (256u * per_thread) + threadIdx.x;
        std::uint32_t p[per_thread];
        std::uint32_t begin[per_thread];
        std::uint32_t count[per_thread];
#pragma unroll
        for (int j = 0; j < per_thread; ++j)
        {
            std::uint32_t const i = base + 256u * j;
            bool const ok = i < i_end;
            p[j] = ok ? as_global(ref_pos)[i] : 0u;
            begin[j] = ok ? as_global(ref_first)[i] : 0u;
            count[j] = ok ? as_global(ref_first)[i + 1] - begin[j] : 0u;
        }
        double sum[per_thread];
        double v[per_thread][most];
#pragma unroll
        for (int j = 0; j < per_thread; ++j)
        {
            sum[j] = count[j] != 0u ? as_global(c.side)[p[j]] : 0.0;
#pragma unroll
            for (int k = 0; k < most; ++k)
                v[j][k] = static_cast<std::uint32_t>(k) < count[j] ? as_global(c.slot)[begin[j] + k] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < per_thread; ++j)
        {
            if (count[j] == 0u)
                continue;
#pragma unroll
            for (int k = 0; k < most; ++k)
            {
                if (static_cast<std::uint32_t>(k) < count[j])
                    sum[j] += v[j][k];
            }
            as_global(c.side)[p[j]] = sum[j];
        }
    }

    template<int OP, typename T>
    __device__ inline ChainLane<OP, T> load_chain_lane(SweepArgs<T> const& a, ChainRun<OP, T> const& c,
                                                       std::uint32_t start, std::uint32_t count, std::uint32_t k)
    {
        ChainLane<OP, T> d{};
        d.valid = k < count;
        if (d.valid)
        {
            std::uint32_t const p = start + k;
            d.start = as_global(c.seg_start)[p] != 0;
            d.extra = as_global(c.side)[p];
            d.push = as_global(c.push_slot)[p];
            if constexpr (OP != op_accu)
                d.in = load_staged<OP, T>(a, c.staged, p);  // accu: the inflow is in the side sum, d.in.x stays 0
        }
        return d;
    }

    // One work item of a level on one wavefront, 64 positions at a time. The data of the next `ahead` blocks
    // is in flight while a block runs: 2 for the kinematic wave, whose block solve is long enough to cover
    // the memory latency, 6 for the scans of the other operations. The ring of blocks rotates through a loop
    // unrolled over it, so no load is waited for before its block is due.
    template<int OP, typename T>
    __device__ inline void run_item_blocks(SweepArgs<T> const& a, ChainRun<OP, T> const& c, std::uint32_t item)
    {
        constexpr bool two_outputs = is_split_op(OP);
        constexpr int ahead = OP == op_kinematic_wave ? 2 : 6;
        int const lane = threadIdx.x & 63;
        std::uint32_t const start = as_global(c.item_start)[item];
        std::uint32_t const count = as_global(c.item_count)[item];
        BlockState st;
        ChainLane<OP, T> ring[ahead];
#pragma unroll
        for (int k = 0; k < ahead; ++k)
            ring[k] = load_chain_lane<OP, T>(a, c, start, count, 64 * k + lane);
        for (std::uint32_t base = 0; base < count; base += 64 * ahead)
        {
#pragma unroll
            for (int k = 0; k < ahead; ++k)
            {
                std::uint32_t const block = base + 64 * k;
                if (block >= count)
                    break;
                ChainLane<OP, T> const d = ring[k];
                if (block + 64 * ahead < count)
                    ring[k] = load_chain_lane<OP, T>(a, c, start, count, block + 64 * ahead + lane);
                int const nb = static_cast<int>(count - block < 64 ? count - block : 64);
                double value;
                T rem{};
                solve_block<OP, T>(a, d, nb, st, value, rem);
                if (d.valid)
                {
                    std::uint32_t const p = start + block + lane;
                    as_global(c.res)[p] = static_cast<T>(value);
                    if constexpr (two_outputs)
                        as_global(c.res_rem)[p] = rem;
                    if (d.push != no_slot)
                        as_global(c.slot)[d.push] = value;
                }
            }
        }
    }

    // One level of chains, a wavefront per work item (4 per workgroup).
    template<int OP, typename T>
    __global__ void __launch_bounds__(256)
        chains_run_kernel(SweepArgs<T> a, ChainRun<OP, T> c, std::uint32_t item_begin, std::uint32_t item_end)
    {
        std::uint32_t const item = item_begin + blockIdx.x * 4 + (threadIdx.x >> 6);
        if (item >= item_end)
            return;
        run_item_blocks<OP, T>(a, c, item);
    }

    // Lane walk: a wavefront takes 64 work items (one workgroup of one wavefront), each lane walks its item
    // position after position. Every cell sees its upstream outflow exactly, as in LUE's walk, and the cell
    // function runs once per cell instead of once per Newton step of a 64-cell block, which suits the many
    // short items of the low levels. The items' data moves in chunks of lane_chunk positions per item through
    // LDS: the wavefront loads (and finally stores) runs of consecutive positions of one item with
    // consecutive lanes, so that its memory instructions touch few lines, and each lane then reads its own
    // item's positions from LDS.
    template<int OP, typename T>
    constexpr int lane_chunk()
    {
        return OP == op_kinematic_wave ? (std::is_same_v<T, float> ? 4 : 2) : 8;
    }

    template<int OP, typename T>
    struct LaneWalkLds
    {
        static constexpr int C = lane_chunk<OP, T>();
        static constexpr int stride = C + 1;  // per item (padded against bank conflicts)
        static constexpr int n = 64 * stride;
        double value[n];  // side sums, then the outflows
        std::uint32_t push[n];
        T rem[is_split_op(OP) ? n : 1];
        CellInput<OP, T> in[OP == op_accu ? 1 : n];
        std::uint8_t start[n];
    };

    template<int OP, typename T>
    __device__ inline void run_items_lane_walk(SweepArgs<T> const& a, ChainRun<OP, T> const& c, LaneWalkLds<OP, T>& s,
                                               std::uint32_t first_item, std::uint32_t item_end)
    {
        using L = LaneWalkLds<OP, T>;
        constexpr int C = L::C;
        constexpr bool two_outputs = is_split_op(OP);
        int const lane = threadIdx.x & 63;
        std::uint32_t const item = first_item + static_cast<std::uint32_t>(lane);
        bool const mine = item < item_end;
        std::uint32_t const start = mine ? as_global(c.item_start)[item] : 0u;
        std::uint32_t const count = mine ? as_global(c.item_count)[item] : 0u;
        std::uint32_t most = count;
        for (int offset = 32; offset > 0; offset >>= 1)
        {
            std::uint32_t const other = __shfl_xor(most, offset, 64);
            most = other > most ? other : most;
        }
        double carry = 0.0;
        for (std::uint32_t k0 = 0; k0 < most; k0 += C)
        {
            // load: record r = j * 64 + lane is position k0 + r % C of the wavefront's item r / C
#pragma unroll
            for (int j = 0; j < C; ++j)
            {
                int const r = j * 64 + lane;
                int const il = r / C;
                int const pos = r % C;
                std::uint32_t const s_il = __shfl(start, il, 64);
                std::uint32_t const n_il = __shfl(count, il, 64);
                std::uint32_t const k = k0 + static_cast<std::uint32_t>(pos);
                if (k < n_il)
                {
                    std::uint32_t const p = s_il + k;
                    int const idx = il * L::stride + pos;
                    s.value[idx] = as_global(c.side)[p];
                    s.push[idx] = as_global(c.push_slot)[p];
                    s.start[idx] = as_global(c.seg_start)[p];
                    if constexpr (OP != op_accu)
                        s.in[idx] = load_staged<OP, T>(a, c.staged, p);  // accu: the inflow is in the side sum
                }
            }
            __syncthreads();
            // walk: this lane's item, positions k0 .. k0 + C - 1
#pragma unroll
            for (int pos = 0; pos < C; ++pos)
            {
                if (k0 + static_cast<std::uint32_t>(pos) >= count)
                    break;
                int const idx = lane * L::stride + pos;
                double const up = (s.start[idx] ? 0.0 : carry) + s.value[idx];
                CellInput<OP, T> in{};
                if constexpr (OP != op_accu)
                    in = s.in[idx];
                double value;
                T rem{};
                solve_lane<OP, T>(a, in, up, value, rem);
                s.value[idx] = value;
                if constexpr (two_outputs)
                    s.rem[idx] = rem;
                std::uint32_t const push = s.push[idx];
                if (push != no_slot)
                    as_global(c.slot)[push] = value;
                carry = value;
            }
            __syncthreads();
            // store the outflows (and remainders) as they were loaded
#pragma unroll
            for (int j = 0; j < C; ++j)
            {
                int const r = j * 64 + lane;
                int const il = r / C;
                int const pos = r % C;
                std::uint32_t const s_il = __shfl(start, il, 64);
                std::uint32_t const n_il = __shfl(count, il, 64);
                std::uint32_t const k = k0 + static_cast<std::uint32_t>(pos);
                if (k < n_il)
                {
                    std::uint32_t const p = s_il + k;
                    int const idx = il * L::stride + pos;
                    as_global(c.res)[p] = static_cast<T>(s.value[idx]);
                    if constexpr (two_outputs)
                        as_global(c.res_rem)[p] = s.rem[idx];
                }
            }
            __syncthreads();
        }
    }

    // One level of chains (or its short items), 64 work items per workgroup of one wavefront.
    template<int OP, typename T>
    __global__ void __launch_bounds__(64)
        chains_run_lanes_kernel(SweepArgs<T> a, ChainRun<OP, T> c, std::uint32_t item_begin, std::uint32_t item_end)
    {
        __shared__ LaneWalkLds<OP, T> s;
        run_items_lane_walk<OP, T>(a, c, s, item_begin + blockIdx.x * 64u, item_end);
    }

    // One level of chains whose work items are sorted longest first: items [item_begin, item_split) one per
    // workgroup of one wavefront (the first workgroups, so the long items start first), the short ones
    // [item_split, item_end) 64 per workgroup by the lane walk.
    template<int OP, typename T>
    __global__ void __launch_bounds__(64)
        chains_run_mixed_kernel(SweepArgs<T> a, ChainRun<OP, T> c, std::uint32_t item_begin, std::uint32_t item_split,
                                std::uint32_t item_end)
    {
        __shared__ LaneWalkLds<OP, T> s;
        std::uint32_t const block_groups = item_split - item_begin;
        if (blockIdx.x < block_groups)
        {
            run_item_blocks<OP, T>(a, c, item_begin + blockIdx.x);
            return;
        }
        run_items_lane_walk<OP, T>(a, c, s, item_split + (blockIdx.x - block_groups) * 64u, item_end);
    }

    // Write the chain results of the pending cells of every tile to the output rasters. One wavefront per tile
    // (a tile has about a third of its cells pending), scatter_tiles_per_block<...>() tiles per workgroup: the
    // results are read in chain-position order (sorted_pos: runs of consecutive positions, so whole lines) into
    // the wavefront's part of LDS, then written in cell order (the scatter entries in cell order give the cell
    // and the result's place in LDS), so that the stores of a row of the tile are one line. The tile's
    // catchment and IO entry are loaded with its bounds, so a wavefront waits for three rounds of loads.
    template<int OP, typename T, int TW, int TH>
    constexpr int scatter_tiles_per_block()
    {
        constexpr int bytes = TW * TH * static_cast<int>(sizeof(T)) * (is_split_op(OP) ? 2 : 1);
        constexpr int fit = 16384 / bytes;
        return fit < 1 ? 1 : (fit > 4 ? 4 : fit);
    }

    template<int OP, typename T, int TW, int TH>
    __global__ void __launch_bounds__((64 * scatter_tiles_per_block<OP, T, TW, TH>()))
        chains_scatter_kernel(SweepArgs<T> a, std::uint32_t const* tile_off, std::uint32_t const* scatter_entry,
                              std::uint32_t const* sorted_pos, T const* res, T const* res_rem, std::int64_t ntiles)
    {
        using G = Tile<TW, TH>;
        constexpr bool two_outputs = is_split_op(OP);
        constexpr int waves = scatter_tiles_per_block<OP, T, TW, TH>();
        constexpr int chunk = 8;  // entries per lane with loads in flight together
        static_assert(G::n % (64 * chunk) == 0, "tile cells a multiple of 512");
        __shared__ T value[waves][G::n];
        __shared__ T remainder[waves][two_outputs ? G::n : 1];
        int const lane = static_cast<int>(threadIdx.x & 63);
        int const wave = static_cast<int>(threadIdx.x >> 6);
        std::int64_t const tile = static_cast<std::int64_t>(blockIdx.x) * waves + wave;
        if (tile >= ntiles)
            return;
        std::uint32_t const begin = as_global(tile_off)[tile];
        std::uint32_t const npending = as_global(tile_off)[tile + 1] - begin;
        Catchment cat;
        CatchmentIO<T> io;
        tile_catchment(a, tile, cat, io);
        if (npending == 0)
            return;
        T* const v_lds = value[wave];
        T* const r_lds = remainder[wave];
        // the scatter entries of the first chunk of ranks, loaded with the first positions (they do not depend
        // on the results)
        std::uint32_t entry0[chunk];
#pragma unroll
        for (int j = 0; j < chunk; ++j)
        {
            std::uint32_t const rank = static_cast<std::uint32_t>(64 * j + lane);
            entry0[j] = rank < npending ? as_global(scatter_entry)[begin + rank] : 0u;
        }
        // chain-position order: positions, then results, a chunk of entries per lane at a time
        for (std::uint32_t base = 0; base < npending; base += 64 * chunk)
        {
            std::uint32_t pos[chunk];
#pragma unroll
            for (int j = 0; j < chunk; ++j)
            {
                std::uint32_t const k = base + static_cast<std::uint32_t>(64 * j + lane);
                pos[j] = k < npending ? as_global(sorted_pos)[begin + k] : 0u;
            }
            T v[chunk];
            T r[chunk];
#pragma unroll
            for (int j = 0; j < chunk; ++j)
            {
                std::uint32_t const k = base + static_cast<std::uint32_t>(64 * j + lane);
                if (k < npending)
                {
                    v[j] = as_global(res)[pos[j]];
                    if constexpr (two_outputs)
                        r[j] = as_global(res_rem)[pos[j]];
                }
            }
#pragma unroll
            for (int j = 0; j < chunk; ++j)
            {
                std::uint32_t const k = base + static_cast<std::uint32_t>(64 * j + lane);
                if (k < npending)
                {
                    v_lds[k] = v[j];
                    if constexpr (two_outputs)
                        r_lds[k] = r[j];
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        std::uint32_t const local = static_cast<std::uint32_t>(tile - cat.tile_base);
        std::uint32_t const tiles_x = static_cast<std::uint32_t>(cat.tiles_x);
        std::uint32_t const ty = local / tiles_x;
        std::uint32_t const tx = local - ty * tiles_x;
        std::uint32_t const cols32 = static_cast<std::uint32_t>(cat.cols);
        std::int64_t const origin = static_cast<std::int64_t>(ty * TH) * cat.cols + tx * TW;
        // cell order: entries of a chunk of ranks per lane, then the stores
        for (std::uint32_t base = 0; base < npending; base += 64 * chunk)
        {
            std::uint32_t cell[chunk];
            std::uint32_t at[chunk];
#pragma unroll
            for (int j = 0; j < chunk; ++j)
            {
                std::uint32_t const rank = base + static_cast<std::uint32_t>(64 * j + lane);
                std::uint32_t e = entry0[j];
                if (base != 0)
                    e = rank < npending ? as_global(scatter_entry)[begin + rank] : 0u;
                cell[j] = e & (G::n - 1);
                at[j] = (e >> 16) & (G::n - 1);
            }
#pragma unroll
            for (int j = 0; j < chunk; ++j)
            {
                std::uint32_t const rank = base + static_cast<std::uint32_t>(64 * j + lane);
                if (rank >= npending)
                    break;
                std::int64_t const g = origin + static_cast<std::int64_t>(
                                                    static_cast<std::uint64_t>(cell[j] / TW) * cols32 + cell[j] % TW);
                as_global(io.out)[g] = v_lds[at[j]];
                if constexpr (two_outputs)
                    as_global(io.rem)[g] = r_lds[at[j]];
            }
        }
    }

}  // namespace lue_amd
