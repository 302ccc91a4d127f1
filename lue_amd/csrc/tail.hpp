// Chain-streaming tail of the routing sweep.
//
// After the first rounds of the tiled sweep (sweep.hpp) almost every cell is solved; what is left are
// the long flow paths that cross many tiles (main streams). Solving them one tile per round would cost
// one launch per tile crossing. Instead the pending cells are cut into chains and streamed:
//   1. compact: list the pending cells (pending bitmaps of the tiles), map cell -> list index;
//   2. links: per pending cell, its downstream pending cell and the number of pending upstream cells;
//      a chain continues from a cell into its downstream cell when that cell has exactly one pending
//      upstream cell, and starts at every other pending cell (no pending upstream: ready; two or more:
//      a confluence of chains);
//   3. list ranking by pointer jumping gives every cell its distance to the end of its chain and the
//      chain's last cell, from which chains are laid out contiguously in flow order;
//   4. execute: one wavefront per chain. The wavefront loads 64 consecutive chain cells at a time
//      (cell index, the outflow of the upstream neighbours that are not on the chain, the inputs),
//      then runs the cell function serially along the block (values broadcast with readlane).
//      When a chain ends at a confluence, the wavefront decrements the confluence's arrival counter
//      (agent-scope acquire/release, once per chain); the last arriving wavefront continues with
//      the chain that starts there. No wavefront ever waits, so the order of work cannot deadlock.
// Kinematic wave, FP32 path: the serial part is the FP32 Newton solve only. After each block the 64
// lanes compute, in parallel, the FP64 residual correction of their cell and its sensitivity to the
// upstream discharge; the corrections are propagated along the block (affine recurrence, first order
// in the FP32 error) and the corrected discharge becomes the start of the next block.
#pragma once

#include "kinematic_wave.hpp"
#include "routing_common.hpp"
#include "sweep.hpp"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace lue_amd {

    struct TailChains
    {
        std::uint32_t const* cells;     // chain cells (global cell ids) in flow order, chains back to back
        std::uint32_t const* chain_cat; // per chain: catchment
        std::uint32_t const* base;      // per chain: offset into cells
        std::uint32_t const* length;    // per chain
        std::int32_t const* downstream; // per chain: chain starting at the confluence below the last cell, or -1
        std::int32_t* arrivals;         // per chain: chains still to arrive at its first cell
        std::int32_t const* ready;      // chains without pending upstream
        std::uint32_t nready;
        std::uint32_t* cursor;          // next index into `ready`
        unsigned long long* solved;     // cells solved by the tail
    };

    template<int TW, int TH>
    __device__ inline bool cell_pending(std::uint32_t const* pending_bits, Catchment const& cat, std::int64_t r,
                                        std::int64_t c)
    {
        using G = Tile<TW, TH>;
        std::int64_t const tile = cat.tile_base + (r / TH) * cat.tiles_x + (c / TW);
        int const i = static_cast<int>((r % TH) * TW + (c % TW));
        return (pending_bits[tile * G::words + (i >> 5)] >> (i & 31)) & 1u;
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

    // Pending cells per tile.
    template<int TW, int TH>
    __global__ void tail_count_kernel(std::uint32_t const* pending_bits, std::int64_t ntiles, std::uint32_t* counts)
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

    // List the pending cells of each tile at its offset; record global cell id -> list index.
    template<int TW, int TH>
    __global__ void tail_compact_kernel(std::uint32_t const* pending_bits, Catchment const* cats,
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

    // Downstream pending cell of every pending cell, and pending upstream counts.
    template<int TW, int TH>
    __global__ void tail_links_kernel(Catchment const* cats, int ncats, std::uint32_t const* pending_bits,
                                      std::uint32_t const* cells, std::uint32_t const* index_of, std::uint32_t n,
                                      std::int32_t* next, std::int32_t* upstream_count)
    {
        std::uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
        if (p >= n)
            return;
        std::uint32_t const cell = cells[p];
        Catchment const cat = cats[find_catchment(cats, ncats, cell)];
        std::int64_t const local = cell - cat.cell_base;
        std::int64_t const r = local / cat.cols;
        std::int64_t const c = local % cat.cols;
        int const dir = ldd_direction(cat.ldd[local]);
        std::int32_t nx = -1;
        if (dir >= 0)
        {
            std::int64_t const r2 = r + neighbour_dr(dir);
            std::int64_t const c2 = c + neighbour_dc(dir);
            if (r2 >= 0 && r2 < cat.rows && c2 >= 0 && c2 < cat.cols && cat.ldd[r2 * cat.cols + c2] != no_data_ldd &&
                cell_pending<TW, TH>(pending_bits, cat, r2, c2))
            {
                nx = static_cast<std::int32_t>(index_of[cat.cell_base + r2 * cat.cols + c2]);
                atomicAdd(&upstream_count[nx], 1);
            }
        }
        next[p] = nx;
    }

    // Chain successor (link), and the initial state of list ranking.
    __global__ void tail_rank_init_kernel(std::int32_t const* next, std::int32_t const* upstream_count, std::uint32_t n,
                                          std::int32_t* succ, std::int32_t* rank, std::int32_t* last)
    {
        std::uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
        if (p >= n)
            return;
        std::int32_t const nx = next[p];
        bool const linked = nx >= 0 && upstream_count[nx] == 1;
        succ[p] = linked ? nx : -1;
        rank[p] = linked ? 1 : 0;
        last[p] = linked ? nx : static_cast<std::int32_t>(p);
    }

    // One pointer-jumping step (double buffered).
    __global__ void tail_rank_step_kernel(std::int32_t const* succ, std::int32_t const* rank, std::int32_t const* last,
                                          std::uint32_t n, std::int32_t* succ2, std::int32_t* rank2,
                                          std::int32_t* last2)
    {
        std::uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
        if (p >= n)
            return;
        std::int32_t const s = succ[p];
        if (s >= 0)
        {
            succ2[p] = succ[s];
            rank2[p] = rank[p] + rank[s];
            last2[p] = last[s];
        }
        else
        {
            succ2[p] = -1;
            rank2[p] = rank[p];
            last2[p] = last[p];
        }
    }

    // Chain ids: chains are numbered by their last cell; is_last[p] = 1 for last cells.
    __global__ void tail_flag_last_kernel(std::int32_t const* next, std::int32_t const* upstream_count, std::uint32_t n,
                                          std::uint32_t* is_last)
    {
        std::uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
        if (p >= n)
            return;
        std::int32_t const nx = next[p];
        is_last[p] = (nx >= 0 && upstream_count[nx] == 1) ? 0u : 1u;
    }

    // Per chain (numbered by the exclusive scan of is_last): length, first cell, downstream link.
    __global__ void tail_chain_kernel(std::int32_t const* next, std::int32_t const* upstream_count,
                                      std::int32_t const* rank, std::int32_t const* last, std::uint32_t const* chain_of_last,
                                      std::uint32_t const* cells, Catchment const* cats, int ncats, std::uint32_t n,
                                      std::uint32_t* length, std::int32_t* first, std::int32_t* arrivals,
                                      std::uint32_t* chain_cat)
    {
        std::uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
        if (p >= n)
            return;
        std::int32_t const u = upstream_count[p];
        if (u != 1)
        {
            // p starts a chain; its end is a chain end (a cycle without entry has neither)
            std::uint32_t const chain = chain_of_last[last[p]];
            length[chain] = static_cast<std::uint32_t>(rank[p]) + 1;
            first[chain] = static_cast<std::int32_t>(p);
            arrivals[chain] = u;
            chain_cat[chain] = static_cast<std::uint32_t>(find_catchment(cats, ncats, cells[p]));
        }
    }

    __global__ void tail_downstream_kernel(std::int32_t const* next, std::int32_t const* last,
                                           std::uint32_t const* chain_of_last, std::uint32_t const* is_last,
                                           std::uint32_t n, std::int32_t* downstream)
    {
        std::uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
        if (p >= n || !is_last[p])
            return;
        std::uint32_t const chain = chain_of_last[p];
        std::int32_t const nx = next[p];
        // the chain below starts at nx (a confluence); it is numbered by its own last cell
        downstream[chain] = nx >= 0 ? static_cast<std::int32_t>(chain_of_last[last[nx]]) : -1;
    }

    // Lay the cells out per chain in flow order: position = length - 1 - rank.
    // Cells on a cycle without entry have no chain end; they are skipped here and never solved (the
    // caller reports an invalid network).
    __global__ void tail_layout_kernel(std::uint32_t const* cells, std::int32_t const* rank, std::int32_t const* last,
                                       std::uint32_t const* is_last, std::uint32_t const* chain_of_last,
                                       std::uint32_t const* base, std::uint32_t const* length, std::uint32_t n,
                                       std::uint32_t* chain_cells)
    {
        std::uint32_t const p = blockIdx.x * blockDim.x + threadIdx.x;
        if (p >= n)
            return;
        std::int32_t const end = last[p];
        if (end < 0 || static_cast<std::uint32_t>(end) >= n || !is_last[end])
            return;
        std::uint32_t const chain = chain_of_last[end];
        std::uint32_t const len = length[chain];
        std::int32_t const r = rank[p];
        if (len == 0 || r < 0 || static_cast<std::uint32_t>(r) >= len)
            return;
        std::uint32_t const pos = base[chain] + len - 1 - static_cast<std::uint32_t>(r);
        if (pos < n)
            chain_cells[pos] = cells[p];
    }

    __global__ void tail_ready_kernel(std::int32_t const* arrivals, std::uint32_t nchains, std::int32_t* ready,
                                      std::uint32_t* nready)
    {
        std::uint32_t const chain = blockIdx.x * blockDim.x + threadIdx.x;
        if (chain >= nchains || arrivals[chain] != 0)
            return;
        ready[atomicAdd(nready, 1u)] = static_cast<std::int32_t>(chain);
    }

    // ---------------------------------------------------------------------------------------------------
    // Execution

    __device__ inline double read_lane(double v, int lane)
    {
        std::uint64_t const bits = __double_as_longlong(v);
        std::uint32_t const lo = __builtin_amdgcn_readlane(static_cast<std::uint32_t>(bits), lane);
        std::uint32_t const hi = __builtin_amdgcn_readlane(static_cast<std::uint32_t>(bits >> 32), lane);
        return __longlong_as_double((static_cast<std::uint64_t>(hi) << 32) | lo);
    }

    __device__ inline float read_lane(float v, int lane)
    {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
    }

    __device__ inline int read_lane(int v, int lane)
    {
        return __builtin_amdgcn_readlane(v, lane);
    }

    // Sum (FP64, fixed neighbour order) of the outflow of the upstream neighbours of local cell `cell`,
    // except local cell `skip`.
    template<typename T>
    __device__ inline double side_inflow(Catchment const& cat, CatchmentIO<T> const& io, float const* out_lo,
                                         std::int64_t cell, std::int64_t skip, bool& poisoned)
    {
        std::int64_t const r = cell / cat.cols;
        std::int64_t const c = cell % cat.cols;
        double sum = 0.0;
        for (int k = 0; k < 8; ++k)
        {
            std::int64_t const r2 = r + neighbour_dr(k);
            std::int64_t const c2 = c + neighbour_dc(k);
            if (r2 < 0 || r2 >= cat.rows || c2 < 0 || c2 >= cat.cols)
                continue;
            std::int64_t const u = r2 * cat.cols + c2;
            std::uint8_t const code = cat.ldd[u];
            if (code == no_data_ldd || ldd_direction(code) != opposite(k) || u == skip)
                continue;
            double v = static_cast<double>(io.out[u]);
            if constexpr (std::is_same_v<T, float>)
                v += static_cast<double>(out_lo[u]);
            if (isnan(v))
                poisoned = true;
            else
                sum += v;
        }
        return sum;
    }

    // Inputs of one lane's cell in a 64-cell block of a chain.
    template<int OP, typename T>
    struct BlockLane
    {
        std::int64_t cell;  // local cell index in its catchment
        bool valid;
        bool bad;           // an upstream neighbour off the chain is no-data
        double side;        // FP64 sum of the outflow of the upstream neighbours off the chain
        CellInput<OP, T> in;
    };

    template<int OP, typename T>
    __device__ inline BlockLane<OP, T> load_block_lane(SweepArgs<T> const& a, TailChains const& t, Catchment const& cat,
                                                       CatchmentIO<T> const& io, float const* out_lo,
                                                       std::uint32_t base, std::uint32_t length, std::uint32_t k)
    {
        BlockLane<OP, T> d{};
        d.valid = k < length;
        if (d.valid)
        {
            d.cell = static_cast<std::int64_t>(t.cells[base + k]) - cat.cell_base;
            std::int64_t const prev = k > 0 ? static_cast<std::int64_t>(t.cells[base + k - 1]) - cat.cell_base : -1;
            d.side = side_inflow<T>(cat, io, out_lo, d.cell, prev, d.bad);
            stage_input<OP, T>(a, io, d.cell, d.in);
        }
        return d;
    }

    // One wavefront per chain, 64-cell blocks. The data of the next block is loaded while the current
    // block runs its serial part.
    template<int OP, typename T>
    __global__ void __launch_bounds__(256) tail_execute_kernel(SweepArgs<T> a, TailChains t)
    {
        constexpr bool compensated = std::is_same_v<T, float>;
        constexpr bool two_outputs = OP == op_threshold || OP == op_fraction;
        int const lane = threadIdx.x & 63;
        int chain = -1;
        unsigned long long solved = 0;
        double const nan64 = __longlong_as_double(0x7ff8000000000000ll);
        while (true)
        {
            if (chain < 0)
            {
                int next = -1;
                if (lane == 0)
                {
                    std::uint32_t const slot = atomicAdd(t.cursor, 1u);
                    next = slot < t.nready ? t.ready[slot] : -1;
                }
                chain = read_lane(next, 0);
                if (chain < 0)
                    break;
            }
            std::uint32_t const base = t.base[chain];
            std::uint32_t const length = t.length[chain];
            std::uint32_t const ci = t.chain_cat[chain];
            Catchment const cat = a.cats[ci];
            CatchmentIO<T> const io = a.io[ci];
            float* const out_lo = compensated ? a.out_lo + cat.cell_base : nullptr;
            double carry = 0.0;  // outflow of the previous chain cell (FP64)
            float warm = 0.0f;   // kinematic wave FP32: an iterate of the previous cell, to start Newton from
            float warm_pow = 0.0f;
            bool have_warm = false;
            BlockLane<OP, T> next_block = load_block_lane<OP, T>(a, t, cat, io, out_lo, base, length, lane);
            for (std::uint32_t block = 0; block < length; block += 64)
            {
                BlockLane<OP, T> const d = next_block;
                if (block + 64 < length)
                    next_block = load_block_lane<OP, T>(a, t, cat, io, out_lo, base, length, block + 64 + lane);
                int const nb = static_cast<int>(length - block < 64 ? length - block : 64);
                double value = 0.0;  // this lane's cell outflow (FP64)
                T rem{};

                if constexpr (OP == op_accu)
                {
                    // linear: out_k = out_{k-1} + side_k + x_k, an inclusive scan along the block
                    double term = (!d.valid) ? 0.0 : ((d.bad || isnan(d.in.x)) ? nan64 : d.side + static_cast<double>(d.in.x));
                    for (int offset = 1; offset < 64; offset <<= 1)
                    {
                        double const prev = __shfl_up(term, offset, 64);
                        if (lane >= offset)
                            term += prev;
                    }
                    value = carry + term;
                    carry = read_lane(value, nb - 1);
                }
                else if constexpr (OP == op_threshold || OP == op_fraction)
                {
                    double const term = (d.bad || isnan(d.in.x) || isnan(d.in.p)) ? nan64 : d.side + static_cast<double>(d.in.x);
                    for (int j = 0; j < nb; ++j)
                    {
                        double const total = carry + read_lane(term, j);
                        double const p = static_cast<double>(read_lane(d.in.p, j));
                        double out;
                        double r;
                        if constexpr (OP == op_threshold)
                        {
                            bool const over = total > p;
                            out = isnan(total) ? total : (over ? total - p : 0.0);
                            r = over ? p : total;
                        }
                        else
                        {
                            out = p * total;
                            r = total - out;
                        }
                        if (lane == j)
                        {
                            value = out;
                            rem = isnan(total) ? quiet_nan<T>() : static_cast<T>(r);
                        }
                        carry = out;
                    }
                }
                else if constexpr (std::is_same_v<T, double>)
                {
                    // plain FP64 kinematic wave: LUE's Newton per cell along the block
                    for (int j = 0; j < nb; ++j)
                    {
                        double const s = read_lane(d.side, j);
                        int const bad = read_lane(static_cast<int>(d.bad), j);
                        CellInput<OP, T> cj{read_lane(d.in.q_cur, j), read_lane(d.in.q_lat, j), read_lane(d.in.alpha, j),
                                            read_lane(d.in.beta, j), read_lane(d.in.b, j), read_lane(d.in.c1, j)};
                        double v;
                        T r{};
                        solve_cell<OP, T>(a, cj, carry + s, bad != 0 || isnan(carry), v, r);
                        if (lane == j)
                            value = v;
                        carry = v;
                    }
                }
                else
                {
                    // FP32 kinematic wave: serial FP32 Newton (warm start from the previous cell), then the
                    // FP64 correction of every cell in parallel, propagated along the block to first order
                    double q_up_mine = 0.0;
                    float q32_mine = 0.0f;
                    bool const uniform_beta = io.beta == nullptr;
                    for (int j = 0; j < nb; ++j)
                    {
                        double const q_up = carry + read_lane(d.side, j);
                        float const q_cur = read_lane(d.in.q_cur, j);
                        int const bad = read_lane(static_cast<int>(d.bad), j);
                        float q;
                        if (bad || isnan(q_up) || isnan(q_cur))
                        {
                            q = __int_as_float(0x7fc00000);
                            have_warm = false;
                        }
                        else
                        {
                            float const q_lat = read_lane(d.in.q_lat, j);
                            float const alpha = read_lane(d.in.alpha, j);
                            float const beta = read_lane(d.in.beta, j);
                            double const b = read_lane(d.in.b, j);
                            double const c1 = read_lane(d.in.c1, j);
                            have_warm = have_warm && uniform_beta;
                            q = kinematic_wave_root_f32(static_cast<float>(c1), alpha, beta, a.dt, q_up, q_cur, q_lat, b,
                                                        c1, have_warm, warm, warm_pow);
                        }
                        if (lane == j)
                        {
                            q_up_mine = q_up;
                            q32_mine = q;
                        }
                        carry = static_cast<double>(q);
                    }
                    // FP64 correction in parallel: eps = G(q_up) - q, sens = dG/dq_up (per lane)
                    double eps = 0.0;
                    double sens = 0.0;
                    double const q = static_cast<double>(q32_mine);
                    if (d.valid && q > 0.0 && !isnan(q))
                    {
                        kinematic_wave_correction(q, q_up_mine, static_cast<double>(d.in.alpha),
                                                  static_cast<double>(d.in.beta), d.in.b, d.in.c1, eps, sens);
                    }
                    // affine scan along the block: delta_k = eps_k + sens_k * delta_{k-1}
                    double mul = sens;
                    double add = eps;
                    for (int offset = 1; offset < 64; offset <<= 1)
                    {
                        double const m_prev = __shfl_up(mul, offset, 64);
                        double const a_prev = __shfl_up(add, offset, 64);
                        if (lane >= offset)
                        {
                            add = add + mul * a_prev;
                            mul = mul * m_prev;
                        }
                    }
                    value = isnan(q) ? q : (d.valid ? (q + add > 0.0 ? q + add : q) : 0.0);
                    // the corrected last value starts the next block
                    carry = read_lane(value, nb - 1);
                }

                if (d.valid)
                {
                    T const hi = static_cast<T>(value);
                    io.out[d.cell] = hi;
                    if constexpr (compensated)
                        out_lo[d.cell] = isnan(value) ? 0.0f : static_cast<float>(value - static_cast<double>(hi));
                    if constexpr (two_outputs)
                        io.rem[d.cell] = rem;
                }
                solved += d.valid ? 1 : 0;
            }
            // end of the chain: arrive at the confluence below; the last arrival continues there.
            // The release covers the stores of all lanes of the wavefront.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            int next = -1;
            if (lane == 0)
            {
                int const down = t.downstream[chain];
                if (down >= 0)
                {
                    int const before = __hip_atomic_fetch_add(&t.arrivals[down], -1, __ATOMIC_ACQ_REL,
                                                              __HIP_MEMORY_SCOPE_AGENT);
                    next = before == 1 ? down : -1;
                }
            }
            chain = read_lane(next, 0);
            if (chain >= 0)
            {
                // acquire for every lane of the wavefront, not only the one that did the atomic
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
        }
        // count solved cells (lanes hold their own counts)
        for (int offset = 32; offset > 0; offset >>= 1)
            solved += __shfl_down(solved, offset, 64);
        if (lane == 0)
            atomicAdd(t.solved, solved);
    }

}  // namespace lue_amd
