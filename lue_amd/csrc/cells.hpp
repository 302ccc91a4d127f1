// Per-cell definitions shared by the tiled sweep (sweep.hpp) and the chain phase (chains.hpp): the
// catchments of a batch, the arguments of one operation, the per-cell inputs and the cell functions of
// LUE's accumulating router (reference: source/framework/algorithm/include/lue/framework/algorithm/
// definition/{accu,accu_threshold,accu_fraction,kinematic_wave}.hpp).
#pragma once

#include "kinematic_wave.hpp"
#include "routing_common.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <limits>
#include <type_traits>

namespace lue_amd {

    constexpr std::uint32_t stamp_pending = 0xFFFFFFFFu;

    // Workgroup of the tile kernels: 4 wavefronts per tile (measured at 16384^2: faster than 8, whose
    // shorter tile latency is lost to fewer tiles resident per CU).
    constexpr int sweep_threads = 256;
    constexpr int sweep_waves = sweep_threads / 64;

    // One raster of a batch.
    struct Catchment
    {
        std::uint8_t const* ldd;
        std::int64_t rows;
        std::int64_t cols;
        int tiles_x;
        int tiles_y;
        std::int64_t tile_base;  // global id of the first tile
        std::int64_t cell_base;  // offset of the catchment's cells in per-cell scratch of the batch
    };

    // Inputs and outputs of one catchment (raster pointers; nullptr: the scalar of SweepArgs).
    template<typename T>
    struct CatchmentIO
    {
        T const* in1;  // accu*: inflow        kinematic wave: current outflow
        T const* in2;  // threshold | fraction | kinematic wave: lateral inflow
        T const* alpha;
        T const* beta;
        T const* dx;
        T* out;        // outflow
        T* rem;        // remainder (threshold, fraction)
    };

    template<typename T>
    struct SweepArgs
    {
        Catchment const* cats;
        CatchmentIO<T> const* io;
        std::uint32_t const* tile_cat;  // catchment of every global tile
        int ncats;
        std::int64_t ntiles;

        T in1_s;
        T in2_s;
        T alpha_s;
        T beta_s;
        T dx_s;
        T dt;
        T c1_s;  // dt / dx_s: the kinematic wave's c1 of every cell where the channel length is the scalar

        float* out_lo;  // compensation of `out` (float path) by global cell, for cells read across rounds

        std::uint32_t* pending_bits;   // per tile: TW*TH/32 words, bit set = cell not solved yet
        std::uint32_t* border_stamp;   // per tile: perimeter cells, round in which the cell was solved
        std::uint32_t* tile_flag;      // 3 x tiles: queued for round r (index r % 3); recording only
        std::uint32_t* worklist;       // 3 x tiles; recording only
        std::uint32_t* work_count;     // 3 counters; recording only
        unsigned long long* counters;  // [0] solved cells, [1] no-data cells; recording only
        int record;                    // queue tiles and count cells (schedule recording)

        // Chain schedule, round 0 of a run: the inputs of the cells left pending are staged at their
        // chain positions (null stage_off: no staging).
        std::uint32_t const* stage_off;  // per tile (+1): first compact index of its pending cells
        std::uint32_t const* stage_pos;  // per staging entry (tile after tile): chain position
        void* staged;                    // per chain position: CellInput<OP, T> (or its compact form)
        // the FP32 kinematic wave with beta and the channel length the scalars of every catchment: the staged
        // inputs are CompactKwInput (16 B), beta and c1 being beta_s and c1_s
        int compact_stage;
        double* stage_side;              // per chain position: FP64 sum of the same-tile upstream cells solved

        // Chain schedule, round 0 of a run (tile_pass.hpp): in-tile chain layout and the rest of the tiles'
        // recorded programs (tile_layout.hpp)
        std::uint16_t const* tile_order;   // per tile and position: local cell | order_start_bit
        std::uint8_t const* tile_flags;    // per tile and position: light children
        std::uint16_t const* tile_split;
        std::uint8_t const* tile_nlevels;
        std::uint32_t const* stage_cell;   // per staging entry: local cell | upstream mask solved in the tile << 16
        std::uint32_t const* lo_off;       // per tile (+1): first entry of its cells that feed side slots
        std::uint64_t const* lo_entry;     // per such cell: side slot << 16 | local cell
        double* slot;                      // per side slot: the source's outflow (FP64)

        // A network of one catchment: that catchment and its rasters, so that kernels need not load them
        // (one_catchment set; otherwise cats, io and tile_cat hold them)
        int one_catchment;
        Catchment cat0;
        CatchmentIO<T> io0;
    };

    // Catchment index, catchment and rasters of a tile.
    template<typename T>
    __device__ inline std::uint32_t tile_catchment(SweepArgs<T> const& a, std::int64_t tile, Catchment& cat,
                                                   CatchmentIO<T>& io)
    {
        if (a.one_catchment)
        {
            cat = a.cat0;
            io = a.io0;
            return 0u;
        }
        std::uint32_t const index = as_global(a.tile_cat)[tile];
        cat = a.cats[index];
        io = a.io[index];
        return index;
    }

    template<int TW, int TH>
    struct Tile
    {
        static constexpr int n = TW * TH;
        static constexpr int hw = TW + 2;
        static constexpr int hn = (TW + 2) * (TH + 2);
        static constexpr int perimeter = 2 * TW + 2 * (TH - 2);
        static constexpr int words = n / 32;
        static_assert(n % 32 == 0, "tile must hold a multiple of 32 cells");
        static_assert(TW >= 2 && TH >= 2, "tile too small");

        // Index of a border cell on the tile perimeter, -1 for interior cells.
        __host__ __device__ static int perimeter_index(int lr, int lc)
        {
            if (lr == 0)
                return lc;
            if (lr == TH - 1)
                return TW + lc;
            if (lc == 0)
                return 2 * TW + (lr - 1);
            if (lc == TW - 1)
                return 2 * TW + (TH - 2) + (lr - 1);
            return -1;
        }
    };

    // Per-cell inputs staged in LDS.
    template<int OP, typename T>
    struct CellInput;

    template<typename T>
    struct CellInput<op_accu, T>
    {
        T x;  // external inflow, NaN when no-data or outside the domain
    };

    template<typename T>
    struct CellInput<op_threshold, T>
    {
        T x;  // NaN when either argument is no-data or outside the domain
        T p;
    };

    template<typename T>
    struct CellInput<op_fraction, T>
    {
        T x;
        T p;
    };

    template<typename T>
    struct CellInput<op_capacity, T>
    {
        T x;
        T p;
    };

    template<typename T>
    struct CellInput<op_trigger, T>
    {
        T x;
        T p;
    };

    // Kinematic wave, plain FP64 path.
    template<typename T>
    struct CellInput<op_kinematic_wave, T>
    {
        T q_cur;    // NaN when any argument is no-data or outside the domain
        T q_lat;
        T alpha;
        T beta;
        double b;   // alpha q_cur^beta + dt q_lat
        double c1;  // dt / dx
    };

    // Kinematic wave, FP32 path (24 bytes, so that a 32x32 tile's layout kernel fits four workgroups per
    // CU): b as an unevaluated pair of floats (about 48 significant bits); the lateral inflow itself is
    // recovered from b where LUE's iteration needs it (kw_q_lat).
    template<>
    struct CellInput<op_kinematic_wave, float>
    {
        float q_cur;  // NaN when any argument is no-data or outside the domain
        float alpha;
        float beta;
        float c1;     // dt / dx
        float b_hi;   // b = alpha q_cur^beta + dt q_lat = b_hi + b_lo
        float b_lo;
    };

    // The staged form of CellInput<op_kinematic_wave, float> where beta and c1 are the scalars of the call.
    struct alignas(16) CompactKwInput
    {
        float q_cur;
        float alpha;
        float b_hi;
        float b_lo;
    };

    // Store / load the staged inputs of a chain position (SweepArgs::compact_stage selects the compact form for
    // the FP32 kinematic wave).
    template<int OP, typename T>
    __device__ inline void store_staged(SweepArgs<T> const& a, std::uint32_t p, CellInput<OP, T> const& in)
    {
        if constexpr (OP == op_kinematic_wave && std::is_same_v<T, float>)
        {
            if (a.compact_stage)
            {
                static_cast<CompactKwInput*>(a.staged)[p] = CompactKwInput{in.q_cur, in.alpha, in.b_hi, in.b_lo};
                return;
            }
        }
        static_cast<CellInput<OP, T>*>(a.staged)[p] = in;
    }

    template<int OP, typename T>
    __device__ inline CellInput<OP, T> load_staged(SweepArgs<T> const& a, void const* staged, std::uint32_t p)
    {
        if constexpr (OP == op_kinematic_wave && std::is_same_v<T, float>)
        {
            if (a.compact_stage)
            {
                CompactKwInput const v = static_cast<CompactKwInput const*>(staged)[p];
                return CellInput<OP, T>{v.q_cur, v.alpha, a.beta_s, a.c1_s, v.b_hi, v.b_lo};
            }
        }
        return static_cast<CellInput<OP, T> const*>(staged)[p];
    }

    template<typename T>
    __device__ inline double kw_b(CellInput<op_kinematic_wave, T> const& c)
    {
        if constexpr (std::is_same_v<T, float>)
            return static_cast<double>(c.b_hi) + static_cast<double>(c.b_lo);
        else
            return c.b;
    }

    template<typename T>
    __device__ inline double kw_c1(CellInput<op_kinematic_wave, T> const& c)
    {
        return static_cast<double>(c.c1);
    }

    // Q_cur^beta of the FP32 path's b = alpha Q_cur^beta + dt q, in FP64 (the short pow_positive, relative
    // error below about 1e-12): b enters every root, and where LUE's own iteration stops short of the root
    // (see lue_guess_converges) its result amplifies input errors about 20-fold, so b is formed far more
    // accurately than FP32 (the hardware FP32 power loses |beta log2 Q_cur| 2^-24). pow for Q_cur = 0 and
    // non-finite arguments.
    __device__ inline double kw_b_pow(float q_cur, float beta)
    {
        if (!(q_cur > 0.0f) || !isfinite(q_cur) || !isfinite(beta))
            return static_cast<double>(powf(q_cur, beta));
        return pow_positive<true, true>(static_cast<double>(q_cur), static_cast<double>(beta));
    }

    // Lateral inflow of a cell: stored (FP64 path) or recovered from b = alpha Q_cur^beta + dt q (FP32 path,
    // where b was formed with the same power, so the difference is dt q up to FP64 rounding; exact for
    // Q_cur = 0, the one case where LUE's dryness test reads q). The iteration uses q only as dt q, so for
    // dt = 0 it is taken as 0.
    template<typename T>
    __device__ inline T kw_q_lat(CellInput<op_kinematic_wave, T> const& c, T dt)
    {
        if constexpr (std::is_same_v<T, float>)
            return dt != 0.0f ? static_cast<float>((kw_b(c) - static_cast<double>(c.alpha) *
                                                                   kw_b_pow(c.q_cur, c.beta)) /
                                                   static_cast<double>(dt))
                              : 0.0f;
        else
            return c.q_lat;
    }


    template<typename T>
    __device__ inline T value_of(T const* raster, T scalar, std::int64_t idx)
    {
        return raster ? as_global(raster)[idx] : scalar;
    }

    template<typename T>
    __device__ inline T quiet_nan()
    {
        return std::numeric_limits<T>::quiet_NaN();
    }

    // Raw input values of one cell (before the domain checks), loaded apart from their use so that the
    // loads of several cells are in flight together.
    template<int OP, typename T>
    struct RawInput
    {
        T v[OP == op_accu ? 1 : (OP == op_kinematic_wave ? 5 : 2)];
    };

    template<int OP, typename T>
    __device__ inline void load_raw(SweepArgs<T> const& a, CatchmentIO<T> const& io, std::int64_t g, RawInput<OP, T>& r)
    {
        r.v[0] = value_of(io.in1, a.in1_s, g);
        if constexpr (OP != op_accu)
            r.v[1] = value_of(io.in2, a.in2_s, g);
        if constexpr (OP == op_kinematic_wave)
        {
            r.v[2] = value_of(io.alpha, a.alpha_s, g);
            r.v[3] = value_of(io.beta, a.beta_s, g);
            r.v[4] = value_of(io.dx, a.dx_s, g);
        }
    }

    // Domain checks and derived coefficients of one cell from its raw inputs. dx_scalar: the channel length
    // is the scalar of the arguments (its c1 = dt / dx is then a.c1_s, the same quotient computed once).
    template<int OP, typename T>
    __device__ inline void stage_raw(SweepArgs<T> const& a, RawInput<OP, T> const& r, CellInput<OP, T>& c,
                                     bool dx_scalar = false)
    {
        if constexpr (OP == op_accu)
        {
            T const x = r.v[0];
            // value_policies/accu.hpp:9-19: inflow >= 0
            c.x = (x >= T{0}) ? x : quiet_nan<T>();
        }
        else if constexpr (is_split_op(OP))
        {
            T const x = r.v[0];
            T const p = r.v[1];
            // value_policies/accu_threshold.hpp:9-20, value_policies/accu_fraction.hpp:9-20,
            // value_policies/accu_capacity.hpp:15-19, value_policies/accu_trigger.hpp:15-19
            bool const ok = OP == op_fraction ? (x >= T{0} && p >= T{0} && p <= T{1}) : (x >= T{0} && p >= T{0});
            c.x = ok ? x : quiet_nan<T>();
            c.p = p;
        }
        else
        {
            T const q_cur = r.v[0];
            T const q_lat = r.v[1];
            T const alpha = r.v[2];
            T const beta = r.v[3];
            T const dx = r.v[4];
            // value_policies/kinematic_wave.hpp:9-22 plus no-data (NaN) in any argument
            bool const ok = q_cur >= T{0} && q_lat >= T{0} && dx > T{0} && !isnan(alpha) && !isnan(beta) &&
                            !isnan(a.dt);
            double const dt = static_cast<double>(a.dt);
            if constexpr (std::is_same_v<T, float>)
            {
                // b in FP64 (kw_b_pow), kept as an FP32 pair
                double const b = ok ? static_cast<double>(alpha) * kw_b_pow(q_cur, beta) +
                                          dt * static_cast<double>(q_lat)
                                    : 0.0;
                c.q_cur = ok ? q_cur : quiet_nan<T>();
                c.alpha = alpha;
                c.beta = beta;
                c.c1 = dx_scalar ? a.c1_s : a.dt / dx;  // correctly rounded FP32 division
                c.b_hi = static_cast<float>(b);
                c.b_lo = static_cast<float>(b - static_cast<double>(c.b_hi));
            }
            else
            {
                c.q_cur = ok ? q_cur : quiet_nan<T>();
                c.q_lat = q_lat;
                c.alpha = alpha;
                c.beta = beta;
                c.c1 = dx_scalar ? static_cast<double>(a.c1_s) : dt / dx;
                double const qb = q_cur > T{0} ? dpow(q_cur, beta) : pow(q_cur, beta);
                c.b = ok ? alpha * qb + dt * q_lat : 0.0;
            }
        }
    }

    template<int OP, typename T>
    __device__ inline void stage_input(SweepArgs<T> const& a, CatchmentIO<T> const& io, std::int64_t g,
                                       CellInput<OP, T>& c)
    {
        RawInput<OP, T> r;
        load_raw<OP, T>(a, io, g, r);
        stage_raw<OP, T>(a, r, c);
    }

    // Solve one cell given the FP64 sum of its upstream outflows.
    template<int OP, typename T>
    __device__ inline void solve_cell(
        SweepArgs<T> const& a, CellInput<OP, T> const& in, double upstream, bool poisoned, double& out, T& rem)
    {
        if constexpr (OP == op_accu)
        {
            // accu.hpp:102-123: out = sum(upstream) + external inflow
            out = (poisoned || isnan(in.x)) ? quiet_nan<double>() : upstream + static_cast<double>(in.x);
        }
        else if constexpr (OP == op_threshold)
        {
            // accu_threshold.hpp:124-169
            if (poisoned || isnan(in.x) || isnan(in.p))
            {
                out = quiet_nan<double>();
                rem = quiet_nan<T>();
            }
            else
            {
                double const total = upstream + static_cast<double>(in.x);
                double const thr = static_cast<double>(in.p);
                if (total > thr)
                {
                    rem = in.p;
                    out = total - thr;
                }
                else
                {
                    rem = static_cast<T>(total);
                    out = 0.0;
                }
            }
        }
        else if constexpr (OP == op_capacity)
        {
            // accu_capacity.hpp:124-166: the outflow is at most the capacity, the rest remains
            if (poisoned || isnan(in.x) || isnan(in.p))
            {
                out = quiet_nan<double>();
                rem = quiet_nan<T>();
            }
            else
            {
                double const total = upstream + static_cast<double>(in.x);
                double const cap = static_cast<double>(in.p);
                if (total > cap)
                {
                    rem = static_cast<T>(total - cap);
                    out = cap;
                }
                else
                {
                    rem = T{0};
                    out = total;
                }
            }
        }
        else if constexpr (OP == op_trigger)
        {
            // accu_trigger.hpp:124-164: below the trigger all material remains, from it on all flows
            if (poisoned || isnan(in.x) || isnan(in.p))
            {
                out = quiet_nan<double>();
                rem = quiet_nan<T>();
            }
            else
            {
                double const total = upstream + static_cast<double>(in.x);
                if (total < static_cast<double>(in.p))
                {
                    rem = static_cast<T>(total);
                    out = 0.0;
                }
                else
                {
                    rem = T{0};
                    out = total;
                }
            }
        }
        else if constexpr (OP == op_fraction)
        {
            // accu_fraction.hpp:137-162
            if (poisoned || isnan(in.x) || isnan(in.p))
            {
                out = quiet_nan<double>();
                rem = quiet_nan<T>();
            }
            else
            {
                double const total = upstream + static_cast<double>(in.x);
                double const delta = static_cast<double>(in.p) * total;
                rem = static_cast<T>(total - delta);
                out = delta;
            }
        }
        else
        {
            // kinematic_wave.hpp:416-456
            if (poisoned || isnan(in.q_cur))
            {
                out = quiet_nan<double>();
            }
            else if constexpr (std::is_same_v<T, double>)
            {
                KinematicWaveCell<double> cell{upstream, in.q_cur, in.q_lat, in.alpha, in.beta, a.dt, in.c1, in.b};
                out = new_discharge_f64(cell);
            }
            else
            {
                out = new_discharge_f32_path(upstream, in.q_cur, kw_q_lat(in, a.dt), in.alpha, in.beta, a.dt, kw_b(in),
                                             kw_c1(in));
            }
        }
    }


}  // namespace lue_amd
