// LUE flow-routing oracle: a CPU restatement of LUE's accumulating router.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP implementation in lue_amd/: it is
// loaded by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg, never by the
// product path.
//
// What it restates (reference paths relative to source/framework/algorithm/include/lue/framework/algorithm/):
//   - ldd codes and downstream offsets ............ flow_direction.hpp:7-35, detail/flow_direction.hpp:11-127
//   - partition shapes (last one absorbs the rest)  ../../../core/include/lue/framework/core/domain_decomposition.hpp:91-183
//   - inflow count (intra + cross-partition) ...... detail/inflow_count.hpp:9-352, definition/inflow_count.hpp:527-577
//   - three-phase router ........................... definition/accumulating_router.hpp:525-599 (walk),
//                                                    :866-882 (row-major ridge scan), :987-1010 (inter-partition)
//   - cell functors ................................ definition/accu.hpp:62-170, definition/accu_threshold.hpp:124-230,
//                                                    definition/accu_fraction.hpp:126-190, definition/kinematic_wave.hpp:416-515,
//                                                    definition/accu_capacity.hpp:124-166, definition/accu_trigger.hpp:124-164,
//                                                    definition/accu_info.hpp:53-145 (cell classes, accu_info.hpp:27-63)
//   - value-policy domains ......................... value_policies/{accu,accu_threshold,accu_fraction,kinematic_wave,
//                                                    accu_capacity,accu_trigger}.hpp
//   - kinematic-wave Newton solve .................. definition/kinematic_wave.hpp:13-254
//   - D8 flow direction ............................ definition/d8_flow_direction.hpp:12-143,
//                                                    value_policies/d8_flow_direction.hpp:31-41 (halo value)
//   - downstream_distance .......................... definition/downstream_distance.hpp:13-75, with the value
//                                                    policies of python/source/algorithm/routing_operation/
//                                                    downstream_distance.cpp:14-23 (no-data ldd -> NaN) or without
//   - upstream / downstream ........................ definition/upstream.hpp:36-150, definition/downstream.hpp:36-155,
//                                                    flow_accumulation.hpp:47-84 (policies: halo ldd drains away,
//                                                    halo material 0), policy/default_value_policies.hpp:19-29 and
//                                                    policy/detect_no_data_by_value.hpp:25-50 (no-data values)
//
// Arithmetic order. Inside one partition the order of floating point operations is LUE's: cells are
// scanned row-major, a walk starts at every ridge cell (inflow count 0), upstream outflow is added to a
// downstream cell in walk-arrival order and the external inflow is added when the cell is entered.
// Across partitions LUE applies incoming fluxes in channel-arrival order, which is not deterministic.
// Here the inter-partition phase runs in bulk-synchronous rounds: in each round every partition applies
// the fluxes it received in the previous round, ordered by sending partition and send order. The
// results therefore equal LUE's up to the floating point order at partition-input cells.
//
// Newton solve. kinematic_wave calls boost::math::tools::newton_raphson_iterate (Boost >= 1.70, not
// pinned by LUE: environment/cmake/LueConfiguration.cmake:515). Boost is not vendored in the
// reference. The iteration below restates Boost's documented contract (roots_deriv.html): Newton
// steps from the guess, kept inside a shrinking bracket [min, max] (a step that leaves the bracket
// is replaced by halving the distance to the violated bound), a step that fails to at least halve
// relative to the step before last is replaced by a bounded jump, and iteration stops when
// |step| <= |x| * 2^(1-digits) or after max_iter evaluations. digits = floor(0.6 * digits(T)) and
// max_iter = 10 as LUE passes them (definition/kinematic_wave.hpp:195-213).
//
// Invalid networks. LUE asserts that a network is valid (no cycles, no flow into a no-data cell).
// Here flow into a no-data cell leaves the network like flow off the array edge, and a cycle is
// reported as error 5 (cells on or below a cycle are never resolved).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <type_traits>
#include <vector>

namespace {

    using Index = std::int64_t;

    constexpr std::uint8_t no_data_ldd = 255;

    enum : int { ok = 0, err_invalid_argument = 1, err_invalid_network = 5 };

    // Keypad encoding: 7 8 9 / 4 5 6 / 1 2 3, 5 = sink; any other code behaves as a sink.
    inline void ldd_offset(std::uint8_t const code, int& d0, int& d1)
    {
        switch (code)
        {
            case 7: d0 = -1; d1 = -1; return;
            case 8: d0 = -1; d1 = 0; return;
            case 9: d0 = -1; d1 = 1; return;
            case 4: d0 = 0; d1 = -1; return;
            case 6: d0 = 0; d1 = 1; return;
            case 1: d0 = 1; d1 = -1; return;
            case 2: d0 = 1; d1 = 0; return;
            case 3: d0 = 1; d1 = 1; return;
            default: d0 = 0; d1 = 0; return;
        }
    }

    // Partition grid per domain_decomposition.hpp:91-183: #partitions = floor(extent / partition
    // extent) along each dimension, the last partition absorbs the remainder.
    struct Grid
    {
        Index rows, cols, prows, pcols, npr, npc;

        auto prow_of(Index r) const -> Index { return std::min(r / prows, npr - 1); }
        auto pcol_of(Index c) const -> Index { return std::min(c / pcols, npc - 1); }
        auto part_of(Index r, Index c) const -> Index { return prow_of(r) * npc + pcol_of(c); }
        auto r_begin(Index pr) const -> Index { return pr * prows; }
        auto r_end(Index pr) const -> Index { return pr == npr - 1 ? rows : (pr + 1) * prows; }
        auto c_begin(Index pc) const -> Index { return pc * pcols; }
        auto c_end(Index pc) const -> Index { return pc == npc - 1 ? cols : (pc + 1) * pcols; }
    };

    auto make_grid(Index rows, Index cols, Index prows, Index pcols, Grid& g) -> bool
    {
        if (rows <= 0 || cols <= 0 || prows <= 0 || pcols <= 0 || prows > rows || pcols > cols)
        {
            return false;
        }
        g = Grid{rows, cols, prows, pcols, rows / prows, cols / pcols};
        return true;
    }

    template<typename F>
    void parallel_for(Index n, int nthreads, F&& f)
    {
        if (nthreads <= 1 || n <= 1)
        {
            for (Index i = 0; i < n; ++i)
            {
                f(i);
            }
            return;
        }
        std::atomic<Index> next{0};
        std::vector<std::thread> pool;
        int const nt = static_cast<int>(std::min<Index>(nthreads, n));
        pool.reserve(nt);
        for (int t = 0; t < nt; ++t)
        {
            pool.emplace_back(
                [&]()
                {
                    for (Index i = next++; i < n; i = next++)
                    {
                        f(i);
                    }
                });
        }
        for (auto& th : pool)
        {
            th.join();
        }
    }

    // Global inflow count: number of in-array, non-no-data neighbours that drain into a cell. This is
    // what the intra-partition count plus the cross-partition input-cell increments add up to
    // (detail/inflow_count.hpp:24-47 and definition/inflow_count.hpp:527-577). Flow into a no-data
    // cell is not counted.
    void inflow_counts(Grid const& g, std::uint8_t const* ldd, std::vector<std::int32_t>& count, int nthreads)
    {
        count.assign(static_cast<std::size_t>(g.rows * g.cols), 0);
        // Row-parallel gather: a cell counts the neighbours pointing at it.
        parallel_for(
            g.rows,
            nthreads,
            [&](Index r)
            {
                for (Index c = 0; c < g.cols; ++c)
                {
                    Index const idx = r * g.cols + c;
                    if (ldd[idx] == no_data_ldd)
                    {
                        continue;
                    }
                    std::int32_t n = 0;
                    for (int dr = -1; dr <= 1; ++dr)
                    {
                        for (int dc = -1; dc <= 1; ++dc)
                        {
                            if (dr == 0 && dc == 0)
                            {
                                continue;
                            }
                            Index const r2 = r + dr, c2 = c + dc;
                            if (r2 < 0 || r2 >= g.rows || c2 < 0 || c2 >= g.cols)
                            {
                                continue;
                            }
                            std::uint8_t const code = ldd[r2 * g.cols + c2];
                            if (code == no_data_ldd)
                            {
                                continue;
                            }
                            int d0, d1;
                            ldd_offset(code, d0, d1);
                            if (d0 == -dr && d1 == -dc && !(d0 == 0 && d1 == 0))
                            {
                                ++n;
                            }
                        }
                    }
                    count[idx] = n;
                }
            });
    }

    // ---------------------------------------------------------------------------------------------
    // Kinematic wave (definition/kinematic_wave.hpp:13-254)

    template<typename T>
    struct KinematicWaveEquation
    {
        T upstream, current, lateral, alpha, beta, alpha_beta, dt, dt_over_dx, known;

        KinematicWaveEquation(T q_up, T q_cur, T q_lat, T a, T b, T t, T dx):
            upstream{q_up},
            current{q_cur},
            lateral{q_lat},
            alpha{a},
            beta{b},
            alpha_beta{a * b},
            dt{t}
        {
            dt_over_dx = dt / dx;
            known = (dt_over_dx * upstream) + (alpha * std::pow(current, beta)) + (dt * lateral);
        }

        // definition/kinematic_wave.hpp:61-91
        auto guess() const -> T
        {
            T const min_discharge = std::numeric_limits<T>::min();
            T q = min_discharge;
            if ((current + upstream != T{0}) || beta >= T{1})
            {
                T const a_b_pq = alpha_beta * std::pow((current + upstream) / T{2}, beta - T{1});
                q = (dt_over_dx * upstream + a_b_pq * current + dt * lateral) / (dt_over_dx + a_b_pq);
                q = std::max<T>(q, min_discharge);
            }
            return q;
        }

        auto f(T q) const -> T { return (dt_over_dx * q) + (alpha * std::pow(q, beta)) - known; }
        auto df(T q) const -> T { return dt_over_dx + (alpha_beta * std::pow(q, beta - T{1})); }
    };

    // Restatement of the documented newton_raphson_iterate contract (see file header).
    template<typename T, typename Eq>
    auto newton_raphson(Eq const& eq, T guess, T lo, T hi, int digits, int max_iter) -> T
    {
        T const factor = std::ldexp(T{1}, 1 - digits);
        T x = guess;
        T step = std::numeric_limits<T>::max();
        T step_prev = step;
        T step_prev2 = step;
        int left = max_iter;

        while (true)
        {
            step_prev2 = step_prev;
            step_prev = step;
            T const f0 = eq.f(x);
            T const f1 = eq.df(x);
            --left;
            if (f0 == T{0})
            {
                break;
            }
            if (f1 == T{0})
            {
                // A zero derivative does not occur for this equation (df > 0 for q > 0); fall back to
                // a bisection step towards the bracket.
                step = (f0 > T{0}) ? (x - lo) / 2 : (x - hi) / 2;
            }
            else
            {
                step = f0 / f1;
            }
            if (std::fabs(step * 2) > std::fabs(step_prev2))
            {
                // The last two steps did not converge: jump instead, bounded by the bracket and by
                // the size of the current iterate.
                T const shift = (step > 0) ? (x - lo) / 2 : (x - hi) / 2;
                if ((x != 0) && (std::fabs(shift) > std::fabs(x)))
                {
                    step = std::copysign(std::fabs(x) * T(1.1f), step);
                }
                else
                {
                    step = shift;
                }
                step_prev = 3 * step;
                step_prev2 = 3 * step;
            }
            T const previous = x;
            x -= step;
            if (x <= lo)
            {
                step = T(0.5f) * (previous - lo);
                x = previous - step;
                if (x == lo || x == hi)
                {
                    break;
                }
            }
            else if (x >= hi)
            {
                step = T(0.5f) * (previous - hi);
                x = previous - step;
                if (x == lo || x == hi)
                {
                    break;
                }
            }
            if (step > 0)
            {
                hi = previous;
            }
            else
            {
                lo = previous;
            }
            if (left == 0 || !(std::fabs(x * factor) < std::fabs(step)))
            {
                break;
            }
        }
        return x;
    }

    // definition/kinematic_wave.hpp:145-254
    template<typename T>
    auto iterate_to_new_discharge(T q_up, T q_cur, T q_lat, T alpha, T beta, T dt, T dx) -> T
    {
        T q_new{0};
        if (q_up + q_cur > 0 || q_lat > 0)
        {
            T const inflow = q_lat >= 0 ? q_lat : T{0};
            KinematicWaveEquation<T> const eq{q_up, q_cur, inflow, alpha, beta, dt, dx};
            int const digits = static_cast<int>(std::numeric_limits<T>::digits * 0.6);
            q_new = newton_raphson<T>(eq, eq.guess(), T{0}, std::numeric_limits<T>::max(), digits, 10);
        }
        if (q_lat < T{0})
        {
            T const extraction = std::min(dx * std::fabs(q_lat), q_new);
            q_new -= extraction;
        }
        return q_new;
    }

    // ---------------------------------------------------------------------------------------------
    // Cell functors. Each holds the inputs (raster or broadcast scalar, accumulating_router.hpp:727-786)
    // and the outputs; the router calls them in LUE's order.

    template<typename T>
    struct Input
    {
        T const* raster;
        T scalar;
        auto operator()(Index idx) const -> T { return raster ? raster[idx] : scalar; }
    };

    template<typename T>
    inline auto is_nan(T v) -> bool
    {
        return std::isnan(v);
    }

    template<typename T>
    inline auto nan_value() -> T
    {
        return std::numeric_limits<T>::quiet_NaN();
    }

    enum class Kind { accu, threshold, fraction, kinematic_wave, capacity, trigger };

    constexpr bool two_outputs(Kind kind)
    {
        return kind == Kind::threshold || kind == Kind::fraction || kind == Kind::capacity || kind == Kind::trigger;
    }

    template<typename T, Kind kind>
    struct Cells
    {
        Input<T> in1;  // material | current outflow
        Input<T> in2;  // threshold | fraction | capacity | trigger | lateral inflow
        Input<T> alpha, beta, dx;
        T dt;
        T* out;  // outflow; holds the summed upstream outflow until the cell is entered
        T* rem;  // remainder (threshold, fraction, capacity, trigger)

        void mark_no_data(Index i)
        {
            out[i] = nan_value<T>();
            if constexpr (two_outputs(kind))
            {
                rem[i] = nan_value<T>();
            }
        }

        // Router events with empty bodies in the material functors (accu.hpp:62-72,96-99,156-170)
        void enter_intra_partition_stream(Index) {}
        void leave_intra_partition_stream(Index) {}
        void leave_inter_partition_stream(Index) {}
        void stop_at_sink_cell(Index) {}
        void stop_at_confluence_cell(Index) {}
        void stop_at_partition_output_cell(Index) {}

        auto outflow(Index i) const -> T { return out[i]; }

        void leave_cell(Index from, Index to) { add_upstream(out[from], to); }

        void enter_inter_partition_stream(T upstream_outflow, Index to) { add_upstream(upstream_outflow, to); }

        // leave_cell / enter_inter_partition_stream (accu.hpp:74-93,126-153)
        void add_upstream(T upstream_outflow, Index to)
        {
            if (!is_nan(out[to]))
            {
                if (is_nan(upstream_outflow))
                {
                    mark_no_data(to);
                }
                else
                {
                    out[to] += upstream_outflow;
                }
            }
        }

        void enter_cell(Index i)
        {
            if (is_nan(out[i]))
            {
                return;
            }
            if constexpr (kind == Kind::accu)
            {
                // accu.hpp:102-123; domain inflow >= 0 (value_policies/accu.hpp:9-19)
                T const ext = in1(i);
                if (is_nan(ext) || !(ext >= 0))
                {
                    mark_no_data(i);
                }
                else
                {
                    out[i] += ext;
                }
            }
            else if constexpr (kind == Kind::threshold)
            {
                // accu_threshold.hpp:124-169; domain inflow >= 0 and threshold >= 0
                T const ext = in1(i);
                T const thr = in2(i);
                if (is_nan(ext) || is_nan(thr) || !(ext >= 0 && thr >= 0))
                {
                    mark_no_data(i);
                }
                else
                {
                    out[i] += ext;
                    if (out[i] > thr)
                    {
                        rem[i] = thr;
                        out[i] -= thr;
                    }
                    else
                    {
                        rem[i] = out[i];
                        out[i] = 0;
                    }
                }
            }
            else if constexpr (kind == Kind::fraction)
            {
                // accu_fraction.hpp:137-162; domain inflow >= 0 and 0 <= fraction <= 1
                T const ext = in1(i);
                T const frac = in2(i);
                if (is_nan(ext) || is_nan(frac) || !(ext >= 0 && frac >= 0 && frac <= 1))
                {
                    mark_no_data(i);
                }
                else
                {
                    out[i] += ext;
                    T const delta = frac * out[i];
                    rem[i] = out[i] - delta;
                    out[i] = delta;
                }
            }
            else if constexpr (kind == Kind::capacity)
            {
                // accu_capacity.hpp:124-166; domain inflow >= 0 and capacity >= 0
                // (value_policies/accu_capacity.hpp:9-21)
                T const ext = in1(i);
                T const cap = in2(i);
                if (is_nan(ext) || is_nan(cap) || !(ext >= 0 && cap >= 0))
                {
                    mark_no_data(i);
                }
                else
                {
                    out[i] += ext;
                    rem[i] = 0;
                    if (out[i] > cap)
                    {
                        rem[i] = out[i] - cap;
                        out[i] = cap;
                    }
                }
            }
            else if constexpr (kind == Kind::trigger)
            {
                // accu_trigger.hpp:124-164; domain inflow >= 0 and trigger >= 0
                // (value_policies/accu_trigger.hpp)
                T const ext = in1(i);
                T const trg = in2(i);
                if (is_nan(ext) || is_nan(trg) || !(ext >= 0 && trg >= 0))
                {
                    mark_no_data(i);
                }
                else
                {
                    out[i] += ext;
                    rem[i] = 0;
                    if (out[i] < trg)
                    {
                        rem[i] = out[i];
                        out[i] = 0;
                    }
                }
            }
            else
            {
                // kinematic_wave.hpp:416-456; domain current >= 0, lateral >= 0, channel length > 0
                T const q_cur = in1(i);
                T const q_lat = in2(i);
                T const a = alpha(i);
                T const b = beta(i);
                T const l = dx(i);
                if (is_nan(q_cur) || is_nan(q_lat) || is_nan(a) || is_nan(b) || is_nan(dt) || is_nan(l) ||
                    !(q_cur >= 0 && q_lat >= 0 && l > 0))
                {
                    mark_no_data(i);
                }
                else
                {
                    out[i] = iterate_to_new_discharge<T>(out[i], q_cur, q_lat, a, b, dt, l);
                }
            }
        }
    };

    // Cell classes of accu_info (accu_info.hpp:27-63); the no-data value is the uint8 maximum.
    constexpr std::uint8_t ridge_cell = 11;
    constexpr std::uint8_t intra_partition_stream_cell = 12;
    constexpr std::uint8_t partition_output_cell = 13;
    constexpr std::uint8_t partition_input_cell = 14;
    constexpr std::uint8_t partition_input_output_cell = 15;
    constexpr std::uint8_t inter_partition_stream_cell = 16;
    constexpr std::uint8_t sink_cell = 17;
    constexpr std::uint8_t confluence_cell = 18;
    constexpr std::uint8_t no_data_class = 255;

    // accu_info's functor (definition/accu_info.hpp:53-145): each router event writes a class, leave_cell
    // passes the class of the cell left to the cell entered.
    struct CellInfo
    {
        std::uint8_t* out;
        std::uint8_t* rem;  // unused (one output)

        void mark_no_data(Index i) { out[i] = no_data_class; }
        void enter_intra_partition_stream(Index i) { out[i] = intra_partition_stream_cell; }
        void leave_intra_partition_stream(Index i) { out[i] = ridge_cell; }
        void enter_inter_partition_stream(std::uint8_t, Index i) { out[i] = inter_partition_stream_cell; }
        void leave_inter_partition_stream(Index i)
        {
            out[i] = out[i] != partition_output_cell ? partition_input_cell : partition_input_output_cell;
        }
        void enter_cell(Index) {}
        void leave_cell(Index from, Index to)
        {
            if (out[to] != no_data_class)
            {
                out[to] = out[from] == no_data_class ? no_data_class : out[from];
            }
        }
        void stop_at_sink_cell(Index i) { out[i] = sink_cell; }
        void stop_at_confluence_cell(Index i) { out[i] = confluence_cell; }
        void stop_at_partition_output_cell(Index i) { out[i] = partition_output_cell; }
        auto outflow(Index i) const -> std::uint8_t { return out[i]; }
    };

    struct Message
    {
        Index dest;
        double value_f64;  // wide enough for either element type
    };

    // The router: phases 2 and 3 of accumulating_router.hpp, for any functor.
    template<typename T, typename C>
    auto route(Grid const& g, std::uint8_t const* ldd, C& cells, int nthreads) -> int
    {
        Index const n = g.rows * g.cols;
        std::vector<std::int32_t> count0;
        inflow_counts(g, ldd, count0, nthreads);
        std::vector<std::int32_t> count{count0};  // deep copy (accumulating_router.hpp:844)

        // Zero-initialised results (accumulating_router.hpp:1486-1491)
        std::fill(cells.out, cells.out + n, T{0});
        if (cells.rem)
        {
            std::fill(cells.rem, cells.rem + n, T{0});
        }

        Index const nparts = g.npr * g.npc;
        std::vector<std::vector<Message>> outbox(static_cast<std::size_t>(nparts));

        // Walk from a ready cell inside partition p (accumulating_router.hpp:525-599).
        auto walk = [&](Index p, Index start, std::vector<Message>& sent)
        {
            Index const pr = p / g.npc, pc = p % g.npc;
            Index const rb = g.r_begin(pr), re = g.r_end(pr), cb = g.c_begin(pc), ce = g.c_end(pc);
            Index idx = start;
            while (true)
            {
                cells.enter_cell(idx);
                Index const r = idx / g.cols, c = idx % g.cols;
                int d0, d1;
                ldd_offset(ldd[idx], d0, d1);
                if (d0 == 0 && d1 == 0)
                {
                    cells.stop_at_sink_cell(idx);
                    break;
                }
                Index const r2 = r + d0, c2 = c + d1;
                bool const in_array = r2 >= 0 && r2 < g.rows && c2 >= 0 && c2 < g.cols;
                bool const in_part = r2 >= rb && r2 < re && c2 >= cb && c2 < ce;
                if (!in_part)
                {
                    // Partition output cell (also at the array edge): send to the neighbour, or drop at
                    // the array edge (accumulating_router.hpp:467-510).
                    cells.stop_at_partition_output_cell(idx);
                    if (in_array && ldd[r2 * g.cols + c2] != no_data_ldd)
                    {
                        sent.push_back(Message{r2 * g.cols + c2, static_cast<double>(cells.outflow(idx))});
                    }
                    break;
                }
                Index const to = r2 * g.cols + c2;
                if (ldd[to] == no_data_ldd)
                {
                    break;  // flow into no-data leaves the network
                }
                cells.leave_cell(idx, to);
                if (--count[to] > 0)
                {
                    cells.stop_at_confluence_cell(to);
                    break;
                }
                idx = to;
            }
        };

        // Phase 2: intra-partition streams, partitions in parallel (accumulating_router.hpp:866-882)
        parallel_for(
            nparts,
            nthreads,
            [&](Index p)
            {
                Index const pr = p / g.npc, pc = p % g.npc;
                for (Index r = g.r_begin(pr); r < g.r_end(pr); ++r)
                {
                    for (Index c = g.c_begin(pc); c < g.c_end(pc); ++c)
                    {
                        Index const idx = r * g.cols + c;
                        if (ldd[idx] == no_data_ldd)
                        {
                            cells.mark_no_data(idx);
                        }
                        else if (count0[idx] == 0)
                        {
                            cells.enter_intra_partition_stream(idx);
                            walk(p, idx, outbox[static_cast<std::size_t>(p)]);
                            cells.leave_intra_partition_stream(idx);
                        }
                    }
                }
            });

        // Phase 3: inter-partition streams in bulk-synchronous rounds (accumulating_router.hpp:987-1010)
        std::vector<std::vector<Message>> inbox(static_cast<std::size_t>(nparts));
        while (true)
        {
            bool any = false;
            for (auto& in : inbox)
            {
                in.clear();
            }
            for (Index p = 0; p < nparts; ++p)
            {
                for (Message const& m : outbox[static_cast<std::size_t>(p)])
                {
                    Index const dp = g.part_of(m.dest / g.cols, m.dest % g.cols);
                    inbox[static_cast<std::size_t>(dp)].push_back(m);
                    any = true;
                }
                outbox[static_cast<std::size_t>(p)].clear();
            }
            if (!any)
            {
                break;
            }
            parallel_for(
                nparts,
                nthreads,
                [&](Index p)
                {
                    for (Message const& m : inbox[static_cast<std::size_t>(p)])
                    {
                        cells.enter_inter_partition_stream(static_cast<T>(m.value_f64), m.dest);
                        if (--count[m.dest] == 0)
                        {
                            walk(p, m.dest, outbox[static_cast<std::size_t>(p)]);
                            cells.leave_inter_partition_stream(m.dest);
                        }
                    }
                });
        }

        // Cells on or below a cycle never become ready.
        for (Index i = 0; i < n; ++i)
        {
            if (ldd[i] != no_data_ldd && count[i] > 0)
            {
                return err_invalid_network;
            }
        }
        return ok;
    }

    template<typename T>
    auto input(T const* raster, T scalar) -> Input<T>
    {
        return Input<T>{raster, scalar};
    }

    // D8 flow direction of one cell from its 3x3 window: definition/d8_flow_direction.hpp:12-143. No-data
    // (NaN) centre -> no-data ldd; else sink, then the neighbours in the order N, W, E, S, NW, NE, SW, SE,
    // each taken when its raw drop (centre - neighbour, no distance weighting) exceeds the best so far
    // (initially 0); no-data neighbours are skipped. Outside the array the window holds
    // numeric_limits<T>::min() (value_policies/d8_flow_direction.hpp:31-41: FillHaloWithConstantValue).
    template<typename T>
    void d8_flow_direction(T const* elevation, Index rows, Index cols, std::uint8_t* out)
    {
        T const halo = std::numeric_limits<T>::min();
        auto at = [&](Index r, Index c) -> T
        { return (r < 0 || r >= rows || c < 0 || c >= cols) ? halo : elevation[r * cols + c]; };
        // (row offset, column offset, ldd code) in LUE's order
        static constexpr int order[8][3] = {{-1, 0, 8}, {0, -1, 4}, {0, 1, 6}, {1, 0, 2},
                                            {-1, -1, 7}, {-1, 1, 9}, {1, -1, 1}, {1, 1, 3}};
        for (Index r = 0; r < rows; ++r)
        {
            for (Index c = 0; c < cols; ++c)
            {
                T const centre = elevation[r * cols + c];
                if (is_nan(centre))
                {
                    out[r * cols + c] = no_data_ldd;
                    continue;
                }
                std::uint8_t direction = 5;
                T drop{0};
                for (auto const& o : order)
                {
                    T const neighbour = at(r + o[0], c + o[1]);
                    if (is_nan(neighbour))
                        continue;
                    T const other_drop = centre - neighbour;
                    if (other_drop > drop)
                    {
                        drop = other_drop;
                        direction = static_cast<std::uint8_t>(o[2]);
                    }
                }
                out[r * cols + c] = direction;
            }
        }
    }

    // Material no-data under the default value policies: NaN for floating point, numeric_limits<T>::min()
    // for signed and ::max() for unsigned integers (policy/default_value_policies.hpp:19-29,
    // policy/detect_no_data_by_value.hpp:25-50).
    template<typename T>
    constexpr auto material_no_data() -> T
    {
        if constexpr (std::is_floating_point_v<T>)
            return std::numeric_limits<T>::quiet_NaN();
        else if constexpr (std::is_signed_v<T>)
            return std::numeric_limits<T>::min();
        else
            return std::numeric_limits<T>::max();
    }

    template<typename T>
    auto material_is_no_data(T v) -> bool
    {
        if constexpr (std::is_floating_point_v<T>)
            return is_nan(v);
        else
            return v == material_no_data<T>();
    }

    // a + b in T; integers wrap (two's complement), as the reference's `material += ...` does for the
    // narrow types and its hardware does for the wide ones
    template<typename T>
    auto material_add(T a, T b) -> T
    {
        if constexpr (std::is_floating_point_v<T>)
            return a + b;
        else
        {
            using U = std::make_unsigned_t<T>;
            return static_cast<T>(static_cast<U>(static_cast<U>(a) + static_cast<U>(b)));
        }
    }

    // upstream: definition/upstream.hpp:36-150. A no-data ldd or material in the cell gives no-data; else the
    // sum, starting from 0, of the material of the neighbours that drain into the cell, taken in the window's
    // row-major order (NW, N, NE, W, E, SW, S, SE neighbour); a no-data material among them gives no-data.
    // Cells outside the array never drain into it (FlowDirectionHalo: the halo points away).
    template<typename T>
    void upstream(std::uint8_t const* ldd, T const* material, Index rows, Index cols, T* out)
    {
        // (row offset, column offset, code that drains into the centre) in window order
        static constexpr int window[8][3] = {{-1, -1, 3}, {-1, 0, 2}, {-1, 1, 1}, {0, -1, 6},
                                             {0, 1, 4},   {1, -1, 9}, {1, 0, 8},  {1, 1, 7}};
        for (Index r = 0; r < rows; ++r)
        {
            for (Index c = 0; c < cols; ++c)
            {
                Index const i = r * cols + c;
                if (ldd[i] == no_data_ldd || material_is_no_data(material[i]))
                {
                    out[i] = material_no_data<T>();
                    continue;
                }
                T sum{0};
                for (auto const& w : window)
                {
                    Index const rr = r + w[0];
                    Index const cc = c + w[1];
                    if (rr < 0 || rr >= rows || cc < 0 || cc >= cols)
                        continue;
                    Index const j = rr * cols + cc;
                    if (ldd[j] != w[2])
                        continue;
                    if (material_is_no_data(material[j]))
                    {
                        sum = material_no_data<T>();
                        break;
                    }
                    sum = material_add(sum, material[j]);
                }
                out[i] = sum;
            }
        }
    }

    // downstream: definition/downstream.hpp:36-155. A no-data ldd or material in the cell gives no-data; else
    // the material of the downstream neighbour (no-data if that is no-data; 0 where the ldd leaves the
    // array: FillHaloWithConstantValue{0}), the cell's own material for a sink (or any code that is no
    // direction).
    template<typename T>
    void downstream(std::uint8_t const* ldd, T const* material, Index rows, Index cols, T* out)
    {
        for (Index r = 0; r < rows; ++r)
        {
            for (Index c = 0; c < cols; ++c)
            {
                Index const i = r * cols + c;
                std::uint8_t const code = ldd[i];
                if (code == no_data_ldd || material_is_no_data(material[i]))
                {
                    out[i] = material_no_data<T>();
                    continue;
                }
                if (code < 1 || code > 9 || code == 5)
                {
                    out[i] = material[i];
                    continue;
                }
                int const dr = code <= 3 ? 1 : (code >= 7 ? -1 : 0);
                int const dc = (code % 3 == 1) ? -1 : (code % 3 == 0 ? 1 : 0);
                Index const rr = r + dr;
                Index const cc = c + dc;
                if (rr < 0 || rr >= rows || cc < 0 || cc >= cols)
                {
                    out[i] = T{0};
                    continue;
                }
                T const m = material[rr * cols + cc];
                out[i] = material_is_no_data(m) ? material_no_data<T>() : m;
            }
        }
    }

    // downstream_distance: definition/downstream_distance.hpp:13-75. cell_size for N, W, E, S;
    // sqrt(2) * cell_size (in T) for NW, NE, SW, SE; 0 otherwise (pit, unknown code). Under value policies a
    // no-data ldd gives NaN; under the default policies it is an unknown code (0).
    template<typename T>
    void downstream_distance(std::uint8_t const* ldd, Index n, T cell_size, bool mark_no_data, T* out)
    {
        T const diagonal = std::sqrt(T{2}) * cell_size;
        for (Index i = 0; i < n; ++i)
        {
            std::uint8_t const code = ldd[i];
            if (mark_no_data && code == no_data_ldd)
                out[i] = std::numeric_limits<T>::quiet_NaN();
            else if (code == 7 || code == 9 || code == 1 || code == 3)
                out[i] = diagonal;
            else if (code == 8 || code == 4 || code == 6 || code == 2)
                out[i] = cell_size;
            else
                out[i] = T{0};
        }
    }

}  // namespace


extern "C" {

// Number of upstream neighbours per cell (uint8), no-data (255) where the ldd is no-data.
// ALG/definition/inflow_count.hpp:665-726 with DefaultValuePolicies<uint8, uint8>.
int oracle_inflow_count(
    std::uint8_t const* ldd,
    Index rows,
    Index cols,
    Index prows,
    Index pcols,
    std::uint8_t* out,
    int nthreads)
{
    Grid g;
    if (!make_grid(rows, cols, prows, pcols, g) || !ldd || !out)
    {
        return err_invalid_argument;
    }
    std::vector<std::int32_t> count;
    inflow_counts(g, ldd, count, nthreads);
    for (Index i = 0; i < rows * cols; ++i)
    {
        out[i] = ldd[i] == no_data_ldd ? std::uint8_t{255} : static_cast<std::uint8_t>(count[i]);
    }
    return ok;
}

#define ORACLE_ROUTING_OPS(T, SUFFIX)                                                                        \
    int oracle_accu_##SUFFIX(                                                                                \
        std::uint8_t const* ldd, Index rows, Index cols, Index prows, Index pcols,                           \
        T const* inflow, T inflow_scalar, T* outflow, int nthreads)                                          \
    {                                                                                                        \
        Grid g;                                                                                              \
        if (!make_grid(rows, cols, prows, pcols, g) || !ldd || !outflow)                                     \
            return err_invalid_argument;                                                                     \
        Cells<T, Kind::accu> cells{input(inflow, inflow_scalar), input<T>(nullptr, 0), input<T>(nullptr, 0), \
                                   input<T>(nullptr, 0), input<T>(nullptr, 0), T{0}, outflow, nullptr};     \
        return route<T>(g, ldd, cells, nthreads);                                                            \
    }                                                                                                        \
    int oracle_accu_threshold_##SUFFIX(                                                                      \
        std::uint8_t const* ldd, Index rows, Index cols, Index prows, Index pcols,                           \
        T const* inflow, T inflow_scalar, T const* threshold, T threshold_scalar, T* outflow, T* remainder,   \
        int nthreads)                                                                                        \
    {                                                                                                        \
        Grid g;                                                                                              \
        if (!make_grid(rows, cols, prows, pcols, g) || !ldd || !outflow || !remainder)                       \
            return err_invalid_argument;                                                                     \
        Cells<T, Kind::threshold> cells{input(inflow, inflow_scalar), input(threshold, threshold_scalar),    \
                                        input<T>(nullptr, 0), input<T>(nullptr, 0), input<T>(nullptr, 0),    \
                                        T{0}, outflow, remainder};                                           \
        return route<T>(g, ldd, cells, nthreads);                                                            \
    }                                                                                                        \
    int oracle_accu_fraction_##SUFFIX(                                                                       \
        std::uint8_t const* ldd, Index rows, Index cols, Index prows, Index pcols,                           \
        T const* inflow, T inflow_scalar, T const* fraction, T fraction_scalar, T* outflow, T* remainder,     \
        int nthreads)                                                                                        \
    {                                                                                                        \
        Grid g;                                                                                              \
        if (!make_grid(rows, cols, prows, pcols, g) || !ldd || !outflow || !remainder)                       \
            return err_invalid_argument;                                                                     \
        Cells<T, Kind::fraction> cells{input(inflow, inflow_scalar), input(fraction, fraction_scalar),       \
                                       input<T>(nullptr, 0), input<T>(nullptr, 0), input<T>(nullptr, 0),     \
                                       T{0}, outflow, remainder};                                            \
        return route<T>(g, ldd, cells, nthreads);                                                            \
    }                                                                                                        \
    int oracle_accu_capacity_##SUFFIX(                                                                       \
        std::uint8_t const* ldd, Index rows, Index cols, Index prows, Index pcols,                           \
        T const* inflow, T inflow_scalar, T const* capacity, T capacity_scalar, T* outflow, T* remainder,     \
        int nthreads)                                                                                        \
    {                                                                                                        \
        Grid g;                                                                                              \
        if (!make_grid(rows, cols, prows, pcols, g) || !ldd || !outflow || !remainder)                       \
            return err_invalid_argument;                                                                     \
        Cells<T, Kind::capacity> cells{input(inflow, inflow_scalar), input(capacity, capacity_scalar),       \
                                       input<T>(nullptr, 0), input<T>(nullptr, 0), input<T>(nullptr, 0),     \
                                       T{0}, outflow, remainder};                                            \
        return route<T>(g, ldd, cells, nthreads);                                                            \
    }                                                                                                        \
    int oracle_accu_trigger_##SUFFIX(                                                                        \
        std::uint8_t const* ldd, Index rows, Index cols, Index prows, Index pcols,                           \
        T const* inflow, T inflow_scalar, T const* trigger, T trigger_scalar, T* outflow, T* remainder,       \
        int nthreads)                                                                                        \
    {                                                                                                        \
        Grid g;                                                                                              \
        if (!make_grid(rows, cols, prows, pcols, g) || !ldd || !outflow || !remainder)                       \
            return err_invalid_argument;                                                                     \
        Cells<T, Kind::trigger> cells{input(inflow, inflow_scalar), input(trigger, trigger_scalar),          \
                                      input<T>(nullptr, 0), input<T>(nullptr, 0), input<T>(nullptr, 0),      \
                                      T{0}, outflow, remainder};                                             \
        return route<T>(g, ldd, cells, nthreads);                                                            \
    }                                                                                                        \
    int oracle_kinematic_wave_##SUFFIX(                                                                      \
        std::uint8_t const* ldd, Index rows, Index cols, Index prows, Index pcols,                           \
        T const* current_outflow, T const* inflow, T const* alpha, T alpha_scalar, T const* beta,            \
        T beta_scalar, T time_step_duration, T const* channel_length, T channel_length_scalar, T* outflow,    \
        int nthreads)                                                                                        \
    {                                                                                                        \
        Grid g;                                                                                              \
        if (!make_grid(rows, cols, prows, pcols, g) || !ldd || !outflow || !current_outflow || !inflow)      \
            return err_invalid_argument;                                                                     \
        Cells<T, Kind::kinematic_wave> cells{input<T>(current_outflow, 0), input<T>(inflow, 0),               \
                                             input(alpha, alpha_scalar), input(beta, beta_scalar),           \
                                             input(channel_length, channel_length_scalar),                   \
                                             time_step_duration, outflow, nullptr};                          \
        return route<T>(g, ldd, cells, nthreads);                                                            \
    }                                                                                                        \
    T oracle_iterate_to_new_discharge_##SUFFIX(T q_up, T q_cur, T q_lat, T alpha, T beta, T dt, T dx)       \
    {                                                                                                        \
        return iterate_to_new_discharge<T>(q_up, q_cur, q_lat, alpha, beta, dt, dx);                         \
    }

ORACLE_ROUTING_OPS(float, f32)
ORACLE_ROUTING_OPS(double, f64)

// Cell classes of accu_info (uint8): the router's events replayed with accu_info's functor
// (definition/accu_info.hpp:53-145, value policies: a no-data ldd gives class 255).
int oracle_accu_info(std::uint8_t const* ldd, Index rows, Index cols, Index prows, Index pcols,
                     std::uint8_t* cell_class, int nthreads)
{
    Grid g;
    if (!make_grid(rows, cols, prows, pcols, g) || !ldd || !cell_class)
    {
        return err_invalid_argument;
    }
    CellInfo cells{cell_class, nullptr};
    return route<std::uint8_t>(g, ldd, cells, nthreads);
}

int oracle_d8_flow_direction_f32(float const* elevation, Index rows, Index cols, std::uint8_t* out)
{
    if (rows < 0 || cols < 0 || (rows * cols > 0 && (!elevation || !out)))
        return err_invalid_argument;
    d8_flow_direction<float>(elevation, rows, cols, out);
    return ok;
}

int oracle_d8_flow_direction_f64(double const* elevation, Index rows, Index cols, std::uint8_t* out)
{
    if (rows < 0 || cols < 0 || (rows * cols > 0 && (!elevation || !out)))
        return err_invalid_argument;
    d8_flow_direction<double>(elevation, rows, cols, out);
    return ok;
}

#define ORACLE_NEIGHBOUR_OPS(T, SUFFIX)                                                                       \
    int oracle_upstream_##SUFFIX(std::uint8_t const* ldd, T const* material, Index rows, Index cols, T* out)  \
    {                                                                                                        \
        if (rows < 0 || cols < 0 || (rows * cols > 0 && (!ldd || !material || !out)))                      \
            return err_invalid_argument;                                                                     \
        upstream<T>(ldd, material, rows, cols, out);                                                         \
        return ok;                                                                                           \
    }                                                                                                        \
    int oracle_downstream_##SUFFIX(std::uint8_t const* ldd, T const* material, Index rows, Index cols, T* out) \
    {                                                                                                        \
        if (rows < 0 || cols < 0 || (rows * cols > 0 && (!ldd || !material || !out)))                      \
            return err_invalid_argument;                                                                     \
        downstream<T>(ldd, material, rows, cols, out);                                                       \
        return ok;                                                                                           \
    }

int oracle_downstream_distance_f32(std::uint8_t const* ldd, Index rows, Index cols, float cell_size,
                                   int mark_no_data, float* out)
{
    if (rows < 0 || cols < 0 || (rows * cols > 0 && (!ldd || !out)))
        return err_invalid_argument;
    downstream_distance<float>(ldd, rows * cols, cell_size, mark_no_data != 0, out);
    return ok;
}

int oracle_downstream_distance_f64(std::uint8_t const* ldd, Index rows, Index cols, double cell_size,
                                   int mark_no_data, double* out)
{
    if (rows < 0 || cols < 0 || (rows * cols > 0 && (!ldd || !out)))
        return err_invalid_argument;
    downstream_distance<double>(ldd, rows * cols, cell_size, mark_no_data != 0, out);
    return ok;
}

ORACLE_NEIGHBOUR_OPS(std::uint8_t, u8)
ORACLE_NEIGHBOUR_OPS(std::uint32_t, u32)
ORACLE_NEIGHBOUR_OPS(std::uint64_t, u64)
ORACLE_NEIGHBOUR_OPS(std::int8_t, i8)
ORACLE_NEIGHBOUR_OPS(std::int32_t, i32)
ORACLE_NEIGHBOUR_OPS(std::int64_t, i64)
ORACLE_NEIGHBOUR_OPS(float, f32)
ORACLE_NEIGHBOUR_OPS(double, f64)

}  // extern "C"
