// Long chains with a whole workgroup.
//
// A level of the chain schedule takes as long as its longest work item: one wavefront walks it 64 positions
// at a time, each block waiting for the carry of the one before. The last levels of a large network hold
// few but long chains (at 16384^2 cells: 150 chains up to 16 385 positions in the last two levels), so most
// of the device idles there. For levels whose longest work item reaches wide_min_positions, a workgroup of
// wide_waves wavefronts takes each work item and advances 64 * wide_waves positions at a time:
//   - accu: each wavefront scans its 64 positions; the wavefronts' aggregates (sum, chain-start flag) are
//     composed in order through LDS, starting from the carry. This is exactly the composition the one-wavefront
//     kernel performs block after block, so the results are the same bits.
//   - threshold, capacity, fraction: each wavefront scans its 64 positions (the maps of solve_block); the value
//     entering wavefront w + 1 is the cell function of wavefront w's last position applied to the map of the
//     positions before it and the value entering w, one after the other by one thread (16 short steps). That
//     is the arithmetic of the one-wavefront kernel block after block, so the results are the same bits;
//   - kinematic wave: the block Newton of kinematic_wave_block over all 64 * wide_waves positions at once (the
//     Jacobian of the chain stays lower bidiagonal; every affine scan is a wavefront scan followed by the
//     composition of the wavefront aggregates, and the value before each wavefront's first lane comes from the
//     previous wavefront's last lane through LDS). It converges to the same roots with the same FP64
//     correction and the same check of LUE's own iteration; where any lane needs LUE's per-cell iteration the
//     whole group of positions is solved by the one-wavefront block solver instead.
// Workgroup-wide agreements (convergence, fallbacks) use __syncthreads_or, so every wavefront of the
// workgroup takes the same branches and reaches the same barriers.
#pragma once

#include "block_solve.hpp"
#include "chains.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace lue_amd {

    constexpr int wide_waves = 16;
    constexpr int wide_threads = 64 * wide_waves;
    // levels whose longest work item has at least this many positions run on the wide path
    constexpr std::uint32_t wide_min_positions = 2048;

    template<int OP>
    constexpr bool has_wide_path()
    {
        return OP == op_accu || OP == op_kinematic_wave || OP == op_threshold || OP == op_capacity ||
               OP == op_fraction;
    }

    // Exchange slots of the wavefronts. Every exchange writes one buffer of a pair and is read after one
    // barrier; the next write to that buffer comes two exchanges later, after another barrier, so no barrier
    // is needed to free a slot.
    struct WideLds
    {
        double last[2][wide_waves];  // lane 63's value per wavefront
        int flags[2][wide_waves];    // per wavefront: bit 0 a lane did not converge, bit 1 a lane is out of range
        double add[2][wide_waves];   // wavefront aggregates of a scan
        double mul[2][wide_waves];
        int flag[wide_waves];        // accu: chain-start flag of the wavefront aggregate
        double carry;                // value of the last position of a group, for the next group
        // threshold, capacity, fraction: per wavefront its last position's cell and the value entering it
        int cell_valid[wide_waves];
        int cell_start[wide_waves];
        double cell_extra[wide_waves];
        double cell_x[wide_waves];
        double cell_p[wide_waves];
        double entering[wide_waves];
    };

    constexpr int wide_not_converged = 1;
    constexpr int wide_out_of_range = 2;

    __device__ inline int wide_wave()
    {
        return static_cast<int>(threadIdx.x >> 6);
    }

    // Value of the previous chain position for every lane (within a wavefront by a lane shift, for lane 0 of
    // wavefront w from lane 63 of wavefront w - 1, for lane 0 of wavefront 0 `before`), and the OR of the
    // wavefronts' flags (`mine`: this wavefront's, the same in all its lanes). One barrier.
    __device__ inline double wide_previous(double x, double before, int mine, int& all, WideLds& s, int& buffer)
    {
        int const lane = threadIdx.x & 63;
        int const wave = wide_wave();
        int const b = buffer;
        buffer ^= 1;
        if (lane == 63)
        {
            s.last[b][wave] = x;
            s.flags[b][wave] = mine;
        }
        __syncthreads();
        double prev = wave_shr1(x);
        if (lane == 0)
            prev = wave == 0 ? before : s.last[b][wave - 1];
        int const f = lane < wide_waves ? s.flags[b][lane] : 0;
        all = __reduce_or_sync(~0ull, f);
        return prev;
    }

    // Affine scan across the wavefronts: (add, mul) hold this wavefront's inclusive scan (x_k = add + mul x_in).
    // Returns x_in of this wavefront: x0 for wavefront 0, else the aggregates of the wavefronts before it
    // composed in order (an FP64 scan over lanes 0..15, identity for the wavefronts from this one on), applied
    // to x0. One barrier.
    __device__ inline double wide_affine_in(double add, double mul, double x0, WideLds& s, int& buffer)
    {
        int const lane = threadIdx.x & 63;
        int const wave = wide_wave();
        int const b = buffer;
        buffer ^= 1;
        if (lane == 63)
        {
            s.add[b][wave] = add;
            s.mul[b][wave] = mul;
        }
        __syncthreads();
        double a = lane < wave ? s.add[b][lane] : 0.0;
        double m = lane < wave ? s.mul[b][lane] : 1.0;
        affine_scan_nf(a, m);  // lane 15 now holds the composition of the wavefronts 0 .. wave - 1
        double const a_all = read_lane(a, 15);
        double const m_all = read_lane(m, 15);
        return fma(m_all, x0, a_all);
    }

    // kinematic_wave_block over the positions of the whole workgroup (see there for the method and the
    // tolerances). Returns kw_block_done with q set, or kw_block_serial when any lane needs LUE's per-cell
    // iteration or LUE's own iteration would stop short of the root somewhere (the caller then solves the
    // positions with the one-wavefront block solver); the outcome is the same in every lane of the workgroup.
    template<typename T>
    __device__ inline int kinematic_wave_wide(ChainLane<op_kinematic_wave, T> const& d, double carry, double dt,
                                              double& q, WideLds& s)
    {
        bool const valid = d.valid;
        bool const start = valid && d.start;
        double const q_cur = static_cast<double>(d.in.q_cur);
        if (__syncthreads_or(valid && !(q_cur > 0.0 && !isnan(d.extra))))
            return kw_block_serial;
        double const c1 = kw_c1(d.in);
        double const alpha = static_cast<double>(d.in.alpha);
        double const beta = static_cast<double>(d.in.beta);
        double const b = kw_b(d.in);
        float const c1f = static_cast<float>(c1);
        float const beta32 = static_cast<float>(d.in.beta);
        float const ab32 = static_cast<float>(d.in.alpha) * beta32;
        bool const chained = valid && !start;
        int xbuf = 0;  // exchange buffers of wide_previous and wide_affine_in
        int abuf = 0;

        // start: F linearised at Q_cur, an affine recurrence (FP32 scan, FP64 composition of the wavefronts)
        double x;
        float pc1 = 1.0f;
        {
            float add = 0.0f;
            float mul = 0.0f;
            if (valid)
            {
                float const q_cur32 = static_cast<float>(q_cur);
                pc1 = fast_exp2((beta32 - 1.0f) * fast_log2(q_cur32));
                float const inv = fast_rcp(c1f + ab32 * pc1);
                float const w = c1f * inv;
                mul = chained ? w : 0.0f;
                add = w * static_cast<float>(d.extra) +
                      (static_cast<float>(b) - (1.0f - beta32) * static_cast<float>(alpha) * q_cur32 * pc1) * inv;
            }
            affine_scan_f32_nf(add, mul);
            double const x_in = wide_affine_in(static_cast<double>(add), static_cast<double>(mul), carry, s, abuf);
            x = valid ? fma(static_cast<double>(mul), x_in, static_cast<double>(add)) : 1.0;
        }
        // this wavefront's flags for the next exchange
        int mine = __ballot(valid && !(x > 0.0 && x < __builtin_huge_val())) != 0 ? wide_out_of_range : 0;

        constexpr double tol = std::is_same_v<T, double> ? 0x1p-30 : 0x1p-8;
        float const extra32 = static_cast<float>(d.extra);
        float const alpha32 = static_cast<float>(alpha);
        float const b32 = static_cast<float>(b);
        int all = 0;
        double prev = 0.0;
        bool converged = false;
        for (int iteration = 0; iteration <= 16; ++iteration)
        {
            // the previous positions of x, with the flags of the step that produced x
            prev = wide_previous(x, carry, mine, all, s, xbuf);
            if (all & wide_out_of_range)
                return kw_block_serial;
            if (iteration > 0 && !(all & wide_not_converged))
            {
                converged = true;
                break;
            }
            if (iteration == 16)
                break;
            if (start)
                prev = 0.0;
            double step;
            if constexpr (std::is_same_v<T, double>)
            {
                double const pw1 = pow_positive<false>(x, beta - 1.0);
                double const f = c1 * (x - prev - d.extra) + alpha * x * pw1 - b;
                double const fp = c1 + alpha * beta * pw1;
                double inv = static_cast<double>(fast_rcp(static_cast<float>(fp)));
                inv = inv * fma(-fp, inv, 2.0);
                inv = inv * fma(-fp, inv, 2.0);
                double add = valid ? -f * inv : 0.0;
                double mul = chained ? c1 * inv : 0.0;
                affine_scan_nf(add, mul);
                double const step_in = wide_affine_in(add, mul, 0.0, s, abuf);  // the carry is exact: its step is 0
                step = fma(mul, step_in, add);
            }
            else
            {
                float const x32 = static_cast<float>(x);
                float const prev32 = static_cast<float>(prev);
                float const pw1 = fast_exp2((beta32 - 1.0f) * fast_log2(x32));
                float const f = c1f * (x32 - prev32 - extra32) + alpha32 * x32 * pw1 - b32;
                float const inv = fast_rcp(c1f + ab32 * pw1);
                float add = valid ? -f * inv : 0.0f;
                float mul = chained ? c1f * inv : 0.0f;
                affine_scan_f32_nf(add, mul);
                double const step_in =
                    wide_affine_in(static_cast<double>(add), static_cast<double>(mul), 0.0, s, abuf);
                step = fma(static_cast<double>(mul), step_in, static_cast<double>(add));
            }
            double const xn = x + step;
            mine = (__ballot(valid && !(xn > 0.0 && xn < __builtin_huge_val())) != 0 ? wide_out_of_range : 0) |
                   (__ballot(valid && !(fabs(step) <= fabs(xn) * tol)) != 0 ? wide_not_converged : 0);
            x = xn;
        }
        if (!converged)
            return kw_block_serial;
        // prev now holds the previous positions of the converged x
        if constexpr (std::is_same_v<T, float>)
        {
            // FP64 residual correction (as in kinematic_wave_block)
            double const p0 = start ? 0.0 : prev;
            double const pw1 = pow_positive<true, true>(x, beta - 1.0);
            double const f = c1 * (x - p0 - d.extra) + alpha * x * pw1 - b;
            float const inv = fast_rcp(c1f + ab32 * static_cast<float>(pw1));
            float add = valid ? static_cast<float>(-f) * inv : 0.0f;
            float mul = chained ? c1f * inv : 0.0f;
            affine_scan_f32_nf(add, mul);
            double const step_in = wide_affine_in(static_cast<double>(add), static_cast<double>(mul), 0.0, s, abuf);
            double const xn = x + fma(static_cast<double>(mul), step_in, static_cast<double>(add));
            mine = __ballot(valid && !(xn > 0.0 && xn < __builtin_huge_val())) != 0 ? wide_out_of_range : 0;
            x = xn;
            prev = wide_previous(x, carry, mine, all, s, xbuf);
            if (all & wide_out_of_range)
                return kw_block_serial;
        }
        // LUE's own iteration ends at this root only from a guess near it (lue_guess_converges)
        {
            double const p0 = start ? 0.0 : prev;
            float dtq;
            if constexpr (std::is_same_v<T, float>)
                dtq = static_cast<float>(b - alpha * q_cur * static_cast<double>(pc1));
            else
                dtq = static_cast<float>(dt * d.in.q_lat);
            bool const far = valid && !lue_guess_converges(static_cast<float>(p0 + d.extra), static_cast<float>(q_cur),
                                                           dtq, static_cast<float>(alpha), beta32, c1f,
                                                           static_cast<float>(x));
            if (__syncthreads_or(far))
                return kw_block_serial;
        }
        q = valid ? x : 0.0;
        return kw_block_done;
    }

    // One work item per workgroup of wide_threads threads, 64 * wide_waves positions at a time (the next group's
    // data in flight while a group runs).
    template<int OP, typename T>
    __global__ void __launch_bounds__(wide_threads)
        chains_run_wide_kernel(SweepArgs<T> a, ChainRun<OP, T> c, std::uint32_t item_begin)
    {
        static_assert(has_wide_path<OP>(), "the wide path serves accu, threshold, capacity, fraction and the kinematic wave");
        __shared__ WideLds s;
        int const lane = static_cast<int>(threadIdx.x & 63);
        int const wave = wide_wave();
        std::uint32_t const item = item_begin + blockIdx.x;
        std::uint32_t const start = as_global(c.item_start)[item];
        std::uint32_t const count = as_global(c.item_count)[item];
        double carry = 0.0;
        std::uint32_t const mine = static_cast<std::uint32_t>(64 * wave + lane);
        ChainLane<OP, T> next = load_chain_lane<OP, T>(a, c, start, count, mine);
        for (std::uint32_t base = 0; base < count; base += wide_threads)
        {
            ChainLane<OP, T> const d = next;
            if (base + wide_threads < count)
                next = load_chain_lane<OP, T>(a, c, start, count, base + wide_threads + mine);
            std::uint32_t const p = start + base + mine;
            if constexpr (OP == op_accu)
            {
                // the same composition as block after block in chains_run_kernel
                double const nan64 = __longlong_as_double(0x7ff8000000000000ll);
                double term = !d.valid ? 0.0 : (isnan(d.in.x) ? nan64 : d.extra + static_cast<double>(d.in.x));
                int flag = d.valid && d.start ? 1 : 0;
                sum_scan(term, flag);
                if (lane == 63)
                {
                    s.add[0][wave] = term;
                    s.flag[wave] = flag;
                }
                __syncthreads();
                double x = carry;
                for (int v = 0; v < wave; ++v)
                    x = s.flag[v] ? s.add[0][v] : x + s.add[0][v];
                double const value = flag ? term : x + term;
                double next_carry = carry;
                for (int v = 0; v < wide_waves; ++v)
                    next_carry = s.flag[v] ? s.add[0][v] : next_carry + s.add[0][v];
                __syncthreads();
                carry = next_carry;
                if (d.valid)
                {
                    as_global(c.res)[p] = static_cast<T>(value);
                    if (d.push != no_slot)
                        as_global(c.slot)[d.push] = value;
                }
            }
            else if constexpr (OP == op_threshold || OP == op_capacity || OP == op_fraction)
            {
                // the maps of solve_block (value before this position -> value at it) scanned per wavefront
                int flag = d.valid && d.start ? 1 : 0;
                double const ex = d.extra + static_cast<double>(d.in.x);
                double const pp = static_cast<double>(d.in.p);
                double ma;  // fraction: add, clamped sums: c
                double mb;  // fraction: mul, clamped sums: d
                if constexpr (OP == op_fraction)
                {
                    ma = d.valid ? pp * ex : 0.0;
                    mb = d.valid ? pp : 1.0;
                    affine_scan(ma, mb, flag);
                }
                else
                {
                    constexpr bool is_max = OP == op_threshold;
                    mb = !d.valid ? clamp_identity<is_max>() : (is_max ? 0.0 : pp);
                    ma = !d.valid ? 0.0 : (is_max ? ex - pp : ex);
                    clamp_scan<is_max>(mb, ma, flag);
                }
                auto apply = [](double a_, double b_, int f_, double before) -> double
                {
                    if constexpr (OP == op_fraction)
                        return f_ ? a_ : fma(b_, before, a_);
                    else
                        return extremum_nan<OP == op_threshold>(b_, (f_ ? 0.0 : before) + a_);
                };
                if (lane == 62)
                {
                    s.add[0][wave] = ma;
                    s.mul[0][wave] = mb;
                    s.flag[wave] = flag;
                }
                if (lane == 63)
                {
                    s.cell_valid[wave] = d.valid ? 1 : 0;
                    s.cell_start[wave] = d.valid && d.start ? 1 : 0;
                    s.cell_extra[wave] = d.extra;
                    s.cell_x[wave] = static_cast<double>(d.in.x);
                    s.cell_p[wave] = static_cast<double>(d.in.p);
                }
                __syncthreads();
                if (threadIdx.x == 0)
                {
                    // the value entering each wavefront: its predecessor's last cell, block after block
                    double before = carry;
                    for (int v = 0; v < wide_waves; ++v)
                    {
                        s.entering[v] = before;
                        if (!s.cell_valid[v])
                            continue;  // the work item ends in this wavefront
                        double const x62 = apply(s.add[0][v], s.mul[0][v], s.flag[v], before);
                        double const up = (s.cell_start[v] ? 0.0 : x62) + s.cell_extra[v];
                        CellInput<OP, T> const in{static_cast<T>(s.cell_x[v]), static_cast<T>(s.cell_p[v])};
                        double out = 0.0;
                        T r{};
                        solve_cell<OP, T>(a, in, up, isnan(up), out, r);
                        before = out;
                    }
                    s.carry = before;
                }
                __syncthreads();
                double const before = s.entering[wave];
                double const x = apply(ma, mb, flag, before);
                double prev = wave_shr1(x);
                if (lane == 0)
                    prev = before;
                if (d.start)
                    prev = 0.0;
                double const up = prev + d.extra;
                if (d.valid)
                {
                    double out = 0.0;
                    T r{};
                    solve_cell<OP, T>(a, d.in, up, isnan(up), out, r);
                    as_global(c.res)[p] = static_cast<T>(out);
                    as_global(c.res_rem)[p] = r;
                    if (d.push != no_slot)
                        as_global(c.slot)[d.push] = out;
                }
                carry = s.carry;
                __syncthreads();  // s is written again by the next group
            }
            else
            {
                double value = 0.0;
                int const outcome = kinematic_wave_wide<T>(d, carry, static_cast<double>(a.dt), value, s);
                if (outcome == kw_block_done)
                {
                    if (d.valid)
                    {
                        as_global(c.res)[p] = static_cast<T>(value);
                        if (d.push != no_slot)
                            as_global(c.slot)[d.push] = value;
                    }
                    std::uint32_t const last = (count - base < wide_threads ? count - base : wide_threads) - 1u;
                    if (mine == last)
                        s.carry = value;
                }
                else
                {
                    // the one-wavefront block solver over these positions, 64 at a time (LUE's per-cell
                    // iteration where a lane needs it)
                    if (wave == 0)
                    {
                        BlockState st{carry};
                        for (int k = 0; k < wide_waves; ++k)
                        {
                            std::uint32_t const block = base + static_cast<std::uint32_t>(64 * k);
                            if (block >= count)
                                break;
                            ChainLane<OP, T> const dk = load_chain_lane<OP, T>(a, c, start, count, block + lane);
                            int const nb = static_cast<int>(count - block < 64 ? count - block : 64);
                            double v = 0.0;
                            T r{};
                            solve_block<OP, T>(a, dk, nb, st, v, r);
                            if (dk.valid)
                            {
                                as_global(c.res)[start + block + lane] = static_cast<T>(v);
                                if (dk.push != no_slot)
                                    as_global(c.slot)[dk.push] = v;
                            }
                        }
                        if (lane == 0)
                            s.carry = st.carry;
                    }
                }
                __syncthreads();
                carry = s.carry;
                __syncthreads();
            }
        }
    }

}  // namespace lue_amd
