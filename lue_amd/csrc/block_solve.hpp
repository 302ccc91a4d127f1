// Solving one block of up to 64 consecutive chain positions with one wavefront (lane j: position j).
// Used by the chain phase (chains.hpp) over the cells round 0 leaves pending, and by round 0 itself over
// the in-tile chains (tile_layout.hpp). A chain is a path of cells in flow order; a block may hold several
// short chains (a set `start` marks the first, most upstream, cell of each) or a part of a long one (the
// state of the previous block is carried over).
#pragma once

#include "cells.hpp"
#include "kinematic_wave.hpp"
#include "routing_common.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace lue_amd {

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

    template<int OP, typename T>
    struct ChainLane
    {
        bool valid;
        bool start;    // first cell of a chain
        double extra;  // side inflow (FP64)
        CellInput<OP, T> in;
    };

    // Inclusive scan of affine maps along the lanes of a wavefront. Lane k holds x_k = add + mul x_{k-1};
    // `flag` marks lanes that do not depend on the lanes before them (chain starts). On return
    // x_k = add + mul x_in where x_in is the value before the first lane, or x_k = add where flag is set.
    __device__ inline void affine_scan(double& add, double& mul, int& flag)
    {
        int const lane = threadIdx.x & 63;
        for (int offset = 1; offset < 64; offset <<= 1)
        {
            double const a_prev = __shfl_up(add, offset, 64);
            double const m_prev = __shfl_up(mul, offset, 64);
            int const f_prev = __shfl_up(flag, offset, 64);
            if (lane >= offset && !flag)
            {
                add = fma(mul, a_prev, add);
                mul = mul * m_prev;
                flag = f_prev;
            }
        }
    }

    // Kinematic wave on one block of chain positions, all lanes at once. Newton's method on the coupled
    // system of the block
    //     F_k(Q) = c1_k (Q_k - Q_{k-1} - extra_k) + alpha_k Q_k^beta_k - b_k = 0,
    // Q_{k-1} being the previous position on the chain (`carry` before the first lane, nothing at a chain
    // start); per cell this is LUE's equation with Q_up = Q_{k-1} + extra_k. The Jacobian is lower
    // bidiagonal (d_k = c1_k + alpha_k beta_k Q_k^(beta_k - 1) on the diagonal, -c1_k below it), so a
    // Newton step dQ_k = (-F_k + c1_k dQ_{k-1}) / d_k is one affine scan. The start is the system linearised
    // at Q_cur, an affine recurrence as well. F is concave in every Q_k and the Jacobian inverse-positive,
    // so the iterates approach the root from below after the first step. The iteration stops when every
    // |dQ_k| <= |Q_k| 2^(1 - digits), LUE's criterion (digits 31 for double, 14 for float).
    // T = float: the iterations use the hardware FP32 power (log2, exp2) and approximate reciprocals; one
    // FP64 Newton step follows (the FP64 residual correction).
    // Returns false, with q untouched, when a lane needs LUE's per-cell iteration: no-data, a zero current
    // discharge (dry cells and cells whose iteration starts from numeric_limits::min()), a non-positive
    // iterate, or no convergence within 16 steps.
    template<typename T>
    __device__ inline bool kinematic_wave_block(ChainLane<op_kinematic_wave, T> const& d, double carry, double& q)
    {
        int const lane = threadIdx.x & 63;
        bool const valid = d.valid;
        bool const start = valid && d.start;
        double const q_cur = static_cast<double>(d.in.q_cur);
        if (__ballot(valid && !(q_cur > 0.0 && !isnan(d.extra))))
            return false;
        double const c1 = d.in.c1;
        double const alpha = static_cast<double>(d.in.alpha);
        double const beta = static_cast<double>(d.in.beta);
        double const b = d.in.b;
        float const beta32 = static_cast<float>(d.in.beta);

        // start: F linearised at Q_cur gives Q_k = w_k (Q_{k-1} + extra_k) + g_k
        double add = 0.0;
        double mul = 0.0;
        int flag = (start || !valid) ? 1 : 0;
        if (valid)
        {
            double const pc = static_cast<double>(fast_exp2(beta32 * fast_log2(static_cast<float>(q_cur))));
            double const inv = 1.0 / (c1 + alpha * beta * pc / q_cur);
            mul = c1 * inv;
            add = mul * d.extra + (b - (1.0 - beta) * alpha * pc) * inv;
        }
        affine_scan(add, mul, flag);
        double x = valid ? (flag ? add : fma(mul, carry, add)) : 1.0;
        if (__ballot(valid && !(x > 0.0)))
            return false;

        constexpr double tol = std::is_same_v<T, double> ? 0x1p-30 : 0x1p-13;
        bool converged = false;
        for (int iteration = 0; iteration < 16 && !converged; ++iteration)
        {
            double prev = __shfl_up(x, 1, 64);
            if (lane == 0)
                prev = carry;
            if (start)
                prev = 0.0;
            double xb;
            double inv;
            if constexpr (std::is_same_v<T, double>)
            {
                xb = dpow(x, beta);
            }
            else
            {
                xb = static_cast<double>(fast_exp2(beta32 * fast_log2(static_cast<float>(x))));
            }
            double const ax = alpha * xb;
            double const f = c1 * (x - prev - d.extra) + ax - b;
            double const fp = c1 + beta * ax / x;
            if constexpr (std::is_same_v<T, double>)
                inv = 1.0 / fp;
            else
                inv = static_cast<double>(fast_rcp(static_cast<float>(fp)));
            add = valid ? -f * inv : 0.0;
            mul = valid ? c1 * inv : 0.0;
            flag = (start || !valid) ? 1 : 0;
            affine_scan(add, mul, flag);
            double const xn = x + add;  // the carry is exact: its step is 0
            if (__ballot(valid && !(xn > 0.0)))
                return false;
            converged = __ballot(valid && !(fabs(add) <= fabs(xn) * tol)) == 0;
            x = xn;
        }
        if (!converged)
            return false;
        if constexpr (std::is_same_v<T, float>)
        {
            // FP64 residual correction
            double prev = __shfl_up(x, 1, 64);
            if (lane == 0)
                prev = carry;
            if (start)
                prev = 0.0;
            double const ax = alpha * dpow(x, beta);
            double const f = c1 * (x - prev - d.extra) + ax - b;
            double const inv = 1.0 / (c1 + beta * ax / x);
            add = valid ? -f * inv : 0.0;
            mul = valid ? c1 * inv : 0.0;
            flag = (start || !valid) ? 1 : 0;
            affine_scan(add, mul, flag);
            double const xn = x + add;
            if (__ballot(valid && !(xn > 0.0)))
                return false;
            x = xn;
        }
        q = valid ? x : 0.0;
        return true;
    }

    // State carried from one block of chain positions to the next by the wavefront that runs them.
    struct BlockState
    {
        double carry{0.0};  // outflow of the previous chain position (FP64)
        float warm{0.0f};   // kinematic wave FP32: an iterate of the previous cell, to start Newton from
        float warm_pow{0.0f};
        bool have_warm{false};
    };

    // Run the cell function along one block of up to 64 chain positions (lane j: position j; `nb` valid
    // lanes). Chains start where d.start is set; the first lane continues the chain of st.carry otherwise.
    // Linear functions are scans; the kinematic wave is a block Newton (kinematic_wave_block) with LUE's
    // per-cell iteration for blocks that need it; threshold and fraction run serially with lane broadcasts.
    // On return `value` (FP64 outflow) and `rem` hold this lane's results and st the state for the next
    // block.
    template<int OP, typename T>
    __device__ inline void solve_block(SweepArgs<T> const& a, ChainLane<OP, T> const& d, int nb, bool uniform_beta,
                                       BlockState& st, double& value, T& rem)
    {
        int const lane = threadIdx.x & 63;
        double const nan64 = __longlong_as_double(0x7ff8000000000000ll);
        value = 0.0;
        if constexpr (OP == op_accu)
        {
            // linear: out_k = out_{k-1} + side_k + x_k, a segmented inclusive scan along the block
            double term = !d.valid ? 0.0 : (isnan(d.in.x) ? nan64 : d.extra + static_cast<double>(d.in.x));
            int flag = d.valid && d.start ? 1 : 0;
            for (int offset = 1; offset < 64; offset <<= 1)
            {
                double const t_prev = __shfl_up(term, offset, 64);
                int const f_prev = __shfl_up(flag, offset, 64);
                if (lane >= offset)
                {
                    if (!flag)
                        term += t_prev;
                    flag |= f_prev;
                }
            }
            value = flag ? term : st.carry + term;
            st.carry = read_lane(value, nb - 1);
        }
        else if constexpr (OP == op_threshold || OP == op_fraction)
        {
            for (int j = 0; j < nb; ++j)
            {
                int const first = read_lane(static_cast<int>(d.start), j);
                double const up = (first ? 0.0 : st.carry) + read_lane(d.extra, j);
                CellInput<OP, T> cj{read_lane(d.in.x, j), read_lane(d.in.p, j)};
                double out;
                T r{};
                solve_cell<OP, T>(a, cj, up, isnan(up), out, r);
                if (lane == j)
                {
                    value = out;
                    rem = r;
                }
                st.carry = out;
            }
        }
        else if (kinematic_wave_block<T>(d, st.carry, value))
        {
            st.carry = read_lane(value, nb - 1);
            st.have_warm = false;
        }
        else if constexpr (std::is_same_v<T, double>)
        {
            // plain FP64 kinematic wave: LUE's Newton per cell along the block
            for (int j = 0; j < nb; ++j)
            {
                int const first = read_lane(static_cast<int>(d.start), j);
                double const up = (first ? 0.0 : st.carry) + read_lane(d.extra, j);
                CellInput<OP, T> cj{read_lane(d.in.q_cur, j), read_lane(d.in.q_lat, j), read_lane(d.in.alpha, j),
                                    read_lane(d.in.beta, j), read_lane(d.in.b, j), read_lane(d.in.c1, j)};
                double v;
                T r{};
                solve_cell<OP, T>(a, cj, up, isnan(up), v, r);
                if (lane == j)
                    value = v;
                st.carry = v;
            }
        }
        else
        {
            // FP32 kinematic wave: serial FP32 Newton (warm start from the previous cell), then the FP64
            // correction of every cell in parallel, propagated along the chains to first order
            double q_up_mine = 0.0;
            float q32_mine = 0.0f;
            for (int j = 0; j < nb; ++j)
            {
                int const first = read_lane(static_cast<int>(d.start), j);
                if (first)
                    st.have_warm = false;
                double const q_up = (first ? 0.0 : st.carry) + read_lane(d.extra, j);
                float const q_cur = read_lane(d.in.q_cur, j);
                float q;
                if (isnan(q_up) || isnan(q_cur))
                {
                    q = __int_as_float(0x7fc00000);
                    st.have_warm = false;
                }
                else
                {
                    float const q_lat = read_lane(d.in.q_lat, j);
                    float const alpha = read_lane(d.in.alpha, j);
                    float const beta = read_lane(d.in.beta, j);
                    double const b = read_lane(d.in.b, j);
                    double const c1 = read_lane(d.in.c1, j);
                    st.have_warm = st.have_warm && uniform_beta;
                    q = kinematic_wave_root_f32(static_cast<float>(c1), alpha, beta, a.dt, q_up, q_cur, q_lat, b, c1,
                                                st.have_warm, st.warm, st.warm_pow);
                }
                if (lane == j)
                {
                    q_up_mine = q_up;
                    q32_mine = q;
                }
                st.carry = static_cast<double>(q);
            }
            double eps = 0.0;
            double sens = 0.0;
            double const q = static_cast<double>(q32_mine);
            if (d.valid && q > 0.0 && !isnan(q))
            {
                kinematic_wave_correction(q, q_up_mine, static_cast<double>(d.in.alpha),
                                          static_cast<double>(d.in.beta), d.in.b, d.in.c1, eps, sens);
            }
            // affine scan: delta_k = eps_k + sens_k * delta_{k-1}; a chain start does not depend on
            // the cell before it, and the first cell of the block starts from a corrected value
            double mul = (d.valid && d.start) ? 0.0 : sens;
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
            st.carry = read_lane(value, nb - 1);
        }
    }

}  // namespace lue_amd
