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
        std::uint32_t push;  // chain phase: side slot the position's outflow feeds (0xFFFFFFFF: none)
    };

    // Lane shift inside rows of 16 lanes (DPP row_shr:N): lane j gets the value of lane j - N of its row;
    // the first N lanes of a row get `identity`.
    template<int N>
    __device__ inline double row_shr(double v, double identity)
    {
        int const lo = __builtin_amdgcn_update_dpp(__double2loint(identity), __double2loint(v), 0x110 + N, 0xF, 0xF, false);
        int const hi = __builtin_amdgcn_update_dpp(__double2hiint(identity), __double2hiint(v), 0x110 + N, 0xF, 0xF, false);
        return __hiloint2double(hi, lo);
    }

    template<int N>
    __device__ inline float row_shr(float v, float identity)
    {
        return __int_as_float(
            __builtin_amdgcn_update_dpp(__float_as_int(identity), __float_as_int(v), 0x110 + N, 0xF, 0xF, false));
    }

    template<int N>
    __device__ inline int row_shr(int v, int identity)
    {
        return __builtin_amdgcn_update_dpp(identity, v, 0x110 + N, 0xF, 0xF, false);
    }

    // The same shift with 0 shifted in: bound_ctrl makes the lanes whose source lies before their row read 0,
    // so no register has to hold the identity first.
    template<int N>
    __device__ inline double row_shr0(double v)
    {
        int const lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x110 + N, 0xF, 0xF, true);
        int const hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x110 + N, 0xF, 0xF, true);
        return __hiloint2double(hi, lo);
    }

    template<int N>
    __device__ inline float row_shr0(float v)
    {
        return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x110 + N, 0xF, 0xF, true));
    }

    template<int N>
    __device__ inline int row_shr0(int v)
    {
        return __builtin_amdgcn_mov_dpp(v, 0x110 + N, 0xF, 0xF, true);
    }

    // Lane shift by one across the wavefront (DPP wave_shr:1): lane j gets the value of lane j - 1, lane 0
    // gets 0.
    __device__ inline double wave_shr1(double v)
    {
        int const lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xF, 0xF, true);
        int const hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xF, 0xF, true);
        return __hiloint2double(hi, lo);
    }

    __device__ inline float wave_shr1(float v)
    {
        return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
    }

    // DPP row broadcasts across rows of 16 lanes: Ctrl 0x142 (row_bcast:15) with RowMask 0xA gives rows 1 and
    // 3 the value of lane 15 of the row before them; Ctrl 0x143 (row_bcast:31) with RowMask 0xC gives rows 2
    // and 3 the value of lane 31. Lanes of the other rows get `identity`.
    template<int Ctrl, int RowMask>
    __device__ inline double row_bcast(double v, double identity)
    {
        int const lo = __builtin_amdgcn_update_dpp(__double2loint(identity), __double2loint(v), Ctrl, RowMask, 0xF, false);
        int const hi = __builtin_amdgcn_update_dpp(__double2hiint(identity), __double2hiint(v), Ctrl, RowMask, 0xF, false);
        return __hiloint2double(hi, lo);
    }

    template<int Ctrl, int RowMask>
    __device__ inline float row_bcast(float v, float identity)
    {
        return __int_as_float(
            __builtin_amdgcn_update_dpp(__float_as_int(identity), __float_as_int(v), Ctrl, RowMask, 0xF, false));
    }

    template<int Ctrl, int RowMask>
    __device__ inline int row_bcast(int v, int identity)
    {
        return __builtin_amdgcn_update_dpp(identity, v, Ctrl, RowMask, 0xF, false);
    }

    constexpr int dpp_row_bcast15 = 0x142;
    constexpr int dpp_row_bcast31 = 0x143;

    // Segmented inclusive scans along the 64 lanes of a wavefront: Hillis-Steele inside rows of 16 lanes
    // (DPP shifts by 1, 2, 4, 8), then rows 1 and 3 take the prefix of the row before them (row_bcast:15)
    // and rows 2 and 3 that of rows 0-1 (row_bcast:31); the values stay in vector registers. `flag` marks lanes whose element does not combine with the lanes before
    // them (chain starts); on return it is set where the result does not depend on the value before lane 0.

    // The combine steps below select rather than branch: a lane whose flag is set keeps its element (the
    // combined value is computed and dropped), so the scans compile to straight-line DPP, FP64 and
    // v_cndmask instructions without exec-mask changes.

    // Affine maps: lane k holds x_k = add + mul x_{k-1}. On return x_k = add + mul x_in, x_in being the
    // value before lane 0 (or x_k = add where flag is set).
    __device__ inline void affine_combine(double a_prev, double m_prev, int f_prev, double& add, double& mul, int& flag)
    {
        double const a = fma(mul, a_prev, add);
        double const m = mul * m_prev;
        add = flag ? add : a;
        mul = flag ? mul : m;
        flag |= f_prev;
    }

    template<int N>
    __device__ inline void affine_step(double& add, double& mul, int& flag)
    {
        double const a_prev = row_shr0<N>(add);
        double const m_prev = row_shr<N>(mul, 1.0);
        int const f_prev = row_shr0<N>(flag);
        affine_combine(a_prev, m_prev, f_prev, add, mul, flag);
    }

    __device__ inline void affine_scan(double& add, double& mul, int& flag)
    {
        affine_step<1>(add, mul, flag);
        affine_step<2>(add, mul, flag);
        affine_step<4>(add, mul, flag);
        affine_step<8>(add, mul, flag);
        {
            double const a_prev = row_bcast<dpp_row_bcast15, 0xA>(add, 0.0);
            double const m_prev = row_bcast<dpp_row_bcast15, 0xA>(mul, 1.0);
            int const f_prev = row_bcast<dpp_row_bcast15, 0xA>(flag, 0);
            affine_combine(a_prev, m_prev, f_prev, add, mul, flag);
        }
        {
            double const a_prev = row_bcast<dpp_row_bcast31, 0xC>(add, 0.0);
            double const m_prev = row_bcast<dpp_row_bcast31, 0xC>(mul, 1.0);
            int const f_prev = row_bcast<dpp_row_bcast31, 0xC>(flag, 0);
            affine_combine(a_prev, m_prev, f_prev, add, mul, flag);
        }
    }

    // Inclusive sum over the 64 lanes of a wavefront (all lanes active).
    __device__ inline int wave_inclusive_sum(int x)
    {
        x += row_shr0<1>(x);
        x += row_shr0<2>(x);
        x += row_shr0<4>(x);
        x += row_shr0<8>(x);
        x += row_bcast<dpp_row_bcast15, 0xA>(x, 0);
        x += row_bcast<dpp_row_bcast31, 0xC>(x, 0);
        return x;
    }

    // Affine maps without start flags: a lane whose element does not depend on the lanes before it (a chain
    // start) has mul = 0, and compositions keep that (mul of the composition = product of the muls). On return
    // x_k = add + mul x_in for every lane. Needs finite add and mul (0 * inf or NaN would leak across a start):
    // the kinematic-wave block sends any block with a non-finite value to its per-cell path anyway.
    template<int N>
    __device__ inline void affine_step_nf(double& add, double& mul)
    {
        double const a_prev = row_shr0<N>(add);
        double const m_prev = row_shr<N>(mul, 1.0);
        add = fma(mul, a_prev, add);
        mul = mul * m_prev;
    }

    __device__ inline void affine_scan_nf(double& add, double& mul)
    {
        affine_step_nf<1>(add, mul);
        affine_step_nf<2>(add, mul);
        affine_step_nf<4>(add, mul);
        affine_step_nf<8>(add, mul);
        {
            double const a_prev = row_bcast<dpp_row_bcast15, 0xA>(add, 0.0);
            double const m_prev = row_bcast<dpp_row_bcast15, 0xA>(mul, 1.0);
            add = fma(mul, a_prev, add);
            mul = mul * m_prev;
        }
        {
            double const a_prev = row_bcast<dpp_row_bcast31, 0xC>(add, 0.0);
            double const m_prev = row_bcast<dpp_row_bcast31, 0xC>(mul, 1.0);
            add = fma(mul, a_prev, add);
            mul = mul * m_prev;
        }
    }

    // The FP32 flagless affine scan with the lane shifts folded into the arithmetic (DPP source operands), two
    // instructions per step where separate moves take five: v_fmac_f32_dpp with bound_ctrl reads 0 from a
    // source lane outside the row (add unchanged there), v_mul_f32_dpp without bound_ctrl does not write a
    // lane whose source lies outside the row or whose row the row mask leaves out (mul unchanged there: the
    // identity 1 without a register holding it). Same operations as the step
    //     add = fma(mul, add[lane - N], add); mul = mul * mul[lane - N]
    // (products commute exactly), so the results are the same bits. Every DPP instruction is preceded by
    // s_nop 1: DPP must not read a VGPR written by the VALU instruction just before it (two wait states).
    // All 64 lanes are active where the scans run.
    __device__ inline void affine_scan_f32_nf(float& add, float& mul)
    {
        asm volatile(
            "s_nop 1\n\t"
            "v_fmac_f32_dpp %0, %0, %1 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
            "s_nop 1\n\t"
            "v_mul_f32_dpp %1, %1, %1 row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
            "s_nop 1\n\t"
            "v_fmac_f32_dpp %0, %0, %1 row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
            "s_nop 1\n\t"
            "v_mul_f32_dpp %1, %1, %1 row_shr:2 row_mask:0xf bank_mask:0xf\n\t"
            "s_nop 1\n\t"
            "v_fmac_f32_dpp %0, %0, %1 row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
            "s_nop 1\n\t"
            "v_mul_f32_dpp %1, %1, %1 row_shr:4 row_mask:0xf bank_mask:0xf\n\t"
            "s_nop 1\n\t"
            "v_fmac_f32_dpp %0, %0, %1 row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
            "s_nop 1\n\t"
            "v_mul_f32_dpp %1, %1, %1 row_shr:8 row_mask:0xf bank_mask:0xf\n\t"
            "s_nop 1\n\t"
            "v_fmac_f32_dpp %0, %0, %1 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
            "s_nop 1\n\t"
            "v_mul_f32_dpp %1, %1, %1 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
            "s_nop 1\n\t"
            "v_fmac_f32_dpp %0, %0, %1 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
            "s_nop 1\n\t"
            "v_mul_f32_dpp %1, %1, %1 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
            "s_nop 1"
            : "+v"(add), "+v"(mul));
    }

    // Sums: lane k holds x_k = term + x_{k-1}. On return x_k = term + x_in (or term where flag is set).
    __device__ inline void sum_combine(double t_prev, int f_prev, double& term, int& flag)
    {
        double const t = term + t_prev;
        term = flag ? term : t;
        flag |= f_prev;
    }

    template<int N>
    __device__ inline void sum_step(double& term, int& flag)
    {
        double const t_prev = row_shr0<N>(term);
        int const f_prev = row_shr0<N>(flag);
        sum_combine(t_prev, f_prev, term, flag);
    }

    __device__ inline void sum_scan(double& term, int& flag)
    {
        sum_step<1>(term, flag);
        sum_step<2>(term, flag);
        sum_step<4>(term, flag);
        sum_step<8>(term, flag);
        {
            double const t_prev = row_bcast<dpp_row_bcast15, 0xA>(term, 0.0);
            int const f_prev = row_bcast<dpp_row_bcast15, 0xA>(flag, 0);
            sum_combine(t_prev, f_prev, term, flag);
        }
        {
            double const t_prev = row_bcast<dpp_row_bcast31, 0xC>(term, 0.0);
            int const f_prev = row_bcast<dpp_row_bcast31, 0xC>(flag, 0);
            sum_combine(t_prev, f_prev, term, flag);
        }
    }

    // Clamped sums: lane k holds x_k = ext(d, x_{k-1} + c), ext = max (Max) or min. A composition of such
    // maps is one of the same form: (d, c) after (d_prev, c_prev) is (ext(d, d_prev + c), c_prev + c). A NaN
    // in d or c gives NaN from there on (extremum_nan returns NaN for a NaN operand). On return
    // x_k = ext(d, x_in + c), or ext(d, c) where flag is set (a chain start has no upstream value).
    template<bool Max>
    __device__ inline double extremum_nan(double a, double b)
    {
        double const nan = __longlong_as_double(0x7ff8000000000000ll);
        if constexpr (Max)
            return a >= b ? a : (b > a ? b : nan);
        else
            return a <= b ? a : (b < a ? b : nan);
    }

    template<bool Max>
    __device__ inline void clamp_combine(double d_prev, double c_prev, int f_prev, double& d, double& c, int& flag)
    {
        double const dn = extremum_nan<Max>(d, d_prev + c);
        double const cn = c_prev + c;
        d = flag ? d : dn;
        c = flag ? c : cn;
        flag |= f_prev;
    }

    template<bool Max>
    constexpr double clamp_identity()
    {
        return Max ? -__builtin_huge_val() : __builtin_huge_val();
    }

    template<bool Max, int N>
    __device__ inline void clamp_step(double& d, double& c, int& flag)
    {
        double const d_prev = row_shr<N>(d, clamp_identity<Max>());
        double const c_prev = row_shr0<N>(c);
        int const f_prev = row_shr0<N>(flag);
        clamp_combine<Max>(d_prev, c_prev, f_prev, d, c, flag);
    }

    template<bool Max>
    __device__ inline void clamp_scan(double& d, double& c, int& flag)
    {
        clamp_step<Max, 1>(d, c, flag);
        clamp_step<Max, 2>(d, c, flag);
        clamp_step<Max, 4>(d, c, flag);
        clamp_step<Max, 8>(d, c, flag);
        {
            double const d_prev = row_bcast<dpp_row_bcast15, 0xA>(d, clamp_identity<Max>());
            double const c_prev = row_bcast<dpp_row_bcast15, 0xA>(c, 0.0);
            int const f_prev = row_bcast<dpp_row_bcast15, 0xA>(flag, 0);
            clamp_combine<Max>(d_prev, c_prev, f_prev, d, c, flag);
        }
        {
            double const d_prev = row_bcast<dpp_row_bcast31, 0xC>(d, clamp_identity<Max>());
            double const c_prev = row_bcast<dpp_row_bcast31, 0xC>(c, 0.0);
            int const f_prev = row_bcast<dpp_row_bcast31, 0xC>(flag, 0);
            clamp_combine<Max>(d_prev, c_prev, f_prev, d, c, flag);
        }
    }

    // Outcomes of kinematic_wave_block.
    enum KwBlockResult : int
    {
        kw_block_done = 0,    // q holds the block's outflows
        kw_block_serial = 1,  // a lane needs LUE's per-cell iteration or the block Newton failed: per-cell path
        kw_block_lue_far = 2  // converged, but LUE's own iteration would not end at the root in some lane
    };

    // Kinematic wave on one block of chain positions, all lanes at once. Newton's method on the coupled
    // system of the block
    //     F_k(Q) = c1_k (Q_k - Q_{k-1} - extra_k) + alpha_k Q_k^beta_k - b_k = 0,
    // Q_{k-1} being the previous position on the chain (`carry` before the first lane, nothing at a chain
    // start); per cell this is LUE's equation with Q_up = Q_{k-1} + extra_k. The Jacobian is lower
    // bidiagonal (d_k = c1_k + alpha_k beta_k Q_k^(beta_k - 1) on the diagonal, -c1_k below it), so a
    // Newton step dQ_k = (-F_k + c1_k dQ_{k-1}) / d_k is one affine scan (a chain start multiplies the value
    // before it by 0, so the scans need no start flags). The start is the system linearised at Q_cur, an
    // affine recurrence as well. F is concave in every Q_k and the Jacobian inverse-positive, so the iterates
    // approach the root from below after the first step.
    // T = double: FP64 steps until every |dQ_k| <= |Q_k| 2^-30, LUE's FP64 criterion (digits 31).
    // T = float: FP32 steps with the hardware power (log2, exp2) and reciprocal until every
    // |dQ_k| <= |Q_k| 2^-8, then one Newton step on the FP64 residual (the FP64 residual correction; its
    // scan in FP32). Newton's relative error obeys r' <= |beta - 1| / (2 (1 - r)) (1 + r)^|beta - 2| r^2, so
    // after the step that meets 2^-8 the error is below about 1.5e-5 (beta in [0.3, 3]) and after the FP64
    // step below about 3e-10, far below the FP32 output rounding (6e-8): the correction, not more FP32
    // steps, sets the accuracy.
    // Returns kw_block_serial, with q untouched, when a lane needs LUE's per-cell iteration: no-data, a zero
    // current discharge (dry cells and cells whose iteration starts from numeric_limits::min()), a non-positive
    // or non-finite iterate, or no convergence within 16 steps; kw_block_lue_far when LUE's own iteration
    // would stop short of the root in some lane (lue_guess_converges; q then holds the roots and far_mask
    // those lanes).
    template<typename T>
    __device__ inline int kinematic_wave_block(ChainLane<op_kinematic_wave, T> const& d, double carry, double dt,
                                                double& q, std::uint64_t& far_mask)
    {
        int const lane = threadIdx.x & 63;
        bool const valid = d.valid;
        bool const start = valid && d.start;
        double const q_cur = static_cast<double>(d.in.q_cur);
        if (__builtin_expect(__ballot(valid && !(q_cur > 0.0 && !isnan(d.extra))) != 0, 0))
            return kw_block_serial;
        double const c1 = kw_c1(d.in);
        double const alpha = static_cast<double>(d.in.alpha);
        double const beta = static_cast<double>(d.in.beta);
        double const b = kw_b(d.in);
        float const c1f = static_cast<float>(c1);
        float const beta32 = static_cast<float>(d.in.beta);
        float const ab32 = static_cast<float>(d.in.alpha) * beta32;
        // the value before lane k is multiplied by c1 (its upstream neighbour on the chain), except at a chain
        // start and past the block's positions
        bool const chained = valid && !start;

        // start: F linearised at Q_cur gives Q_k = w_k (Q_{k-1} + extra_k) + g_k, in FP32 (a starting point:
        // b - (1 - beta) alpha Q_cur^beta = beta alpha Q_cur^beta + dt q_lat does not cancel)
        double x;
        float pc1 = 1.0f;  // Q_cur^(beta-1)
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
            x = valid ? fma(static_cast<double>(mul), carry, static_cast<double>(add)) : 1.0;
        }
        if (__builtin_expect(__ballot(valid && !(x > 0.0 && x < __builtin_huge_val())) != 0, 0))
            return kw_block_serial;

        // Newton steps: T = float: FP32 residual, power (hardware log2/exp2), Jacobian and step scan; its
        // rounding (2^-24 of the largest term) moves the iterate by about 2^-24 of Q (the Jacobian is at least
        // c1), far below the stop criterion, and the FP64 residual correction below removes it. T = double:
        // all in FP64
        constexpr double tol = std::is_same_v<T, double> ? 0x1p-30 : 0x1p-8;
        bool converged = false;
        float const carry32 = static_cast<float>(carry);
        float const extra32 = static_cast<float>(d.extra);
        float const alpha32 = static_cast<float>(alpha);
        float const b32 = static_cast<float>(b);
        for (int iteration = 0; iteration < 16 && !converged; ++iteration)
        {
            double step;
            if constexpr (std::is_same_v<T, double>)
            {
                double prev = wave_shr1(x);
                if (lane == 0)
                    prev = carry;
                if (start)
                    prev = 0.0;
                double const pw1 = pow_positive<false>(x, beta - 1.0);  // x^(beta-1)
                double const f = c1 * (x - prev - d.extra) + alpha * x * pw1 - b;
                double const fp = c1 + alpha * beta * pw1;
                double inv = static_cast<double>(fast_rcp(static_cast<float>(fp)));
                inv = inv * fma(-fp, inv, 2.0);
                inv = inv * fma(-fp, inv, 2.0);
                double add = valid ? -f * inv : 0.0;
                double mul = chained ? c1 * inv : 0.0;
                affine_scan_nf(add, mul);
                step = add;  // the carry is exact: its step is 0
            }
            else
            {
                float const x32 = static_cast<float>(x);
                float prev32 = wave_shr1(x32);
                if (lane == 0)
                    prev32 = carry32;
                if (start)
                    prev32 = 0.0f;
                float const pw1 = fast_exp2((beta32 - 1.0f) * fast_log2(x32));
                float const f = c1f * (x32 - prev32 - extra32) + alpha32 * x32 * pw1 - b32;
                float const inv = fast_rcp(c1f + ab32 * pw1);
                float add = valid ? -f * inv : 0.0f;
                float mul = chained ? c1f * inv : 0.0f;
                affine_scan_f32_nf(add, mul);
                step = static_cast<double>(add);
            }
            double const xn = x + step;
            if (__builtin_expect(__ballot(valid && !(xn > 0.0 && xn < __builtin_huge_val())) != 0, 0))
                return kw_block_serial;
            converged = __ballot(valid && !(fabs(step) <= fabs(xn) * tol)) == 0;
            x = xn;
        }
        if (!converged)
            return kw_block_serial;
        if constexpr (std::is_same_v<T, float>)
        {
            // FP64 residual correction: one Newton step on the FP64 residual (the short FP64 power, relative
            // error about 1e-12). The step is below about 1e-6 of x, so its Jacobian and scan in FP32
            // (relative error 1e-7 of the step) leave x within about 1e-13 of the FP64 step's result
            double prev = wave_shr1(x);
            if (lane == 0)
                prev = carry;
            if (start)
                prev = 0.0;
            double const pw1 = pow_positive<true, true>(x, beta - 1.0);
            double const f = c1 * (x - prev - d.extra) + alpha * x * pw1 - b;
            float const inv = fast_rcp(c1f + ab32 * static_cast<float>(pw1));
            float add = valid ? static_cast<float>(-f) * inv : 0.0f;
            float mul = chained ? c1f * inv : 0.0f;
            affine_scan_f32_nf(add, mul);
            double const xn = x + static_cast<double>(add);
            if (__builtin_expect(__ballot(valid && !(xn > 0.0 && xn < __builtin_huge_val())) != 0, 0))
                return kw_block_serial;
            x = xn;
        }
        // LUE's own iteration ends at this root only from a guess near it (lue_guess_converges); otherwise
        // the block takes the per-cell path, which follows LUE's iteration there
        {
            double prev = wave_shr1(x);
            if (lane == 0)
                prev = carry;
            if (start)
                prev = 0.0;
            float dtq;  // dt q_lat; for T = float recovered from b with the hardware power (enough for a check)
            if constexpr (std::is_same_v<T, float>)
                dtq = static_cast<float>(b - alpha * q_cur * static_cast<double>(pc1));
            else
                dtq = static_cast<float>(dt * d.in.q_lat);
            bool const far = valid && !lue_guess_converges(static_cast<float>(prev + d.extra), static_cast<float>(q_cur),
                                                           dtq, static_cast<float>(alpha), beta32, c1f,
                                                           static_cast<float>(x));
            q = valid ? x : 0.0;
            far_mask = __ballot(far);
            if (__builtin_expect(far_mask != 0, 0))
                return kw_block_lue_far;
        }
        return kw_block_done;
    }

    // State carried from one block of chain positions to the next by the wavefront that runs them.
    struct BlockState
    {
        double carry{0.0};  // outflow of the previous chain position (FP64)
    };

    // Run the cell function along one block of up to 64 chain positions (lane j: position j; `nb` valid
    // lanes). Chains start where d.start is set; the first lane continues the chain of st.carry otherwise.
    // Linear functions are scans; the kinematic wave is a block Newton (kinematic_wave_block) with LUE's
    // per-cell iteration for blocks that need it; threshold, capacity and fraction are scans of their cell maps;
    // trigger (a step function) runs serially with lane broadcasts.
    // On return `value` (FP64 outflow) and `rem` hold this lane's results and st the state for the next
    // block.
    template<int OP, typename T>
    __device__ inline void solve_block(SweepArgs<T> const& a, ChainLane<OP, T> const& d, int nb, BlockState& st,
                                       double& value, T& rem)
    {
        int const lane = threadIdx.x & 63;
        double const nan64 = __longlong_as_double(0x7ff8000000000000ll);
        value = 0.0;
        if constexpr (OP == op_accu)
        {
            // linear: out_k = out_{k-1} + side_k + x_k, a segmented inclusive scan along the block
            double term = !d.valid ? 0.0 : (isnan(d.in.x) ? nan64 : d.extra + static_cast<double>(d.in.x));
            int flag = d.valid && d.start ? 1 : 0;
            sum_scan(term, flag);
            value = flag ? term : st.carry + term;
            st.carry = read_lane(value, nb - 1);
        }
        else if constexpr (OP == op_threshold || OP == op_capacity || OP == op_fraction)
        {
            // The outflow as a function of the outflow u of the previous chain position is a clamped sum
            // (threshold: max(0, u + extra + x - p), capacity: min(p, u + extra + x)) or an affine map
            // (fraction: p (u + extra + x)). Compositions keep the form, so one segmented scan gives every
            // lane its upstream outflow; the lane then runs the cell function on it (the reference's
            // formula from there on; the scan only reorders the sums that lead to the upstream value).
            int flag = d.valid && d.start ? 1 : 0;
            double const ex = d.extra + static_cast<double>(d.in.x);  // NaN (no-data, domain) propagates
            double const p = static_cast<double>(d.in.p);
            double x;
            if constexpr (OP == op_fraction)
            {
                double add = d.valid ? p * ex : 0.0;
                double mul = d.valid ? p : 1.0;
                affine_scan(add, mul, flag);
                x = flag ? add : fma(mul, st.carry, add);
            }
            else
            {
                constexpr bool is_max = OP == op_threshold;
                double dd = !d.valid ? clamp_identity<is_max>() : (is_max ? 0.0 : p);
                double c = !d.valid ? 0.0 : (is_max ? ex - p : ex);
                clamp_scan<is_max>(dd, c, flag);
                x = extremum_nan<is_max>(dd, (flag ? 0.0 : st.carry) + c);
            }
            double prev = wave_shr1(x);
            if (lane == 0)
                prev = st.carry;
            if (d.start)
                prev = 0.0;
            double const up = prev + d.extra;
            double out = 0.0;
            T r{};
            if (d.valid)
                solve_cell<OP, T>(a, d.in, up, isnan(up), out, r);
            value = out;
            rem = r;
            st.carry = read_lane(value, nb - 1);
        }
        else if constexpr (OP == op_trigger)
        {
            // a step function of the upstream outflow: serial, with lane broadcasts
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
        else
        {
            std::uint64_t far_mask = 0;
            int const outcome = kinematic_wave_block<T>(d, st.carry, static_cast<double>(a.dt), value, far_mask);
            if (__builtin_expect(outcome == kw_block_done, 1))
            {
                st.carry = read_lane(value, nb - 1);
            }
            else
            {
                // per-cell path (LUE's iteration on the FP64 path, new_discharge_f32_path on the FP32 path):
                // along the whole block, or, where only LUE's own iteration stops short of the root, from the
                // first such lane through the positions of its chain after it (the other lanes keep the roots,
                // which are LUE's results there)
                bool const all = outcome == kw_block_serial;
                int const j0 = all ? 0 : __builtin_ctzll(far_mask);
                double carry = j0 == 0 ? st.carry : read_lane(value, j0 - 1);
                [[maybe_unused]] float q_lat_mine = 0.0f;
                [[maybe_unused]] double b_mine = 0.0;
                if constexpr (std::is_same_v<T, float>)
                {
                    q_lat_mine = kw_q_lat(d.in, a.dt);
                    b_mine = kw_b(d.in);
                }
                bool affected = false;
                for (int j = j0; j < nb; ++j)
                {
                    int const first = read_lane(static_cast<int>(d.start), j);
                    affected = all || ((far_mask >> j) & 1u) != 0 || (affected && !first);
                    if (!affected)
                    {
                        carry = read_lane(value, j);
                        continue;
                    }
                    double const up = (first ? 0.0 : carry) + read_lane(d.extra, j);
                    double v;
                    if constexpr (std::is_same_v<T, double>)
                    {
                        CellInput<OP, T> cj{read_lane(d.in.q_cur, j), read_lane(d.in.q_lat, j),
                                            read_lane(d.in.alpha, j), read_lane(d.in.beta, j),
                                            read_lane(d.in.b, j), read_lane(d.in.c1, j)};
                        T r{};
                        solve_cell<OP, T>(a, cj, up, isnan(up), v, r);
                    }
                    else
                    {
                        float const q_cur = read_lane(d.in.q_cur, j);
                        if (isnan(up) || isnan(q_cur))
                            v = nan64;
                        else
                            v = new_discharge_f32_path_call(up, q_cur, read_lane(q_lat_mine, j),
                                                            read_lane(d.in.alpha, j), read_lane(d.in.beta, j), a.dt,
                                                            read_lane(b_mine, j), read_lane(kw_c1(d.in), j));
                    }
                    if (lane == j)
                        value = v;
                    carry = v;
                }
                st.carry = carry;
            }
        }
    }

    // Kinematic wave, FP32 path, one cell whose upstream sum `up` (FP64) is known: the per-cell form of
    // kinematic_wave_block, for a lane that walks a chain by itself. FP32 Newton steps with the hardware power
    // from the equation linearised at Q_cur until |dQ| <= |Q| 2^-8, one Newton step on the FP64 residual (the
    // short FP64 power), then the block's check that LUE's own iteration ends at this root. A cell that needs
    // LUE's iteration (zero current discharge, a non-positive or non-finite iterate, no convergence within 16
    // steps, a guess from which LUE's iteration stops short of the root) takes new_discharge_f32_path, as the
    // block's per-cell fallback does.
    __device__ inline double kinematic_wave_lane_f32(SweepArgs<float> const& a,
                                                     CellInput<op_kinematic_wave, float> const& in, double up)
    {
        float const q_cur = in.q_cur;
        if (isnan(up) || isnan(q_cur))
            return __longlong_as_double(0x7ff8000000000000ll);
        if (q_cur > 0.0f)
        {
            float const c1f = in.c1;
            float const alpha32 = in.alpha;
            float const beta32 = in.beta;
            float const ab32 = alpha32 * beta32;
            float const b32 = in.b_hi;
            float const up32 = static_cast<float>(up);
            float const pc1 = fast_exp2((beta32 - 1.0f) * fast_log2(q_cur));  // Q_cur^(beta-1)
            float x = (c1f * up32 + (b32 - (1.0f - beta32) * alpha32 * q_cur * pc1)) * fast_rcp(c1f + ab32 * pc1);
            bool ok = x > 0.0f && x < __builtin_huge_valf();
            bool converged = false;
            for (int iteration = 0; ok && iteration < 16; ++iteration)
            {
                float const pw1 = fast_exp2((beta32 - 1.0f) * fast_log2(x));
                float const f = c1f * (x - up32) + alpha32 * x * pw1 - b32;
                float const step = -f * fast_rcp(c1f + ab32 * pw1);
                float const xn = x + step;
                ok = xn > 0.0f && xn < __builtin_huge_valf();
                converged = fabsf(step) <= fabsf(xn) * 0x1p-8f;
                x = xn;
                if (converged)
                    break;
            }
            if (ok && converged)
            {
                // FP64 residual correction (its Jacobian in FP32, as in the block)
                double const b = kw_b(in);
                double const c1 = kw_c1(in);
                double const alpha = static_cast<double>(alpha32);
                double const xd = static_cast<double>(x);
                double const pw1 = pow_positive<true, true>(xd, static_cast<double>(beta32) - 1.0);
                double const f = c1 * (xd - up) + alpha * xd * pw1 - b;
                float const inv = fast_rcp(c1f + ab32 * static_cast<float>(pw1));
                double const xn = xd + static_cast<double>(static_cast<float>(-f) * inv);
                if (xn > 0.0 && xn < __builtin_huge_val())
                {
                    float const dtq = static_cast<float>(b - alpha * static_cast<double>(q_cur) * static_cast<double>(pc1));
                    if (lue_guess_converges(up32, q_cur, dtq, alpha32, beta32, c1f, static_cast<float>(xn)))
                        return xn;
                }
            }
        }
        return new_discharge_f32_path_call(up, q_cur, kw_q_lat(in, a.dt), in.alpha, in.beta, a.dt, kw_b(in), kw_c1(in));
    }

    // The cell function of one chain position whose upstream outflow `up` (FP64) is known, for a lane that walks
    // a chain by itself: the operations' cell functions (solve_cell), the FP32 kinematic wave by
    // kinematic_wave_lane_f32 (the same root as the block Newton). NaN upstream poisons the cell.
    template<int OP, typename T>
    __device__ inline void solve_lane(SweepArgs<T> const& a, CellInput<OP, T> const& in, double up, double& out, T& rem)
    {
        if constexpr (OP == op_kinematic_wave && std::is_same_v<T, float>)
            out = kinematic_wave_lane_f32(a, in, up);
        else
            solve_cell<OP, T>(a, in, up, isnan(up), out, rem);
    }

}  // namespace lue_amd
