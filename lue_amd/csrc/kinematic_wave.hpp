// Per-cell kinematic-wave solve, device side.
//
// Equation (reference: source/framework/algorithm/include/lue/framework/algorithm/definition/
// kinematic_wave.hpp:13-142): with c1 = dt / dx,
//     f(Q) = c1 Q + alpha Q^beta - K = 0,     K = c1 Q_up + alpha Q_cur^beta + dt q,
//     f'(Q) = c1 + alpha beta Q^(beta - 1),
// initial guess linearised at (Q_cur + Q_up) / 2 and clamped to numeric_limits<T>::min() (:61-91).
// LUE solves it with boost::math::tools::newton_raphson_iterate, digits = floor(0.6 digits(T)),
// at most 10 iterations (:195-213); dry cells (no upstream, current or lateral water) give 0 (:181-250).
//
// Two device solvers:
//   - solve_bracketed<T>: Newton-Raphson inside a shrinking bracket [0, max] with Boost's documented
//     step rules (an out-of-bracket step halves the distance to the bound, a step that does not
//     converge becomes a bounded jump, stop when |step| <= |Q| 2^(1-digits)). This is the plain FP64
//     drop-in, and the first stage of the FP32 path.
//   - the FP32 path then applies ONE Newton correction in FP64 to the FP32 root, with K formed in FP64
//     from the compensated upstream sum; this is the only FP64 arithmetic on that path.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace lue_amd {

    template<typename T>
    struct NumericLimits;

    template<>
    struct NumericLimits<float>
    {
        __host__ __device__ static constexpr float min() { return FLT_MIN; }
        __host__ __device__ static constexpr float max() { return FLT_MAX; }
        static constexpr int digits = 24;
    };

    template<>
    struct NumericLimits<double>
    {
        __host__ __device__ static constexpr double min() { return DBL_MIN; }
        __host__ __device__ static constexpr double max() { return DBL_MAX; }
        static constexpr int digits = 53;
    };

    __device__ inline float dpow(float x, float y)
    {
        return powf(x, y);
    }

    // x > 0 here: exp(y log x) avoids pow's special-case handling
    __device__ inline double dpow(double x, double y)
    {
        return exp(y * log(x));
    }

    // a * b + c with c in scalar registers: VOP3 v_fma_f64 takes one SGPR operand, where the compiler's own
    // choice (v_fmac_f64, accumulator = addend) first copies every coefficient into vector registers.
    __host__ __device__ inline double fma_scalar_addend(double a, double b, double c)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        double r;
        asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
        return r;
#else
        return std::fma(a, b, c);
#endif
    }

    // x^y for finite x > 0 and |y ln x| < 700.
    //   ln x = e ln 2 + 2 atanh(s), s = (m - 1) / (m + 1), m in [sqrt(1/2), sqrt(2)), |s| < 0.1716;
    //   e^z = 2^k e^r, k = rint(z / ln 2), |r| <= ln 2 / 2 (ln 2 in two parts, double-double).
    // Full (Short = false): the atanh series to s^15 leaves 2 |s|^17 / 17 < 1.1e-14, Taylor to r^12 leaves
    // |r|^13 / 13! < 2e-16: relative error below 2^-52 |y ln x| + 1e-15. Used by the FP64 path.
    // Short: s from the FP32 hardware reciprocal of m + 1 refined by one Newton step (relative 2^-46, no FP64
    // division), the atanh series to s^13 (2 |s|^15 / 15 < 4.4e-13) and Taylor to r^10 (|r|^11 / 11! <
    // 2.2e-13): relative error below 4.4e-13 |y| + 2.2e-13 + 2^-52 |y ln x|, which moves a kinematic-wave
    // root by at most that over beta, far below the FP32 path's output rounding (6e-8). Used by the FP32
    // path's b and its FP64 residual correction.
    // The coefficients live in constant memory, read with scalar loads in every call: as literals the
    // compiler keeps all of them in vector registers across the kernels' loops (and spills them there).
    // [0, 7]: 1/15, 1/13, ..., 1/3, 1 (atanh series); [8]: ln2_hi, [9]: ln2_lo (ln 2 = ln2_hi + ln2_lo,
    // double-double; the fused products below need no exact k ln2_hi), [10]: 1/ln 2; [11, 23]: 1/12!, ...,
    // 1/2!, 1, 1 (Taylor series of e^r).
    inline __constant__ double pow_positive_coefficients[24] = {
        1.0 / 15.0, 1.0 / 13.0, 1.0 / 11.0, 1.0 / 9.0, 1.0 / 7.0, 1.0 / 5.0, 1.0 / 3.0, 1.0,
        0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56, 1.44269504088896340736,
        1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0,
        1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0, 0.5, 1.0, 1.0};

    // Reload: read the table in every call (an opaque copy of its address keeps the scalar loads in place);
    // otherwise the compiler hoists the loads out of the kernels' loops into scalar registers, which pays
    // where the power runs in a dependent chain of Newton steps and the wave has registers to spare.
    template<bool Reload = true, bool Short = false>
    __host__ __device__ inline double pow_positive(double x, double y)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        auto k = (__attribute__((address_space(4))) double const*)pow_positive_coefficients;
        if constexpr (Reload)
            asm volatile("" : "+s"(k));
#else
        static constexpr double k[24] = {
            1.0 / 15.0, 1.0 / 13.0, 1.0 / 11.0, 1.0 / 9.0, 1.0 / 7.0, 1.0 / 5.0, 1.0 / 3.0, 1.0,
            0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56, 1.44269504088896340736,
            1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0,
            1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0, 1.0 / 6.0, 0.5, 1.0, 1.0};
#endif
        int e;
        double m = frexp(x, &e);  // [0.5, 1)
        if (m < 0.70710678118654752440)
        {
            m *= 2.0;
            e -= 1;
        }
        double s;
        if constexpr (Short)
        {
            double const t = m + 1.0;
#if defined(__HIP_DEVICE_COMPILE__)
            double r = static_cast<double>(__builtin_amdgcn_rcpf(static_cast<float>(t)));
#else
            double r = static_cast<double>(1.0f / static_cast<float>(t));
#endif
            r = fma(r, fma(-t, r, 1.0), r);
            s = (m - 1.0) * r;
        }
        else
        {
            s = (m - 1.0) / (m + 1.0);
        }
        double const s2 = s * s;
        constexpr int atanh_first = Short ? 1 : 0;
        double p = k[atanh_first];
        for (int i = atanh_first + 1; i < 8; ++i)
            p = fma_scalar_addend(p, s2, k[i]);
        double const ln_x = fma(static_cast<double>(e), k[8], fma(static_cast<double>(e), k[9], 2.0 * s * p));
        double const z = y * ln_x;
        double const n = rint(z * k[10]);
        double const r = fma(-n, k[9], fma(-n, k[8], z));
        constexpr int taylor_first = Short ? 13 : 11;
        double q = k[taylor_first];
        for (int i = taylor_first + 1; i < 24; ++i)
            q = fma_scalar_addend(q, r, k[i]);
        return ldexp(q, static_cast<int>(n));
    }

    // Coefficients of one cell's equation.
    template<typename T>
    struct KinematicWaveCell
    {
        T q_up;       // summed new discharge of the upstream cells
        T q_cur;      // current discharge
        T q_lat;      // lateral inflow, >= 0 (value policies make negative inflow no-data)
        T alpha;
        T beta;
        T dt;
        T c1;         // dt / dx
        T b;          // alpha q_cur^beta + dt q_lat, evaluated once per cell outside the dependency chain
    };

    // Solve f(Q) = 0 from the linearised guess. Returns the root; `converged` tells whether the
    // iteration stopped on the step criterion (true) or ran out of iterations (false).
    template<typename T>
    __device__ inline T solve_bracketed(KinematicWaveCell<T> const& cell, bool& converged)
    {
        T const c1 = cell.c1;
        T const known = (c1 * cell.q_up) + cell.b;
        T const alpha_beta = cell.alpha * cell.beta;

        // guess (kinematic_wave.hpp:61-91)
        T x = NumericLimits<T>::min();
        if ((cell.q_cur + cell.q_up != T{0}) || cell.beta >= T{1})
        {
            T const a_b_pq = alpha_beta * dpow((cell.q_cur + cell.q_up) / T{2}, cell.beta - T{1});
            x = (c1 * cell.q_up + a_b_pq * cell.q_cur + cell.dt * cell.q_lat) / (c1 + a_b_pq);
            x = fmax(x, NumericLimits<T>::min());
        }

        int const digits = static_cast<int>(NumericLimits<T>::digits * 0.6);
        T const factor = ldexp(T{1}, 1 - digits);
        T lo = T{0};
        T hi = NumericLimits<T>::max();
        T step = NumericLimits<T>::max();
        T step_prev = step;
        T step_prev2 = step;
        int left = 10;
        converged = false;

        while (true)
        {
            step_prev2 = step_prev;
            step_prev = step;
            T const xb = dpow(x, cell.beta);
            T const f0 = (c1 * x) + (cell.alpha * xb) - known;
            T const f1 = c1 + (alpha_beta * xb / x);
            --left;
            if (f0 == T{0})
            {
                converged = true;
                break;
            }
            step = f0 / f1;
            if (fabs(step * 2) > fabs(step_prev2))
            {
                T const shift = (step > 0) ? (x - lo) / 2 : (x - hi) / 2;
                if ((x != 0) && (fabs(shift) > fabs(x)))
                {
                    step = copysign(fabs(x) * T(1.1f), step);
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
            if (!(fabs(x * factor) < fabs(step)))
            {
                converged = true;
                break;
            }
            if (left == 0)
            {
                break;
            }
        }
        return x;
    }

    // Plain FP64 drop-in: iterate_to_new_discharge with lateral inflow >= 0 (kinematic_wave.hpp:145-254).
    __device__ inline double new_discharge_f64(KinematicWaveCell<double> const& cell)
    {
        if (cell.q_up + cell.q_cur > 0 || cell.q_lat > 0)
        {
            bool converged;
            return solve_bracketed<double>(cell, converged);
        }
        return 0.0;
    }

    // Hardware approximations (about 1 ulp): v_log_f32, v_exp_f32, v_rcp_f32.
    __device__ inline float fast_log2(float x)
    {
        return __builtin_amdgcn_logf(x);
    }

    __device__ inline float fast_exp2(float x)
    {
        return __builtin_amdgcn_exp2f(x);
    }

    __device__ inline float fast_rcp(float x)
    {
        return __builtin_amdgcn_rcpf(x);
    }

    // FP32 Newton from LUE's linearised guess, with the hardware transcendentals; stops on LUE's FP32
    // criterion |step| <= |Q| 2^-13 (digits = floor(0.6 * 24) = 14). Needs q_up + q_cur > 0.
    __device__ inline float solve_fast_f32(
        float c1, float alpha, float beta, float dt, float q_up, float q_cur, float q_lat, float known, bool& converged)
    {
        float const ab = alpha * beta;
        float const a_b_pq = ab * fast_exp2((beta - 1.0f) * fast_log2(0.5f * (q_cur + q_up)));
        float x = (c1 * q_up + a_b_pq * q_cur + dt * q_lat) * fast_rcp(c1 + a_b_pq);
        x = fmaxf(x, FLT_MIN);
        converged = false;
        for (int iteration = 0; iteration < 10; ++iteration)
        {
            float const xb = fast_exp2(beta * fast_log2(x));
            float const f = fmaf(c1, x, fmaf(alpha, xb, -known));
            float const fp = fmaf(ab * xb, fast_rcp(x), c1);
            float const step = f * fast_rcp(fp);
            float xn = x - step;
            xn = xn > 0.0f ? xn : 0.5f * x;
            bool const done = fabsf(step) <= fabsf(xn) * 0x1p-13f;
            x = xn;
            if (done)
            {
                converged = true;
                break;
            }
        }
        return x;
    }

    // LUE's FP64 iteration of one cell, out of line: the rarely taken paths that need it stay out of the
    // register budget of the kernels' main paths.
    __device__ __noinline__ inline double lue_iteration_f64(double q_up, double q_cur, double q_lat, double alpha,
                                                            double beta, double dt, double c1, double b)
    {
        KinematicWaveCell<double> cell{q_up, q_cur, q_lat, alpha, beta, dt, c1, b};
        bool converged;
        return solve_bracketed<double>(cell, converged);
    }

    // Whether LUE's iteration surely ends at the root `root` (alpha, beta > 0): its guess, linearised at
    // (Q_cur + Q_up) / 2 (kinematic_wave.hpp:61-91), lies within theta root, theta = 1/2 / max(1, (beta - 1)^2).
    // Newton's error on f obeys e' = f''(xi) / (2 f'(x)) e^2 with xi between the iterate x and the root, and
    // with x = (1 - r) root (or (1 + r) root) that gives r' <= |beta - 1| / (2 (1 - r)) (1 + r)^|beta - 2| r^2;
    // from r <= theta the relative step falls below LUE's 2^-30 within 7 of its 10 steps. Elsewhere (a guess
    // far from the root, typically a tiny Q_cur under a large inflow with beta well below 1) LUE returns
    // whatever its 10 steps reach, which only its own iteration reproduces. dtq = dt q_lat. FP32 arithmetic
    // with the hardware power: a rounding here can only send a cell to LUE's own iteration or spare it that,
    // and both give LUE's result.
    __device__ inline bool lue_guess_converges(float q_up, float q_cur, float dtq, float alpha, float beta, float c1,
                                               float root)
    {
        float const mean = 0.5f * (q_cur + q_up);
        if (!(mean >= FLT_MIN && mean <= FLT_MAX) || !(root > 0.0f) || !(alpha > 0.0f) || !(beta > 0.0f))
            return false;
        float const a_b_pq = alpha * beta * fast_exp2((beta - 1.0f) * fast_log2(mean));
        float const x0 = fmaxf((c1 * q_up + a_b_pq * q_cur + dtq) * fast_rcp(c1 + a_b_pq), FLT_MIN);
        float const d = fmaxf(1.0f, (beta - 1.0f) * (beta - 1.0f));
        return fabsf(x0 - root) * d <= 0.5f * root;
    }

    // FP32 path of one cell. q_up is the compensated upstream sum (FP64), b = alpha q_cur^beta + dt q_lat
    // and c1 = dt / dx are evaluated in FP64 outside the dependency chain. The FP32 root gets one FP64
    // Newton correction, the only FP64 arithmetic on the chain.
    // Dry cells that receive water with nothing upstream (q_up + q_cur = 0, beta < 1) start from
    // numeric_limits::min() in LUE, where the FP32 and FP64 iterations behave differently; those take the
    // FP64 iteration so that results follow LUE's FP64 arithmetic.
    __device__ inline double new_discharge_f32_path(
        double q_up, float q_cur, float q_lat, float alpha, float beta, float dt, double b, double c1)
    {
        if (!(q_up + q_cur > 0 || q_lat > 0))
        {
            return 0.0;
        }
        if (q_up + q_cur == 0.0 && beta < 1.0f)
            return lue_iteration_f64(q_up, q_cur, q_lat, alpha, beta, dt, c1, b);
        double const known = c1 * q_up + b;
        bool converged;
        float const x32 = solve_fast_f32(static_cast<float>(c1), alpha, beta, dt, static_cast<float>(q_up), q_cur,
                                         q_lat, static_cast<float>(known), converged);
        double const x = static_cast<double>(x32);
        double const xb = dpow(x, static_cast<double>(beta));
        double const ax = static_cast<double>(alpha) * xb;
        double const f0 = (c1 * x) + ax - known;
        double const f1 = c1 + static_cast<double>(beta) * ax / x;
        double corrected = x - f0 / f1;
        corrected = corrected > 0.0 ? corrected : x;
        if (!lue_guess_converges(static_cast<float>(q_up), q_cur, dt * q_lat, alpha, beta, static_cast<float>(c1),
                                 static_cast<float>(corrected)))
            return lue_iteration_f64(q_up, q_cur, q_lat, alpha, beta, dt, c1, b);  // its result need not be the root
        return corrected;
    }

    // new_discharge_f32_path out of line, for the rarely taken per-cell redo of a block.
    __device__ __noinline__ inline double new_discharge_f32_path_call(double q_up, float q_cur, float q_lat, float alpha,
                                                                      float beta, float dt, double b, double c1)
    {
        return new_discharge_f32_path(q_up, q_cur, q_lat, alpha, beta, dt, b, c1);
    }


    // Elementwise FP32 solve with the corrected root, used by iterate_to_new_discharge_f32.
    __device__ inline double new_discharge_f32_corrected(
        double q_up, float q_cur, float q_lat, float alpha, float beta, float dt, float dx, double alpha_pow)
    {
        double const c1 = static_cast<double>(dt) / static_cast<double>(dx);
        double const b = alpha_pow + static_cast<double>(dt) * static_cast<double>(q_lat);
        return new_discharge_f32_path(q_up, q_cur, q_lat, alpha, beta, dt, b, c1);
    }

}  // namespace lue_amd
