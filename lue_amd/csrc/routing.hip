// C ABI of the flow-routing path (include/lue_amd/routing.h): argument checks, device scratch, the
// round loop of the tiled sweep, and the synthetic-input generators.
#include "lue_amd/routing.h"

#include "inflow_count.hpp"
#include "kinematic_wave.hpp"
#include "routing_common.hpp"
#include "sweep.hpp"
#include "synthetic.hpp"
#include "tail.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace lue_amd {
    namespace {

        thread_local std::string last_error;

        // Optional per-call timing of the sweep launches with HIP events on the caller's stream.
        struct Profile
        {
            bool enabled{false};
            double first_round_ms{0.0};
            double next_rounds_ms{0.0};
            long long calls{0};
            long long rounds{0};
            long long cells{0};
        };
        Profile profile;

        // When the round loop hands over to the chain-streaming tail: after at least `min_rounds`
        // rounds, once at most max(tiles * tile_fraction, 4) tiles are queued, and always after
        // `max_rounds` rounds. min_rounds < 0 disables the tail (rounds only).
        struct TailPolicy
        {
            int min_rounds{4};
            int max_rounds{24};
            double tile_fraction{1.0 / 64.0};
        };
        TailPolicy tail_policy;

        int fail(int status, std::string message)
        {
            last_error = std::move(message);
            return status;
        }

        int hip_fail(hipError_t e, char const* what)
        {
            int const status = e == hipErrorOutOfMemory ? status_out_of_memory : status_hip_error;
            return fail(status, std::string(what) + ": " + hipGetErrorString(e));
        }

#define LUE_AMD_HIP(call)                                                                                       \
    do                                                                                                          \
    {                                                                                                           \
        hipError_t const e_ = (call);                                                                           \
        if (e_ != hipSuccess)                                                                                   \
            return hip_fail(e_, #call);                                                                         \
    } while (0)

        int check_shape(void const* ldd, std::int64_t rows, std::int64_t cols)
        {
            if (!ldd)
                return fail(status_invalid_argument, "ldd must not be null");
            if (rows <= 0 || cols <= 0)
                return fail(status_invalid_argument, "array shape must not be empty");
            if (rows > (1ll << 31) || cols > (1ll << 31))
                return fail(status_invalid_argument, "array extent too large");
            return status_success;
        }

        // Scratch of one sweep, allocated stream-ordered.
        struct SweepScratch
        {
            std::uint32_t* pending_bits{};
            std::uint32_t* border_stamp{};
            std::uint32_t* tile_flag{};
            std::uint32_t* worklist{};
            std::uint32_t* work_count{};
            unsigned long long* counters{};
            float* out_lo{};
            hipStream_t stream{};

            ~SweepScratch()
            {
                for (void* p : {static_cast<void*>(pending_bits), static_cast<void*>(border_stamp),
                                static_cast<void*>(tile_flag), static_cast<void*>(worklist),
                                static_cast<void*>(work_count), static_cast<void*>(counters),
                                static_cast<void*>(out_lo)})
                {
                    if (p)
                        (void)hipFreeAsync(p, stream);
                }
            }
        };

        template<typename P>
        hipError_t alloc(P*& p, std::size_t count, hipStream_t stream)
        {
            void* raw = nullptr;
            hipError_t const e = hipMallocAsync(&raw, std::max<std::size_t>(count, 1) * sizeof(P), stream);
            p = static_cast<P*>(raw);
            return e;
        }


        template<typename P>
        struct DeviceArray
        {
            P* ptr{};
            hipStream_t stream{};
            ~DeviceArray()
            {
                if (ptr)
                    (void)hipFreeAsync(ptr, stream);
            }
        };

        template<typename P>
        hipError_t alloc_array(DeviceArray<P>& a, std::size_t count, hipStream_t stream)
        {
            a.stream = stream;
            return alloc(a.ptr, count, stream);
        }

        template<typename In, typename Out>
        hipError_t exclusive_scan(In const* in, Out* out, std::size_t n, hipStream_t stream)
        {
            std::size_t bytes = 0;
            hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, stream);
            if (e != hipSuccess)
                return e;
            void* temp = nullptr;
            e = hipMallocAsync(&temp, std::max<std::size_t>(bytes, 1), stream);
            if (e != hipSuccess)
                return e;
            e = hipcub::DeviceScan::ExclusiveSum(temp, bytes, in, out, n, stream);
            hipError_t const f = hipFreeAsync(temp, stream);
            return e != hipSuccess ? e : f;
        }

        inline unsigned blocks_for(std::uint64_t n, unsigned threads = 256)
        {
            return static_cast<unsigned>(std::max<std::uint64_t>((n + threads - 1) / threads, 1));
        }

        // Solve all pending cells by chain streaming (tail.hpp). Adds the number of solved cells to *solved.
        template<int OP, typename T, int TW, int TH>
        int run_tail(SweepArgs<T> const& a, hipStream_t stream, unsigned long long* solved_total)
        {
            std::int64_t const ntiles = static_cast<std::int64_t>(a.tiles_x) * a.tiles_y;
            std::uint64_t const ncells = static_cast<std::uint64_t>(a.rows * a.cols);
            if (ncells >= (1ull << 32))
                return fail(status_invalid_argument, "chain tail needs fewer than 2^32 cells");

            DeviceArray<std::uint32_t> counts, offsets;
            LUE_AMD_HIP(alloc_array(counts, ntiles + 1, stream));
            LUE_AMD_HIP(alloc_array(offsets, ntiles + 1, stream));
            LUE_AMD_HIP(hipMemsetAsync(counts.ptr + ntiles, 0, sizeof(std::uint32_t), stream));
            hipLaunchKernelGGL((tail_count_kernel<TW, TH>), dim3(blocks_for(ntiles)), dim3(256), 0, stream,
                               a.pending_bits, ntiles, counts.ptr);
            LUE_AMD_HIP(hipGetLastError());
            LUE_AMD_HIP(exclusive_scan(counts.ptr, offsets.ptr, static_cast<std::size_t>(ntiles + 1), stream));
            std::uint32_t n = 0;
            LUE_AMD_HIP(hipMemcpyAsync(&n, offsets.ptr + ntiles, sizeof(n), hipMemcpyDeviceToHost, stream));
            LUE_AMD_HIP(hipStreamSynchronize(stream));
            if (n == 0)
                return status_success;

            DeviceArray<std::uint32_t> cells, index_of, is_last, chain_of_last;
            DeviceArray<std::int32_t> next, upstream_count, succ[2], rank[2], last[2];
            LUE_AMD_HIP(alloc_array(cells, n, stream));
            LUE_AMD_HIP(alloc_array(index_of, ncells, stream));
            LUE_AMD_HIP(alloc_array(next, n, stream));
            LUE_AMD_HIP(alloc_array(upstream_count, n, stream));
            for (int b = 0; b < 2; ++b)
            {
                LUE_AMD_HIP(alloc_array(succ[b], n, stream));
                LUE_AMD_HIP(alloc_array(rank[b], n, stream));
                LUE_AMD_HIP(alloc_array(last[b], n, stream));
            }
            LUE_AMD_HIP(alloc_array(is_last, n + 1, stream));
            LUE_AMD_HIP(alloc_array(chain_of_last, n + 1, stream));
            LUE_AMD_HIP(hipMemsetAsync(upstream_count.ptr, 0, n * sizeof(std::int32_t), stream));
            LUE_AMD_HIP(hipMemsetAsync(is_last.ptr + n, 0, sizeof(std::uint32_t), stream));

            hipLaunchKernelGGL((tail_compact_kernel<TW, TH>), dim3(static_cast<unsigned>(ntiles)), dim3(64), 0, stream,
                               a.pending_bits, a.tiles_x, a.cols, offsets.ptr, cells.ptr, index_of.ptr);
            hipLaunchKernelGGL((tail_links_kernel<TW, TH>), dim3(blocks_for(n)), dim3(256), 0, stream, a.ldd, a.rows,
                               a.cols, a.tiles_x, a.pending_bits, cells.ptr, index_of.ptr, n, next.ptr,
                               upstream_count.ptr);
            hipLaunchKernelGGL(tail_rank_init_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, next.ptr,
                               upstream_count.ptr, n, succ[0].ptr, rank[0].ptr, last[0].ptr);
            int cur = 0;
            for (std::uint32_t reach = 1; reach < n; reach <<= 1)
            {
                hipLaunchKernelGGL(tail_rank_step_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, succ[cur].ptr,
                                   rank[cur].ptr, last[cur].ptr, n, succ[1 - cur].ptr, rank[1 - cur].ptr,
                                   last[1 - cur].ptr);
                cur = 1 - cur;
            }
            hipLaunchKernelGGL(tail_flag_last_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, next.ptr,
                               upstream_count.ptr, n, is_last.ptr);
            LUE_AMD_HIP(hipGetLastError());
            LUE_AMD_HIP(exclusive_scan(is_last.ptr, chain_of_last.ptr, static_cast<std::size_t>(n + 1), stream));
            std::uint32_t nchains = 0;
            LUE_AMD_HIP(hipMemcpyAsync(&nchains, chain_of_last.ptr + n, sizeof(nchains), hipMemcpyDeviceToHost, stream));
            LUE_AMD_HIP(hipStreamSynchronize(stream));
            if (nchains == 0)
                return fail(status_invalid_network, "invalid flow direction network: pending cells form cycles");

            DeviceArray<std::uint32_t> length, base, cursor_and_count, chain_cells;
            DeviceArray<std::int32_t> first, arrivals, downstream, ready;
            DeviceArray<unsigned long long> solved;
            LUE_AMD_HIP(alloc_array(length, nchains + 1, stream));
            LUE_AMD_HIP(alloc_array(base, nchains + 1, stream));
            LUE_AMD_HIP(alloc_array(first, nchains, stream));
            LUE_AMD_HIP(alloc_array(arrivals, nchains, stream));
            LUE_AMD_HIP(alloc_array(downstream, nchains, stream));
            LUE_AMD_HIP(alloc_array(ready, nchains, stream));
            LUE_AMD_HIP(alloc_array(cursor_and_count, 2, stream));
            LUE_AMD_HIP(alloc_array(chain_cells, n, stream));
            LUE_AMD_HIP(alloc_array(solved, 1, stream));
            // a chain whose first cell never appears (cells on a cycle) keeps length 0 and is never ready
            LUE_AMD_HIP(hipMemsetAsync(length.ptr, 0, (nchains + 1) * sizeof(std::uint32_t), stream));
            LUE_AMD_HIP(hipMemsetAsync(arrivals.ptr, 0xFF, nchains * sizeof(std::int32_t), stream));
            LUE_AMD_HIP(hipMemsetAsync(cursor_and_count.ptr, 0, 2 * sizeof(std::uint32_t), stream));
            LUE_AMD_HIP(hipMemsetAsync(solved.ptr, 0, sizeof(unsigned long long), stream));

            hipLaunchKernelGGL(tail_chain_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, next.ptr,
                               upstream_count.ptr, rank[cur].ptr, last[cur].ptr, chain_of_last.ptr, n, length.ptr,
                               first.ptr, arrivals.ptr);
            LUE_AMD_HIP(exclusive_scan(length.ptr, base.ptr, static_cast<std::size_t>(nchains + 1), stream));
            hipLaunchKernelGGL(tail_downstream_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, next.ptr,
                               last[cur].ptr, chain_of_last.ptr, is_last.ptr, n, downstream.ptr);
            hipLaunchKernelGGL(tail_layout_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, cells.ptr, rank[cur].ptr,
                               last[cur].ptr, is_last.ptr, chain_of_last.ptr, base.ptr, length.ptr, n,
                               chain_cells.ptr);
            hipLaunchKernelGGL(tail_ready_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream, arrivals.ptr,
                               nchains, ready.ptr, cursor_and_count.ptr + 1);
            LUE_AMD_HIP(hipGetLastError());
            std::uint32_t nready = 0;
            LUE_AMD_HIP(hipMemcpyAsync(&nready, cursor_and_count.ptr + 1, sizeof(nready), hipMemcpyDeviceToHost, stream));
            LUE_AMD_HIP(hipStreamSynchronize(stream));

            TailChains t{chain_cells.ptr, base.ptr,  length.ptr,           downstream.ptr, arrivals.ptr,
                         ready.ptr,       nready,    cursor_and_count.ptr, solved.ptr};
            // one wavefront per chain; 4 wavefronts per workgroup, enough workgroups to fill the device
            unsigned const grid = static_cast<unsigned>(std::min<std::uint64_t>((nready + 3) / 4, 2048));
            if (nready > 0)
            {
                hipLaunchKernelGGL((tail_execute_kernel<OP, T>), dim3(std::max(grid, 1u)), dim3(256), 0, stream, a, t);
                LUE_AMD_HIP(hipGetLastError());
            }
            unsigned long long tail_solved = 0;
            LUE_AMD_HIP(hipMemcpyAsync(&tail_solved, solved.ptr, sizeof(tail_solved), hipMemcpyDeviceToHost, stream));
            LUE_AMD_HIP(hipStreamSynchronize(stream));
            *solved_total += tail_solved;
            return status_success;
        }

        // Run the tiled sweep to completion: round 0 over all tiles, then queued tiles until none is left.
        template<int OP, typename T, int TW, int TH>
        int run_sweep(SweepArgs<T> a, hipStream_t stream)
        {
            using G = Tile<TW, TH>;
            a.tiles_x = static_cast<int>((a.cols + TW - 1) / TW);
            a.tiles_y = static_cast<int>((a.rows + TH - 1) / TH);
            std::int64_t const ntiles = static_cast<std::int64_t>(a.tiles_x) * a.tiles_y;
            if (ntiles > (1ll << 31) - 1)
                return fail(status_invalid_argument, "too many tiles");

            SweepScratch s;
            s.stream = stream;
            LUE_AMD_HIP(alloc(s.pending_bits, ntiles * G::words, stream));
            LUE_AMD_HIP(alloc(s.border_stamp, ntiles * G::perimeter, stream));
            LUE_AMD_HIP(alloc(s.tile_flag, 3 * ntiles, stream));
            LUE_AMD_HIP(alloc(s.worklist, 3 * ntiles, stream));
            LUE_AMD_HIP(alloc(s.work_count, 4, stream));
            LUE_AMD_HIP(alloc(s.counters, 2, stream));
            if constexpr (std::is_same_v<T, float>)
            {
                LUE_AMD_HIP(alloc(s.out_lo, static_cast<std::size_t>(a.rows * a.cols), stream));
            }
            LUE_AMD_HIP(hipMemsetAsync(s.tile_flag, 0, 3 * ntiles * sizeof(std::uint32_t), stream));
            LUE_AMD_HIP(hipMemsetAsync(s.work_count, 0, 4 * sizeof(std::uint32_t), stream));
            LUE_AMD_HIP(hipMemsetAsync(s.counters, 0, 2 * sizeof(unsigned long long), stream));

            a.pending_bits = s.pending_bits;
            a.border_stamp = s.border_stamp;
            a.tile_flag = s.tile_flag;
            a.worklist = s.worklist;
            a.work_count = s.work_count;
            a.counters = s.counters;
            a.out_lo = s.out_lo;

            hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
            if (profile.enabled)
            {
                for (auto& e : ev)
                    LUE_AMD_HIP(hipEventCreate(&e));
                LUE_AMD_HIP(hipEventRecord(ev[0], stream));
            }
            hipLaunchKernelGGL((sweep_first_round<OP, T, TW, TH>), dim3(static_cast<unsigned>(ntiles)), dim3(256), 0,
                               stream, a);
            LUE_AMD_HIP(hipGetLastError());
            if (profile.enabled)
                LUE_AMD_HIP(hipEventRecord(ev[1], stream));

            unsigned const grid = static_cast<unsigned>(std::min<std::int64_t>(ntiles, 2048));
            std::uint32_t round = 1;
            std::uint32_t pending_tiles = 0;
            bool use_tail = false;
            TailPolicy const policy = tail_policy;
            std::int64_t const tail_tiles =
                std::max<std::int64_t>(static_cast<std::int64_t>(ntiles * policy.tile_fraction), 4);
            while (true)
            {
                std::uint32_t const batch = policy.min_rounds >= 0 ? 4u : std::min<std::uint32_t>(round, 64u);
                for (std::uint32_t k = 0; k < batch; ++k, ++round)
                {
                    hipLaunchKernelGGL((sweep_next_round<OP, T, TW, TH>), dim3(grid), dim3(256), 0, stream, a, round);
                }
                LUE_AMD_HIP(hipGetLastError());
                LUE_AMD_HIP(hipMemcpyAsync(&pending_tiles, s.work_count + (round % 3), sizeof(std::uint32_t),
                                           hipMemcpyDeviceToHost, stream));
                LUE_AMD_HIP(hipStreamSynchronize(stream));
                if (pending_tiles == 0)
                    break;
                if (policy.min_rounds >= 0 && static_cast<int>(round) > policy.min_rounds &&
                    (pending_tiles <= tail_tiles || static_cast<int>(round) > policy.max_rounds))
                {
                    use_tail = true;
                    break;
                }
                if (round > 0xFFFF0000u)
                    return fail(status_invalid_network, "routing did not converge");
            }
            unsigned long long tail_solved = 0;
            if (use_tail)
            {
                if (int const st = run_tail<OP, T, TW, TH>(a, stream, &tail_solved))
                    return st;
            }
            if (profile.enabled)
                LUE_AMD_HIP(hipEventRecord(ev[2], stream));
            unsigned long long counters[2] = {0, 0};
            LUE_AMD_HIP(hipMemcpyAsync(counters, s.counters, sizeof(counters), hipMemcpyDeviceToHost, stream));
            LUE_AMD_HIP(hipStreamSynchronize(stream));
            if (profile.enabled)
            {
                float first = 0.0f, rest = 0.0f;
                LUE_AMD_HIP(hipEventElapsedTime(&first, ev[0], ev[1]));
                LUE_AMD_HIP(hipEventElapsedTime(&rest, ev[1], ev[2]));
                profile.first_round_ms += first;
                profile.next_rounds_ms += rest;
                profile.calls += 1;
                profile.rounds += round;
                profile.cells += a.rows * a.cols;
                for (auto& e : ev)
                    (void)hipEventDestroy(e);
            }
            if (counters[0] + counters[1] + tail_solved != static_cast<unsigned long long>(a.rows * a.cols))
            {
                return fail(status_invalid_network,
                            "invalid flow direction network: " +
                                std::to_string(a.rows * a.cols - counters[0] - counters[1] - tail_solved) +
                                " cells lie on or downstream of a cycle");
            }
            return status_success;
        }

        template<typename T>
        SweepArgs<T> base_args(std::uint8_t const* ldd, std::int64_t rows, std::int64_t cols)
        {
            SweepArgs<T> a{};
            a.ldd = ldd;
            a.rows = rows;
            a.cols = cols;
            return a;
        }

        template<typename T>
        int accu_impl(std::uint8_t const* ldd, std::int64_t rows, std::int64_t cols, T const* inflow, T inflow_s,
                      T* outflow, void* stream)
        {
            if (int const st = check_shape(ldd, rows, cols))
                return st;
            if (!outflow)
                return fail(status_invalid_argument, "outflow must not be null");
            auto a = base_args<T>(ldd, rows, cols);
            a.in1 = inflow;
            a.in1_s = inflow_s;
            a.out = outflow;
            if constexpr (std::is_same_v<T, float>)
                return run_sweep<op_accu, T, 64, 64>(a, static_cast<hipStream_t>(stream));
            else
                return run_sweep<op_accu, T, 64, 32>(a, static_cast<hipStream_t>(stream));
        }

        template<int OP, typename T>
        int split_impl(std::uint8_t const* ldd, std::int64_t rows, std::int64_t cols, T const* inflow, T inflow_s,
                       T const* param, T param_s, T* outflow, T* remainder, void* stream)
        {
            if (int const st = check_shape(ldd, rows, cols))
                return st;
            if (!outflow || !remainder)
                return fail(status_invalid_argument, "outflow and remainder must not be null");
            auto a = base_args<T>(ldd, rows, cols);
            a.in1 = inflow;
            a.in1_s = inflow_s;
            a.in2 = param;
            a.in2_s = param_s;
            a.out = outflow;
            a.rem = remainder;
            if constexpr (std::is_same_v<T, float>)
                return run_sweep<OP, T, 64, 32>(a, static_cast<hipStream_t>(stream));
            else
                return run_sweep<OP, T, 32, 32>(a, static_cast<hipStream_t>(stream));
        }

        template<typename T>
        int kinematic_wave_impl(std::uint8_t const* ldd, std::int64_t rows, std::int64_t cols, T const* current,
                                T const* inflow, T const* alpha, T alpha_s, T const* beta, T beta_s, T dt,
                                T const* dx, T dx_s, T* outflow, void* stream)
        {
            if (int const st = check_shape(ldd, rows, cols))
                return st;
            if (!outflow || !current || !inflow)
                return fail(status_invalid_argument, "current_outflow, inflow and outflow must not be null");
            auto a = base_args<T>(ldd, rows, cols);
            a.in1 = current;
            a.in2 = inflow;
            a.alpha = alpha;
            a.alpha_s = alpha_s;
            a.beta = beta;
            a.beta_s = beta_s;
            a.dx = dx;
            a.dx_s = dx_s;
            a.dt = dt;
            a.out = outflow;
            return run_sweep<op_kinematic_wave, T, 32, 32>(a, static_cast<hipStream_t>(stream));
        }

        // iterate_to_new_discharge, elementwise (kinematic_wave.hpp:145-254, extraction included)
        template<typename T>
        __global__ void iterate_kernel(T const* up, T const* cur, T const* lat, T const* alpha, T const* beta,
                                       T const* dt, T const* dx, T* out, std::int64_t n)
        {
            std::int64_t const i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
            if (i >= n)
                return;
            T q = T{0};
            T const q_lat = lat[i];
            if (up[i] + cur[i] > 0 || q_lat > 0)
            {
                T const inflow = q_lat >= 0 ? q_lat : T{0};
                double const c1 = static_cast<double>(dt[i]) / static_cast<double>(dx[i]);
                double const b = static_cast<double>(alpha[i]) *
                                     pow(static_cast<double>(cur[i]), static_cast<double>(beta[i])) +
                                 static_cast<double>(dt[i]) * static_cast<double>(inflow);
                if constexpr (std::is_same_v<T, double>)
                {
                    KinematicWaveCell<double> cell{up[i], cur[i], inflow, alpha[i], beta[i], dt[i], c1, b};
                    bool converged;
                    q = solve_bracketed<double>(cell, converged);
                }
                else
                {
                    q = static_cast<float>(
                        new_discharge_f32_path(up[i], cur[i], inflow, alpha[i], beta[i], dt[i], b, c1));
                }
            }
            if (q_lat < T{0})
            {
                T const extraction = fmin(dx[i] * fabs(q_lat), q);
                q -= extraction;
            }
            out[i] = q;
        }

        template<typename T>
        int iterate_impl(T const* up, T const* cur, T const* lat, T const* alpha, T const* beta, T const* dt,
                         T const* dx, T* out, std::int64_t n, void* stream)
        {
            if (n < 0 || (n > 0 && (!up || !cur || !lat || !alpha || !beta || !dt || !dx || !out)))
                return fail(status_invalid_argument, "null argument");
            if (n == 0)
                return status_success;
            hipStream_t const s = static_cast<hipStream_t>(stream);
            hipLaunchKernelGGL((iterate_kernel<T>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s,
                               up, cur, lat, alpha, beta, dt, dx, out, n);
            LUE_AMD_HIP(hipGetLastError());
            LUE_AMD_HIP(hipStreamSynchronize(s));
            return status_success;
        }

        template<typename T>
        int uniform_impl(std::uint64_t seed, std::uint64_t stream_id, double lo, double hi, std::int64_t n, T* out,
                         void* stream)
        {
            if (n < 0 || (n > 0 && !out))
                return fail(status_invalid_argument, "null output");
            if (n == 0)
                return status_success;
            hipStream_t const s = static_cast<hipStream_t>(stream);
            unsigned const grid = static_cast<unsigned>(std::min<std::int64_t>((n + 255) / 256, 65536));
            hipLaunchKernelGGL((uniform_kernel<T>), dim3(grid), dim3(256), 0, s, seed, stream_id, lo, hi, n, out);
            LUE_AMD_HIP(hipGetLastError());
            LUE_AMD_HIP(hipStreamSynchronize(s));
            return status_success;
        }

    }  // namespace
}  // namespace lue_amd

using namespace lue_amd;

extern "C" {

const char* lue_amd_status_string(int status)
{
    switch (status)
    {
        case status_success: return "success";
        case status_invalid_argument: return "invalid argument";
        case status_shape_mismatch: return "shapes differ";
        case status_out_of_memory: return "out of device memory";
        case status_hip_error: return "HIP error";
        case status_invalid_network: return "invalid flow direction network";
        case status_no_device: return "no GPU device";
        default: return "unknown status";
    }
}

const char* lue_amd_last_error(void)
{
    return last_error.c_str();
}

const char* lue_amd_version(void)
{
    return "0.1.0";
}

int lue_amd_set_tail_policy(int min_rounds, int max_rounds, double tile_fraction)
{
    if (max_rounds < 0 || !(tile_fraction >= 0.0) || tile_fraction > 1.0)
        return fail(status_invalid_argument, "bad tail policy");
    tail_policy = TailPolicy{min_rounds, max_rounds, tile_fraction};
    return status_success;
}

int lue_amd_profile_enable(int enable)
{
    profile = Profile{};
    profile.enabled = enable != 0;
    return status_success;
}

int lue_amd_profile_read(double* first_round_ms, double* next_rounds_ms, long long* calls, long long* rounds,
                         long long* cells)
{
    if (!first_round_ms || !next_rounds_ms || !calls || !rounds || !cells)
        return fail(status_invalid_argument, "null pointer");
    *first_round_ms = profile.first_round_ms;
    *next_rounds_ms = profile.next_rounds_ms;
    *calls = profile.calls;
    *rounds = profile.rounds;
    *cells = profile.cells;
    return status_success;
}

int lue_amd_device_malloc(void** ptr, size_t bytes)
{
    if (!ptr)
        return fail(status_invalid_argument, "null pointer");
    LUE_AMD_HIP(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
    return status_success;
}

int lue_amd_device_free(void* ptr)
{
    if (ptr)
        LUE_AMD_HIP(hipFree(ptr));
    return status_success;
}

int lue_amd_copy_to_device(void* dst, const void* src, size_t bytes, void* stream)
{
    hipStream_t const s = static_cast<hipStream_t>(stream);
    LUE_AMD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    LUE_AMD_HIP(hipStreamSynchronize(s));
    return status_success;
}

int lue_amd_copy_to_host(void* dst, const void* src, size_t bytes, void* stream)
{
    hipStream_t const s = static_cast<hipStream_t>(stream);
    LUE_AMD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    LUE_AMD_HIP(hipStreamSynchronize(s));
    return status_success;
}

int lue_amd_stream_synchronize(void* stream)
{
    LUE_AMD_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return status_success;
}

int lue_amd_inflow_count_u8(const uint8_t* ldd, int64_t rows, int64_t cols, uint8_t* inflow_count, void* stream)
{
    if (int const st = check_shape(ldd, rows, cols))
        return st;
    if (!inflow_count)
        return fail(status_invalid_argument, "inflow_count must not be null");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    dim3 const grid(static_cast<unsigned>((cols + ic_tw - 1) / ic_tw), static_cast<unsigned>((rows + ic_th - 1) / ic_th));
    hipLaunchKernelGGL(inflow_count_kernel, grid, dim3(256), 0, s, ldd, rows, cols, inflow_count);
    LUE_AMD_HIP(hipGetLastError());
    LUE_AMD_HIP(hipStreamSynchronize(s));
    return status_success;
}

int lue_amd_accu_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow, float inflow_scalar,
                     float* outflow, void* stream)
{
    return accu_impl<float>(ldd, rows, cols, inflow, inflow_scalar, outflow, stream);
}

int lue_amd_accu_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow, double inflow_scalar,
                     double* outflow, void* stream)
{
    return accu_impl<double>(ldd, rows, cols, inflow, inflow_scalar, outflow, stream);
}

int lue_amd_accu_threshold_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow,
                               float inflow_scalar, const float* threshold, float threshold_scalar, float* outflow,
                               float* remainder, void* stream)
{
    return split_impl<op_threshold, float>(ldd, rows, cols, inflow, inflow_scalar, threshold, threshold_scalar,
                                           outflow, remainder, stream);
}

int lue_amd_accu_threshold_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow,
                               double inflow_scalar, const double* threshold, double threshold_scalar,
                               double* outflow, double* remainder, void* stream)
{
    return split_impl<op_threshold, double>(ldd, rows, cols, inflow, inflow_scalar, threshold, threshold_scalar,
                                            outflow, remainder, stream);
}

int lue_amd_accu_fraction_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow,
                              float inflow_scalar, const float* fraction, float fraction_scalar, float* outflow,
                              float* remainder, void* stream)
{
    return split_impl<op_fraction, float>(ldd, rows, cols, inflow, inflow_scalar, fraction, fraction_scalar,
                                          outflow, remainder, stream);
}

int lue_amd_accu_fraction_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow,
                              double inflow_scalar, const double* fraction, double fraction_scalar,
                              double* outflow, double* remainder, void* stream)
{
    return split_impl<op_fraction, double>(ldd, rows, cols, inflow, inflow_scalar, fraction, fraction_scalar,
                                           outflow, remainder, stream);
}

int lue_amd_kinematic_wave_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* current_outflow,
                               const float* inflow, const float* alpha, float alpha_scalar, const float* beta,
                               float beta_scalar, float time_step_duration, const float* channel_length,
                               float channel_length_scalar, float* outflow, void* stream)
{
    return kinematic_wave_impl<float>(ldd, rows, cols, current_outflow, inflow, alpha, alpha_scalar, beta,
                                      beta_scalar, time_step_duration, channel_length, channel_length_scalar,
                                      outflow, stream);
}

int lue_amd_kinematic_wave_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* current_outflow,
                               const double* inflow, const double* alpha, double alpha_scalar, const double* beta,
                               double beta_scalar, double time_step_duration, const double* channel_length,
                               double channel_length_scalar, double* outflow, void* stream)
{
    return kinematic_wave_impl<double>(ldd, rows, cols, current_outflow, inflow, alpha, alpha_scalar, beta,
                                       beta_scalar, time_step_duration, channel_length, channel_length_scalar,
                                       outflow, stream);
}

int lue_amd_iterate_to_new_discharge_f32(const float* upstream_discharge, const float* current_discharge,
                                         const float* lateral_inflow, const float* alpha, const float* beta,
                                         const float* time_step_duration, const float* channel_length,
                                         float* new_discharge, int64_t n, void* stream)
{
    return iterate_impl<float>(upstream_discharge, current_discharge, lateral_inflow, alpha, beta,
                               time_step_duration, channel_length, new_discharge, n, stream);
}

int lue_amd_iterate_to_new_discharge_f64(const double* upstream_discharge, const double* current_discharge,
                                         const double* lateral_inflow, const double* alpha, const double* beta,
                                         const double* time_step_duration, const double* channel_length,
                                         double* new_discharge, int64_t n, void* stream)
{
    return iterate_impl<double>(upstream_discharge, current_discharge, lateral_inflow, alpha, beta,
                                time_step_duration, channel_length, new_discharge, n, stream);
}

int lue_amd_synthetic_ldd(uint64_t seed, int64_t rows, int64_t cols, uint8_t* ldd, void* stream)
{
    if (rows <= 0 || cols <= 0 || !ldd)
        return fail(status_invalid_argument, "bad shape or null output");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    double table[256];
    for (int k = 0; k < 256; ++k)
        table[k] = std::fabs(std::sin(M_PI * static_cast<double>(k) / 256.0));
    double* d_table = nullptr;
    LUE_AMD_HIP(hipMallocAsync(reinterpret_cast<void**>(&d_table), sizeof(table), s));
    LUE_AMD_HIP(hipMemcpyAsync(d_table, table, sizeof(table), hipMemcpyHostToDevice, s));
    dim3 const grid(static_cast<unsigned>((cols + dem_tw - 1) / dem_tw), static_cast<unsigned>((rows + dem_th - 1) / dem_th));
    hipLaunchKernelGGL(dem_ldd_kernel, grid, dim3(256), 0, s, seed, rows, cols, d_table, ldd);
    LUE_AMD_HIP(hipGetLastError());
    LUE_AMD_HIP(hipFreeAsync(d_table, s));
    LUE_AMD_HIP(hipStreamSynchronize(s));
    return status_success;
}

int lue_amd_synthetic_uniform_f32(uint64_t seed, uint64_t stream_id, double lo, double hi, int64_t n, float* out,
                                  void* stream)
{
    return uniform_impl<float>(seed, stream_id, lo, hi, n, out, stream);
}

int lue_amd_synthetic_uniform_f64(uint64_t seed, uint64_t stream_id, double lo, double hi, int64_t n, double* out,
                                  void* stream)
{
    return uniform_impl<double>(seed, stream_id, lo, hi, n, out, stream);
}

}  // extern "C"
