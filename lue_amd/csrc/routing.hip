// C ABI of the flow-routing path (include/lue_amd/routing.h), the parts outside the routing sweep: status
// and error messages, the process-wide caches, inflow count, d8_flow_direction, upstream / downstream /
// downstream_distance, iterate_to_new_discharge and the synthetic-input generators. The sweep operations
// are in op_*.hip (one translation unit per operation).
#include "lue_amd/routing.h"

#include "api.hpp"

#include "d8_flow_direction.hpp"
#include "flow_neighbours.hpp"
#include "inflow_count.hpp"
#include "kinematic_wave.hpp"
#include "routing_common.hpp"
#include "sweep.hpp"
#include "synthetic.hpp"
#include "network.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <list>
#include <memory>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace lue_amd {

    std::string& last_error_message()
    {
        thread_local std::string message;
        return message;
    }

    Profile& profile()
    {
        static Profile p;
        return p;
    }

    ScratchCache& scratch_cache()
    {
        // never destroyed: freeing device memory at process exit is not needed
        static ScratchCache* cache = new ScratchCache;
        return *cache;
    }

    HostStaging& host_staging()
    {
        // never destroyed: pinned slots and events stay valid until the process ends
        static HostStaging* s = new HostStaging;
        return *s;
    }

    int& default_schedule()
    {
        static int schedule = schedule_chains;
        return schedule;
    }

    NetworkCache& network_cache()
    {
        // never destroyed: freeing device memory at process exit is not needed
        static NetworkCache* cache = new NetworkCache;
        return *cache;
    }

    int ldd_fingerprint(std::uint8_t const* ldd, std::int64_t n, hipStream_t stream, std::uint64_t (&fingerprint)[2])
    {
        ScratchArray<unsigned long long> sum;
        LUE_AMD_HIP(sum.allocate(2, stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, reinterpret_cast<std::uint32_t*>(sum.ptr), 0u,
                           4ull);
        unsigned const grid = static_cast<unsigned>(std::min<std::int64_t>((n / 4 + 255) / 256 + 1, 4096));
        hipLaunchKernelGGL(fingerprint_kernel, dim3(grid), dim3(256), 0, stream, ldd, static_cast<std::uint64_t>(n),
                           sum.ptr);
        LUE_AMD_HIP(hipGetLastError());
        unsigned long long value[2] = {0, 0};
        LUE_AMD_HIP(copy_to_host(value, sum.ptr, 2, stream));
        fingerprint[0] = value[0];
        fingerprint[1] = value[1];
        return status_success;
    }

    namespace {

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
            return status_success;
        }

    }  // namespace
}  // namespace lue_amd

using namespace lue_amd;

namespace {

    template<typename T>
    int dem_impl(std::uint64_t seed, std::int64_t rows, std::int64_t cols, T* out, void* stream)
    {
        if (rows <= 0 || cols <= 0 || !out)
            return fail(status_invalid_argument, "bad shape or null output");
        hipStream_t const s = static_cast<hipStream_t>(stream);
        double table[256];
        for (int k = 0; k < 256; ++k)
            table[k] = std::fabs(std::sin(M_PI * static_cast<double>(k) / 256.0));
        ScratchArray<double> d_table;
        LUE_AMD_HIP(d_table.allocate(256, s));
        LUE_AMD_HIP(host_staging().copy_to_device(d_table.ptr, table, sizeof(table), s));
        unsigned const grid = static_cast<unsigned>(std::min<std::int64_t>((rows * cols + 255) / 256, 16384));
        hipLaunchKernelGGL(dem_kernel<T>, dim3(grid), dim3(256), 0, s, seed, rows, cols, d_table.ptr, out);
        LUE_AMD_HIP(hipGetLastError());
        return status_success;
    }

    template<typename T>
    int d8_impl(T const* elevation, std::int64_t rows, std::int64_t cols, std::uint8_t* flow_direction,
                void* stream)
    {
        if (rows < 0 || cols < 0)
            return fail(status_invalid_argument, "array shape must not be negative");
        if (rows == 0 || cols == 0)
            return status_success;
        if (!elevation || !flow_direction)
            return fail(status_invalid_argument, "elevation and flow_direction must not be null");
        hipStream_t const s = static_cast<hipStream_t>(stream);
        bool const aligned = cols % 4 == 0 && reinterpret_cast<std::uintptr_t>(elevation) % 16 == 0 &&
                             reinterpret_cast<std::uintptr_t>(flow_direction) % 4 == 0;
        if (aligned)
        {
            std::int64_t const gx = (cols + d8_strip_cols - 1) / d8_strip_cols;
            std::int64_t const gy = (rows + d8_strip_rows - 1) / d8_strip_rows;
            if (gx > 0x7fffffff || gy > 0xffff)
                return fail(status_invalid_argument, "array too large for one launch");
            hipLaunchKernelGGL(d8_strip_kernel<T>, dim3(static_cast<unsigned>(gx), static_cast<unsigned>(gy)),
                               dim3(d8_threads), 0, s, elevation, rows, cols, flow_direction);
        }
        else
        {
            std::int64_t const gx = (cols + d8_tw - 1) / d8_tw;
            std::int64_t const gy = (rows + d8_th - 1) / d8_th;
            if (gx > 0x7fffffff || gy > 0xffff)
                return fail(status_invalid_argument, "array too large for one launch");
            hipLaunchKernelGGL(d8_flow_direction_kernel<T>,
                               dim3(static_cast<unsigned>(gx), static_cast<unsigned>(gy)), dim3(d8_threads), 0, s,
                               elevation, rows, cols, flow_direction);
        }
        LUE_AMD_HIP(hipGetLastError());
        return status_success;
    }

    template<typename T>
    int downstream_distance_impl(std::uint8_t const* ldd, std::int64_t rows, std::int64_t cols, T cell_size,
                                 int mark_no_data, T* out, void* stream)
    {
        if (rows < 0 || cols < 0)
            return fail(status_invalid_argument, "array shape must not be negative");
        if (rows == 0 || cols == 0)
            return status_success;
        if (!ldd || !out)
            return fail(status_invalid_argument, "flow_direction and distance must not be null");
        if (reinterpret_cast<std::uintptr_t>(ldd) % 4 != 0)
            return fail(status_invalid_argument, "flow_direction must be 4-byte aligned");
        std::int64_t const n = rows * cols;
        T const diagonal = std::sqrt(T{2}) * cell_size;  // as definition/downstream_distance.hpp:31-35 does
        std::int64_t const blocks = std::clamp<std::int64_t>((n / 4 + 255) / 256, 1, 8192);
        hipLaunchKernelGGL(downstream_distance_kernel<T>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), ldd, n, cell_size, diagonal, mark_no_data, out);
        LUE_AMD_HIP(hipGetLastError());
        return status_success;
    }

    template<typename T, bool Upstream>
    int flow_neighbour_impl(std::uint8_t const* ldd, T const* material, std::int64_t rows, std::int64_t cols,
                            T* out, void* stream)
    {
        if (rows < 0 || cols < 0)
            return fail(status_invalid_argument, "array shape must not be negative");
        if (rows == 0 || cols == 0)
            return status_success;
        if (!ldd || !material || !out)
            return fail(status_invalid_argument, "flow_direction, material and output must not be null");
        std::int64_t const gx = (cols + nb_tw - 1) / nb_tw;
        std::int64_t const gy = (rows + nb_th - 1) / nb_th;
        if (gx > 0x7fffffff || gy > 0xffff)
            return fail(status_invalid_argument, "array too large for one launch");
        hipLaunchKernelGGL((flow_neighbour_kernel<T, Upstream>),
                           dim3(static_cast<unsigned>(gx), static_cast<unsigned>(gy)), dim3(nb_threads), 0,
                           static_cast<hipStream_t>(stream), ldd, material, rows, cols, out);
        LUE_AMD_HIP(hipGetLastError());
        return status_success;
    }

}  // namespace

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
    return last_error_message().c_str();
}

const char* lue_amd_version(void)
{
    return "0.1.0";
}

int lue_amd_set_sweep_options(int schedule)
{
    if (schedule < 0 || schedule >= schedule_count)
        return fail(status_invalid_argument, "unknown schedule");
    default_schedule() = schedule;
    return status_success;
}

int lue_amd_set_network_cache(int entries)
{
    if (entries < 0)
        return fail(status_invalid_argument, "entries must be >= 0");
    network_cache().set_capacity(entries);
    return status_success;
}

int lue_amd_trim_scratch(void)
{
    LUE_AMD_HIP(scratch_cache().trim());
    return status_success;
}

int lue_amd_profile_enable(int enable)
{
    LUE_AMD_HIP(profile().reset(enable != 0));
    return status_success;
}

int lue_amd_profile_read(double* first_round_ms, double* next_rounds_ms, long long* calls, long long* rounds,
                         long long* cells)
{
    if (!first_round_ms || !next_rounds_ms || !calls || !rounds || !cells)
        return fail(status_invalid_argument, "null pointer");
    Profile::Totals t;
    LUE_AMD_HIP(profile().read(-1, t));
    *first_round_ms = t.first_round_ms;
    *next_rounds_ms = t.next_rounds_ms;
    *calls = t.calls;
    *rounds = t.rounds;
    *cells = t.cells;
    return status_success;
}

int lue_amd_profile_read_op(int op, int f64, double* first_round_ms, double* next_rounds_ms,
                            double* first_round_min_ms, double* first_round_max_ms, long long* calls)
{
    if (!first_round_ms || !next_rounds_ms || !first_round_min_ms || !first_round_max_ms || !calls)
        return fail(status_invalid_argument, "null pointer");
    if (op < 0 || op >= Profile::kinds / 2)
        return fail(status_invalid_argument, "unknown operation");
    Profile::Totals t;
    LUE_AMD_HIP(profile().read(2 * op + (f64 ? 1 : 0), t));
    *first_round_ms = t.first_round_ms;
    *next_rounds_ms = t.next_rounds_ms;
    *first_round_min_ms = t.first_round_min_ms;
    *first_round_max_ms = t.first_round_max_ms;
    *calls = t.calls;
    return status_success;
}

int lue_amd_plan_launch(lue_amd_plan plan, void* stream)
{
    if (!plan || !plan->plan.exec)
        return fail(status_invalid_argument, "null plan");
    LUE_AMD_HIP(hipGraphLaunch(plan->plan.exec, static_cast<hipStream_t>(stream)));
    return status_success;
}

int lue_amd_plan_destroy(lue_amd_plan plan)
{
    // ~Plan waits for the device before it frees the graph and the buffers
    delete plan;
    return status_success;
}

int lue_amd_plan_info(lue_amd_plan plan, int64_t* info, int n)
{
    if (!plan || !info || n < 0)
        return fail(status_invalid_argument, "null argument");
    int64_t const values[] = {static_cast<int64_t>(plan->plan.nodes), static_cast<int64_t>(plan->plan.bytes)};
    for (int i = 0; i < n && i < static_cast<int>(sizeof(values) / sizeof(values[0])); ++i)
        info[i] = values[i];
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

int lue_amd_event_create(void** event)
{
    if (!event)
        return fail(status_invalid_argument, "null pointer");
    hipEvent_t e = nullptr;
    LUE_AMD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    *event = e;
    return status_success;
}

int lue_amd_event_destroy(void* event)
{
    if (event)
        LUE_AMD_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return status_success;
}

int lue_amd_event_record(void* event, void* stream)
{
    if (!event)
        return fail(status_invalid_argument, "null event");
    LUE_AMD_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
    return status_success;
}

int lue_amd_stream_wait_event(void* stream, void* event)
{
    if (!event)
        return fail(status_invalid_argument, "null event");
    LUE_AMD_HIP(hipStreamWaitEvent(static_cast<hipStream_t>(stream), static_cast<hipEvent_t>(event), 0));
    return status_success;
}

int lue_amd_event_synchronize(void* event)
{
    if (!event)
        return fail(status_invalid_argument, "null event");
    LUE_AMD_HIP(hipEventSynchronize(static_cast<hipEvent_t>(event)));
    return status_success;
}

int lue_amd_event_query(void* event, int* complete)
{
    if (!event || !complete)
        return fail(status_invalid_argument, "null argument");
    hipError_t const e = hipEventQuery(static_cast<hipEvent_t>(event));
    if (e == hipErrorNotReady)
    {
        (void)hipGetLastError();
        *complete = 0;
        return status_success;
    }
    LUE_AMD_HIP(e);
    *complete = 1;
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
    return status_success;
}

int lue_amd_d8_flow_direction_f32(const float* elevation, int64_t rows, int64_t cols, uint8_t* flow_direction,
                                  void* stream)
{
    return d8_impl<float>(elevation, rows, cols, flow_direction, stream);
}

int lue_amd_d8_flow_direction_f64(const double* elevation, int64_t rows, int64_t cols, uint8_t* flow_direction,
                                  void* stream)
{
    return d8_impl<double>(elevation, rows, cols, flow_direction, stream);
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
    ScratchArray<double> d_table;
    LUE_AMD_HIP(d_table.allocate(256, s));
    LUE_AMD_HIP(host_staging().copy_to_device(d_table.ptr, table, sizeof(table), s));
    dim3 const grid(static_cast<unsigned>((cols + dem_tw - 1) / dem_tw), static_cast<unsigned>((rows + dem_th - 1) / dem_th));
    hipLaunchKernelGGL(dem_ldd_kernel, grid, dim3(256), 0, s, seed, rows, cols, d_table.ptr, ldd);
    LUE_AMD_HIP(hipGetLastError());
    return status_success;
}

int lue_amd_downstream_distance_f32(const uint8_t* flow_direction, int64_t rows, int64_t cols, float cell_size,
                                   int mark_no_data, float* distance, void* stream)
{
    return downstream_distance_impl<float>(flow_direction, rows, cols, cell_size, mark_no_data, distance, stream);
}

int lue_amd_downstream_distance_f64(const uint8_t* flow_direction, int64_t rows, int64_t cols, double cell_size,
                                   int mark_no_data, double* distance, void* stream)
{
    return downstream_distance_impl<double>(flow_direction, rows, cols, cell_size, mark_no_data, distance,
                                            stream);
}

#define LUE_AMD_FLOW_NEIGHBOUR_OPS(T, SUFFIX)                                                                 \
    int lue_amd_upstream_##SUFFIX(const uint8_t* flow_direction, const T* material, int64_t rows, int64_t cols, \
                                  T* out, void* stream)                                                      \
    {                                                                                                        \
        return flow_neighbour_impl<T, true>(flow_direction, material, rows, cols, out, stream);              \
    }                                                                                                        \
    int lue_amd_downstream_##SUFFIX(const uint8_t* flow_direction, const T* material, int64_t rows,         \
                                    int64_t cols, T* out, void* stream)                                      \
    {                                                                                                        \
        return flow_neighbour_impl<T, false>(flow_direction, material, rows, cols, out, stream);             \
    }

LUE_AMD_FLOW_NEIGHBOUR_OPS(uint8_t, u8)
LUE_AMD_FLOW_NEIGHBOUR_OPS(uint32_t, u32)
LUE_AMD_FLOW_NEIGHBOUR_OPS(uint64_t, u64)
LUE_AMD_FLOW_NEIGHBOUR_OPS(int8_t, i8)
LUE_AMD_FLOW_NEIGHBOUR_OPS(int32_t, i32)
LUE_AMD_FLOW_NEIGHBOUR_OPS(int64_t, i64)
LUE_AMD_FLOW_NEIGHBOUR_OPS(float, f32)
LUE_AMD_FLOW_NEIGHBOUR_OPS(double, f64)

int lue_amd_synthetic_dem_f32(uint64_t seed, int64_t rows, int64_t cols, float* out, void* stream)
{
    return dem_impl<float>(seed, rows, cols, out, stream);
}

int lue_amd_synthetic_dem_f64(uint64_t seed, int64_t rows, int64_t cols, double* out, void* stream)
{
    return dem_impl<double>(seed, rows, cols, out, stream);
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

