// C ABI of the flow-routing path (include/lue_amd/routing.h): argument checks, device scratch, the
// round loop of the tiled sweep, and the synthetic-input generators.
#include "lue_amd/routing.h"

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

    namespace {

        // Schedule of per-call entry points (lue_amd_set_sweep_options).
        int default_schedule = schedule_chains;

        template<int OP, typename T, int TW, int TH>
        constexpr bool fits()
        {
            return sizeof(SweepLds<OP, T, TW, TH>) <= 160 * 1024;
        }

        // Call f.template operator()<TW, TH>() for a tile shape the operation fits in.
        template<int OP, typename T, typename F>
        int with_tile_shape(int shape, F&& f)
        {
            switch (shape)
            {
                case tile_32x32:
                    if constexpr (fits<OP, T, 32, 32>())
                        return f.template operator()<32, 32>();
                    break;
                case tile_64x32:
                    if constexpr (fits<OP, T, 64, 32>())
                        return f.template operator()<64, 32>();
                    break;
                case tile_64x64:
                    if constexpr (fits<OP, T, 64, 64>())
                        return f.template operator()<64, 64>();
                    break;
                default:
                    break;
            }
            return fail(status_invalid_argument, "tile shape does not fit this operation in LDS");
        }

        // Tile shape of per-call entry points. With the chain schedule the first round's time is set by the
        // latency of a tile's walks and by how many tiles fit a CU at once, which favours the smallest tiles
        // (32x32 measured fastest for every operation at 16384^2 cells).
        template<int OP, typename T>
        constexpr int default_tile_shape()
        {
            return tile_32x32;
        }

        // Scalars of an operation (rasters come per catchment in CatchmentIO).
        template<typename T>
        struct Scalars
        {
            T in1_s{};
            T in2_s{};
            T alpha_s{};
            T beta_s{};
            T dx_s{};
            T dt{};
        };

        template<typename T>
        SweepArgs<T> args_from(Scalars<T> const& sc)
        {
            SweepArgs<T> a{};
            a.in1_s = sc.in1_s;
            a.in2_s = sc.in2_s;
            a.alpha_s = sc.alpha_s;
            a.beta_s = sc.beta_s;
            a.dx_s = sc.dx_s;
            a.dt = sc.dt;
            return a;
        }

        // Networks of the per-call entry points, most recently used first. A network is found again by
        // device, ldd pointer, shape, tile shape, schedule and a fingerprint of the ldd bytes (computed in
        // every call), so a changed ldd is recorded anew.
        struct CacheKey
        {
            int device;
            void const* ldd;
            std::int64_t rows;
            std::int64_t cols;
            int tile_shape;
            int schedule;
            std::uint64_t fingerprint;

            bool operator==(CacheKey const&) const = default;
        };

        struct NetworkCache
        {
            std::mutex mutex;
            std::list<std::pair<CacheKey, std::shared_ptr<Network>>> entries;
            int capacity = 4;

            std::shared_ptr<Network> find(CacheKey const& key)
            {
                std::lock_guard<std::mutex> lock(mutex);
                for (auto it = entries.begin(); it != entries.end(); ++it)
                {
                    if (it->first == key)
                    {
                        entries.splice(entries.begin(), entries, it);
                        return it->second;
                    }
                }
                return nullptr;
            }

            void insert(CacheKey const& key, std::shared_ptr<Network> net)
            {
                std::lock_guard<std::mutex> lock(mutex);
                entries.emplace_front(key, std::move(net));
                trim();
            }

            // Drop entries beyond the capacity. Their buffers are freed on the null stream after the
            // device has finished all work, since a replay may still use them.
            void trim()
            {
                if (static_cast<int>(entries.size()) <= capacity)
                    return;
                (void)hipDeviceSynchronize();
                while (static_cast<int>(entries.size()) > capacity)
                    entries.pop_back();
            }
        };

        NetworkCache& network_cache()
        {
            // never destroyed: freeing device memory at process exit is not needed
            static NetworkCache* cache = new NetworkCache;
            return *cache;
        }

        int ldd_fingerprint(std::uint8_t const* ldd, std::int64_t n, hipStream_t stream, std::uint64_t& fingerprint)
        {
            if (reinterpret_cast<std::uintptr_t>(ldd) % 4 != 0)
                return fail(status_invalid_argument, "ldd must be 4-byte aligned");
            ScratchArray<unsigned long long> sum;
            LUE_AMD_HIP(sum.allocate(1, stream));
            hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, reinterpret_cast<std::uint32_t*>(sum.ptr),
                               0u, 2ull);
            unsigned const grid = static_cast<unsigned>(std::min<std::int64_t>((n / 4 + 255) / 256 + 1, 4096));
            hipLaunchKernelGGL(fingerprint_kernel, dim3(grid), dim3(256), 0, stream, ldd, static_cast<std::uint64_t>(n),
                               sum.ptr);
            LUE_AMD_HIP(hipGetLastError());
            unsigned long long value = 0;
            LUE_AMD_HIP(copy_to_host(&value, sum.ptr, 1, stream));
            fingerprint = value;
            return status_success;
        }

        // One operation on one raster: replay a cached network of the ldd, or record one.
        template<int OP, typename T>
        int route_once(std::uint8_t const* ldd, std::int64_t rows, std::int64_t cols, Scalars<T> const& sc,
                       CatchmentIO<T> const& io, hipStream_t stream)
        {
            std::vector<CatchmentIO<T>> table{io};
            int const shape = default_tile_shape<OP, T>();
            NetworkCache& cache = network_cache();
            CacheKey key{0, ldd, rows, cols, shape, default_schedule, 0};
            bool const caching = cache.capacity > 0;
            if (caching)
            {
                LUE_AMD_HIP(hipGetDevice(&key.device));
                if (int const st = ldd_fingerprint(ldd, rows * cols, stream, key.fingerprint))
                    return st;
                if (std::shared_ptr<Network> net = cache.find(key))
                {
                    return with_tile_shape<OP, T>(net->tile_shape, [&]<int TW, int TH>()
                                                  { return replay_run<OP, T, TW, TH>(*net, args_from(sc), table, stream); });
                }
            }
            auto net = std::make_shared<Network>();
            if (int const st = network_init(*net, 1, &ldd, &rows, &cols, shape, default_schedule, stream))
                return st;
            int const st = with_tile_shape<OP, T>(net->tile_shape, [&]<int TW, int TH>()
                                                  { return record_run<OP, T, TW, TH>(*net, args_from(sc), table, stream); });
            if (st == status_success && caching)
            {
                net->detach_stream();
                cache.insert(key, std::move(net));
            }
            return st;
        }

        // One operation on a recorded network.
        template<int OP, typename T>
        int route_network(Network const& net, Scalars<T> const& sc, std::vector<CatchmentIO<T>> const& io,
                          hipStream_t stream)
        {
            if (io.size() != net.cats_host.size())
                return fail(status_invalid_argument, "one input/output set per catchment");
            return with_tile_shape<OP, T>(net.tile_shape, [&]<int TW, int TH>()
                                          { return replay_run<OP, T, TW, TH>(net, args_from(sc), io, stream); });
        }

        int check_shape(void const* ldd, std::int64_t rows, std::int64_t cols)
        {
            if (!ldd)
                return fail(status_invalid_argument, "ldd must not be null");
            if (rows <= 0 || cols <= 0)
                return fail(status_invalid_argument, "array shape must not be empty");
            return status_success;
        }

        template<typename T>
        int accu_impl(std::uint8_t const* ldd, std::int64_t rows, std::int64_t cols, T const* inflow, T inflow_s,
                      T* outflow, void* stream)
        {
            if (int const st = check_shape(ldd, rows, cols))
                return st;
            if (!outflow)
                return fail(status_invalid_argument, "outflow must not be null");
            Scalars<T> sc{};
            sc.in1_s = inflow_s;
            CatchmentIO<T> io{inflow, nullptr, nullptr, nullptr, nullptr, outflow, nullptr};
            return route_once<op_accu, T>(ldd, rows, cols, sc, io, static_cast<hipStream_t>(stream));
        }

        template<int OP, typename T>
        int split_impl(std::uint8_t const* ldd, std::int64_t rows, std::int64_t cols, T const* inflow, T inflow_s,
                       T const* param, T param_s, T* outflow, T* remainder, void* stream)
        {
            if (int const st = check_shape(ldd, rows, cols))
                return st;
            if (!outflow || !remainder)
                return fail(status_invalid_argument, "outflow and remainder must not be null");
            Scalars<T> sc{};
            sc.in1_s = inflow_s;
            sc.in2_s = param_s;
            CatchmentIO<T> io{inflow, param, nullptr, nullptr, nullptr, outflow, remainder};
            return route_once<OP, T>(ldd, rows, cols, sc, io, static_cast<hipStream_t>(stream));
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
            Scalars<T> sc{};
            sc.alpha_s = alpha_s;
            sc.beta_s = beta_s;
            sc.dx_s = dx_s;
            sc.dt = dt;
            CatchmentIO<T> io{current, inflow, alpha, beta, dx, outflow, nullptr};
            return route_once<op_kinematic_wave, T>(ldd, rows, cols, sc, io, static_cast<hipStream_t>(stream));
        }

        // Per-catchment pointer tables of the network entry points (host arrays of device pointers; a null
        // table means the scalar).
        template<typename T>
        T const* entry(T const* const* table, int c)
        {
            return table ? table[c] : nullptr;
        }

        template<typename T>
        T* entry(T* const* table, int c)
        {
            return table ? table[c] : nullptr;
        }

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
    default_schedule = schedule;
    return status_success;
}

int lue_amd_set_network_cache(int entries)
{
    if (entries < 0)
        return fail(status_invalid_argument, "entries must be >= 0");
    NetworkCache& cache = network_cache();
    cache.capacity = entries;
    cache.trim();
    return status_success;
}

int lue_amd_trim_scratch(void)
{
    LUE_AMD_HIP(scratch_cache().trim());
    return status_success;
}

int lue_amd_profile_enable(int enable)
{
    profile() = Profile{};
    profile().enabled = enable != 0;
    return status_success;
}

int lue_amd_profile_read(double* first_round_ms, double* next_rounds_ms, long long* calls, long long* rounds,
                         long long* cells)
{
    if (!first_round_ms || !next_rounds_ms || !calls || !rounds || !cells)
        return fail(status_invalid_argument, "null pointer");
    *first_round_ms = profile().first_round_ms;
    *next_rounds_ms = profile().next_rounds_ms;
    *calls = profile().calls;
    *rounds = profile().rounds;
    *cells = profile().cells;
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

int lue_amd_accu_capacity_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow,
                              float inflow_scalar, const float* capacity, float capacity_scalar, float* outflow,
                              float* remainder, void* stream)
{
    return split_impl<op_capacity, float>(ldd, rows, cols, inflow, inflow_scalar, capacity, capacity_scalar,
                                          outflow, remainder, stream);
}

int lue_amd_accu_capacity_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow,
                              double inflow_scalar, const double* capacity, double capacity_scalar,
                              double* outflow, double* remainder, void* stream)
{
    return split_impl<op_capacity, double>(ldd, rows, cols, inflow, inflow_scalar, capacity, capacity_scalar,
                                           outflow, remainder, stream);
}

int lue_amd_accu_trigger_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow,
                             float inflow_scalar, const float* trigger, float trigger_scalar, float* outflow,
                             float* remainder, void* stream)
{
    return split_impl<op_trigger, float>(ldd, rows, cols, inflow, inflow_scalar, trigger, trigger_scalar, outflow,
                                         remainder, stream);
}

int lue_amd_accu_trigger_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow,
                             double inflow_scalar, const double* trigger, double trigger_scalar, double* outflow,
                             double* remainder, void* stream)
{
    return split_impl<op_trigger, double>(ldd, rows, cols, inflow, inflow_scalar, trigger, trigger_scalar, outflow,
                                          remainder, stream);
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

// ---------------------------------------------------------------------------------------------------
// Flow networks: a batch of catchments with the recorded sweep schedule

struct lue_amd_network_s
{
    lue_amd::Network net;
};

extern "C" {

int lue_amd_network_create(int n, const uint8_t* const* ldd, const int64_t* rows, const int64_t* cols, int tile_shape,
                           int schedule, void* stream, lue_amd_network* network)
{
    if (!network)
        return fail(status_invalid_argument, "network must not be null");
    *network = nullptr;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    auto handle = std::make_unique<lue_amd_network_s>();
    Network& net = handle->net;
    if (int const st = network_init(net, n, ldd, reinterpret_cast<std::int64_t const*>(rows),
                                    reinterpret_cast<std::int64_t const*>(cols), tile_shape, schedule, s))
        return st;
    // record the schedule with accu(ldd, 1) into scratch (every operation fits every tile shape of accu f32)
    DeviceArray<float> scratch;
    LUE_AMD_HIP(scratch.allocate(static_cast<std::size_t>(net.ncells), s));
    std::vector<CatchmentIO<float>> io;
    for (auto const& cat : net.cats_host)
        io.push_back(CatchmentIO<float>{nullptr, nullptr, nullptr, nullptr, nullptr, scratch.ptr + cat.cell_base, nullptr});
    Scalars<float> sc{};
    sc.in1_s = 1.0f;
    int const st = with_tile_shape<op_accu, float>(net.tile_shape, [&]<int TW, int TH>()
                                                   { return record_run<op_accu, float, TW, TH>(net, args_from(sc), io, s,
                                                                                               false); });
    if (st)
        return st;
    LUE_AMD_HIP(hipStreamSynchronize(s));
    *network = handle.release();
    return status_success;
}

int lue_amd_network_destroy(lue_amd_network network)
{
    if (network)
    {
        int device = 0;
        (void)hipGetDevice(&device);
        (void)hipDeviceSynchronize();
        network->net.detach_stream();
        delete network;
    }
    return status_success;
}

int lue_amd_network_info(lue_amd_network network, int64_t* info, int n)
{
    if (!network || !info || n < 0)
        return fail(status_invalid_argument, "null argument");
    Network const& net = network->net;
    int64_t const values[] = {static_cast<int64_t>(net.cats_host.size()),
                              net.ncells,
                              net.ntiles,
                              net.rounds(),
                              net.pending,
                              net.nchains,
                              net.levels(),
                              net.nodata_cells,
                              net.tile_shape,
                              net.schedule,
                              net.longest_chain,
                              static_cast<int64_t>(net.nrefs)};
    for (int i = 0; i < n && i < static_cast<int>(sizeof(values) / sizeof(values[0])); ++i)
        info[i] = values[i];
    return status_success;
}

#define LUE_AMD_NETWORK_OPS(T, SUFFIX)                                                                          \
    int lue_amd_network_accu_##SUFFIX(lue_amd_network network, const T* const* inflow, T inflow_scalar,         \
                                      T* const* outflow, void* stream)                                          \
    {                                                                                                           \
        if (!network || !outflow)                                                                               \
            return fail(status_invalid_argument, "null argument");                                              \
        Network const& net = network->net;                                                                      \
        Scalars<T> sc{};                                                                                        \
        sc.in1_s = inflow_scalar;                                                                               \
        std::vector<CatchmentIO<T>> io;                                                                         \
        for (int c = 0; c < static_cast<int>(net.cats_host.size()); ++c)                                        \
        {                                                                                                       \
            if (!outflow[c])                                                                                    \
                return fail(status_invalid_argument, "outflow must not be null");                              \
            io.push_back(CatchmentIO<T>{entry(inflow, c), nullptr, nullptr, nullptr, nullptr, outflow[c], nullptr}); \
        }                                                                                                       \
        return route_network<op_accu, T>(net, sc, io, static_cast<hipStream_t>(stream));                        \
    }                                                                                                           \
    int lue_amd_network_accu_threshold_##SUFFIX(lue_amd_network network, const T* const* inflow, T inflow_scalar, \
                                                const T* const* threshold, T threshold_scalar, T* const* outflow, \
                                                T* const* remainder, void* stream)                              \
    {                                                                                                           \
        if (!network || !outflow || !remainder)                                                                 \
            return fail(status_invalid_argument, "null argument");                                              \
        Network const& net = network->net;                                                                      \
        Scalars<T> sc{};                                                                                        \
        sc.in1_s = inflow_scalar;                                                                               \
        sc.in2_s = threshold_scalar;                                                                            \
        std::vector<CatchmentIO<T>> io;                                                                         \
        for (int c = 0; c < static_cast<int>(net.cats_host.size()); ++c)                                        \
            io.push_back(CatchmentIO<T>{entry(inflow, c), entry(threshold, c), nullptr, nullptr, nullptr, outflow[c], \
                                        remainder[c]});                                                         \
        return route_network<op_threshold, T>(net, sc, io, static_cast<hipStream_t>(stream));                   \
    }                                                                                                           \
    int lue_amd_network_accu_fraction_##SUFFIX(lue_amd_network network, const T* const* inflow, T inflow_scalar, \
                                               const T* const* fraction, T fraction_scalar, T* const* outflow,  \
                                               T* const* remainder, void* stream)                               \
    {                                                                                                           \
        if (!network || !outflow || !remainder)                                                                 \
            return fail(status_invalid_argument, "null argument");                                              \
        Network const& net = network->net;                                                                      \
        Scalars<T> sc{};                                                                                        \
        sc.in1_s = inflow_scalar;                                                                               \
        sc.in2_s = fraction_scalar;                                                                             \
        std::vector<CatchmentIO<T>> io;                                                                         \
        for (int c = 0; c < static_cast<int>(net.cats_host.size()); ++c)                                        \
            io.push_back(CatchmentIO<T>{entry(inflow, c), entry(fraction, c), nullptr, nullptr, nullptr, outflow[c], \
                                        remainder[c]});                                                         \
        return route_network<op_fraction, T>(net, sc, io, static_cast<hipStream_t>(stream));                    \
    }                                                                                                           \
    int lue_amd_network_accu_capacity_##SUFFIX(lue_amd_network network, const T* const* inflow, T inflow_scalar, \
                                               const T* const* capacity, T capacity_scalar, T* const* outflow,  \
                                               T* const* remainder, void* stream)                               \
    {                                                                                                           \
        if (!network || !outflow || !remainder)                                                                 \
            return fail(status_invalid_argument, "null argument");                                              \
        Network const& net = network->net;                                                                      \
        Scalars<T> sc{};                                                                                        \
        sc.in1_s = inflow_scalar;                                                                               \
        sc.in2_s = capacity_scalar;                                                                             \
        std::vector<CatchmentIO<T>> io;                                                                         \
        for (int c = 0; c < static_cast<int>(net.cats_host.size()); ++c)                                        \
            io.push_back(CatchmentIO<T>{entry(inflow, c), entry(capacity, c), nullptr, nullptr, nullptr, outflow[c], \
                                        remainder[c]});                                                         \
        return route_network<op_capacity, T>(net, sc, io, static_cast<hipStream_t>(stream));                    \
    }                                                                                                           \
    int lue_amd_network_accu_trigger_##SUFFIX(lue_amd_network network, const T* const* inflow, T inflow_scalar, \
                                               const T* const* trigger, T trigger_scalar, T* const* outflow,    \
                                               T* const* remainder, void* stream)                               \
    {                                                                                                           \
        if (!network || !outflow || !remainder)                                                                 \
            return fail(status_invalid_argument, "null argument");                                              \
        Network const& net = network->net;                                                                      \
        Scalars<T> sc{};                                                                                        \
        sc.in1_s = inflow_scalar;                                                                               \
        sc.in2_s = trigger_scalar;                                                                              \
        std::vector<CatchmentIO<T>> io;                                                                         \
        for (int c = 0; c < static_cast<int>(net.cats_host.size()); ++c)                                        \
            io.push_back(CatchmentIO<T>{entry(inflow, c), entry(trigger, c), nullptr, nullptr, nullptr, outflow[c], \
                                        remainder[c]});                                                         \
        return route_network<op_trigger, T>(net, sc, io, static_cast<hipStream_t>(stream));                     \
    }                                                                                                           \
    int lue_amd_network_kinematic_wave_##SUFFIX(lue_amd_network network, const T* const* current_outflow,       \
                                                const T* const* inflow, const T* const* alpha, T alpha_scalar,  \
                                                const T* const* beta, T beta_scalar, T time_step_duration,      \
                                                const T* const* channel_length, T channel_length_scalar,        \
                                                T* const* outflow, void* stream)                                \
    {                                                                                                           \
        if (!network || !outflow || !current_outflow || !inflow)                                               \
            return fail(status_invalid_argument, "null argument");                                              \
        Network const& net = network->net;                                                                      \
        Scalars<T> sc{};                                                                                        \
        sc.alpha_s = alpha_scalar;                                                                              \
        sc.beta_s = beta_scalar;                                                                                \
        sc.dx_s = channel_length_scalar;                                                                        \
        sc.dt = time_step_duration;                                                                             \
        std::vector<CatchmentIO<T>> io;                                                                         \
        for (int c = 0; c < static_cast<int>(net.cats_host.size()); ++c)                                        \
        {                                                                                                       \
            if (!current_outflow[c] || !inflow[c] || !outflow[c])                                               \
                return fail(status_invalid_argument, "null raster");                                            \
            io.push_back(CatchmentIO<T>{current_outflow[c], inflow[c], entry(alpha, c), entry(beta, c),         \
                                        entry(channel_length, c), outflow[c], nullptr});                        \
        }                                                                                                       \
        return route_network<op_kinematic_wave, T>(net, sc, io, static_cast<hipStream_t>(stream));              \
    }

LUE_AMD_NETWORK_OPS(float, f32)
LUE_AMD_NETWORK_OPS(double, f64)

}  // extern "C"
