// Host-side plumbing shared by the translation units of the C ABI (include/lue_amd/routing.h): argument
// checks, the scalars of an operation, the per-call network cache, and the drivers that run one operation
// on one raster (record or replay) or on a flow network (replay). Each op_*.hip file instantiates these for
// one operation, so the library's kernels compile in parallel.
#pragma once

#include "lue_amd/routing.h"

#include "network.hpp"
#include "routing_common.hpp"
#include "sweep.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <list>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

namespace lue_amd {

    // Schedule of the per-call entry points (lue_amd_set_sweep_options).
    int& default_schedule();

    template<int OP, typename T, int TW, int TH>
    constexpr bool fits()
    {
        return sizeof(SweepLds<OP, T, TW, TH>) <= 160 * 1024 && sizeof(TilePassLds<OP, T, TW, TH>) <= 160 * 1024;
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
            case tile_32x64:
                if constexpr (fits<OP, T, 32, 64>())
                    return f.template operator()<32, 64>();
                break;
            default:
                break;
        }
        return fail(status_invalid_argument, "tile shape does not fit this operation in LDS");
    }

    // Tile shape of per-call entry points, measured at 16384^2 cells on the synthetic DEM. The kinematic wave's
    // tile pass is bound by vector instruction issue and needs the occupancy of the smallest tiles (32x32:
    // 6.1 ms FP32, 12.8 ms FP64; 64x32: 7.4 and 17.1 ms). The accu family's tile pass is bound by latency and
    // memory, and larger tiles leave fewer cells to the chain phase (83 M pending at 32x32, 79 M at 64x32):
    // 64x32 runs accu in 2.59 ms FP32 and 4.04 ms FP64 (32x32: 2.81, 4.36), threshold, fraction and capacity
    // in 4.5 and 6.6 ms (4.8, 8.4).
    template<int OP>
    constexpr int default_tile_shape()
    {
        return OP == op_kinematic_wave ? tile_32x32 : tile_64x32;
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
        a.c1_s = sc.dt / sc.dx_s;  // IEEE division in T, as the device's a.dt / dx
        return a;
    }

    // Networks of the per-call entry points, most recently used first. A network is found again by device,
    // ldd pointer, shape, tile shape, schedule and a fingerprint of the ldd bytes (computed in every call),
    // so a changed ldd is recorded anew.
    struct CacheKey
    {
        int device;
        void const* ldd;
        std::int64_t rows;
        std::int64_t cols;
        int tile_shape;
        int schedule;
        std::uint64_t fingerprint[2];

        bool operator==(CacheKey const&) const = default;
    };

    class NetworkCache
    {
    public:
        std::shared_ptr<Network> find(CacheKey const& key)
        {
            std::lock_guard<std::mutex> lock(mutex_);
            for (auto it = entries_.begin(); it != entries_.end(); ++it)
            {
                if (it->first == key)
                {
                    entries_.splice(entries_.begin(), entries_, it);
                    return it->second;
                }
            }
            return nullptr;
        }

        void insert(CacheKey const& key, std::shared_ptr<Network> net)
        {
            std::vector<std::shared_ptr<Network>> dropped;
            {
                std::lock_guard<std::mutex> lock(mutex_);
                entries_.emplace_front(key, std::move(net));
                trim_locked(dropped);
            }
            // a dropped network is destroyed here, outside the lock, once no caller holds it (~Network waits
            // for the device before it frees its buffers)
        }

        int capacity()
        {
            std::lock_guard<std::mutex> lock(mutex_);
            return capacity_;
        }

        void set_capacity(int entries)
        {
            std::vector<std::shared_ptr<Network>> dropped;
            std::lock_guard<std::mutex> lock(mutex_);
            capacity_ = entries;
            trim_locked(dropped);
        }

    private:
        void trim_locked(std::vector<std::shared_ptr<Network>>& dropped)
        {
            while (static_cast<int>(entries_.size()) > capacity_)
            {
                dropped.push_back(std::move(entries_.back().second));
                entries_.pop_back();
            }
        }

        std::mutex mutex_;
        std::list<std::pair<CacheKey, std::shared_ptr<Network>>> entries_;
        int capacity_ = 4;
    };

    NetworkCache& network_cache();

    // 128-bit fingerprint of the ldd bytes (fingerprint_kernel); waits for it.
    int ldd_fingerprint(std::uint8_t const* ldd, std::int64_t n, hipStream_t stream, std::uint64_t (&fingerprint)[2]);

    // One operation on one raster: replay a cached network of the ldd, or record one.
    template<int OP, typename T>
    int route_once(std::uint8_t const* ldd, std::int64_t rows, std::int64_t cols, Scalars<T> const& sc,
                   CatchmentIO<T> const& io, hipStream_t stream)
    {
        std::vector<CatchmentIO<T>> table{io};
        int const shape = default_tile_shape<OP>();
        NetworkCache& cache = network_cache();
        CacheKey key{0, ldd, rows, cols, shape, default_schedule(), {0, 0}};
        bool const caching = cache.capacity() > 0;
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
        if (int const st = network_init(*net, 1, &ldd, &rows, &cols, shape, default_schedule(), stream))
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

    inline int check_shape(void const* ldd, std::int64_t rows, std::int64_t cols)
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
                            T const* inflow, T const* alpha, T alpha_s, T const* beta, T beta_s, T dt, T const* dx,
                            T dx_s, T* outflow, void* stream)
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

    // Per-catchment pointer tables of the network entry points (host arrays of device pointers; a null table
    // means the scalar).
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

    // Capture one operation on a network into a plan (network.hpp: Plan).
    template<int OP, typename T>
    int plan_network(Network const& net, Scalars<T> const& sc, std::vector<CatchmentIO<T>> const& io, Plan& plan)
    {
        return with_tile_shape<OP, T>(net.tile_shape, [&]<int TW, int TH>()
                                      { return plan_capture<OP, T, TW, TH>(net, args_from(sc), io, plan); });
    }

    // Network operations with two outputs: every catchment needs both output rasters.
    // plan == nullptr: run on `stream`; else capture into *plan.
    template<int OP, typename T>
    int network_split(Network const& net, T const* const* inflow, T inflow_s, T const* const* param, T param_s,
                      T* const* outflow, T* const* remainder, hipStream_t stream, Plan* plan = nullptr)
    {
        if (!outflow || !remainder)
            return fail(status_invalid_argument, "outflow and remainder must not be null");
        Scalars<T> sc{};
        sc.in1_s = inflow_s;
        sc.in2_s = param_s;
        std::vector<CatchmentIO<T>> io;
        for (int c = 0; c < static_cast<int>(net.cats_host.size()); ++c)
        {
            if (!outflow[c] || !remainder[c])
                return fail(status_invalid_argument, "outflow and remainder must not be null");
            io.push_back(CatchmentIO<T>{entry(inflow, c), entry(param, c), nullptr, nullptr, nullptr, outflow[c],
                                        remainder[c]});
        }
        return plan ? plan_network<OP, T>(net, sc, io, *plan) : route_network<OP, T>(net, sc, io, stream);
    }

    // Wrap a capture into a new C ABI plan handle.
    template<typename F>
    int make_plan(lue_amd_plan* out, F&& capture);

}  // namespace lue_amd

struct lue_amd_network_s
{
    lue_amd::Network net;
};

struct lue_amd_plan_s
{
    lue_amd::Plan plan;
};

template<typename F>
int lue_amd::make_plan(lue_amd_plan* out, F&& capture)
{
    if (!out)
        return fail(status_invalid_argument, "plan must not be null");
    *out = nullptr;
    auto handle = std::make_unique<lue_amd_plan_s>();
    if (int const st = capture(handle->plan))
        return st;
    *out = handle.release();
    return status_success;
}
