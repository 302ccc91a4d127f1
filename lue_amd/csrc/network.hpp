// Host side of the routing sweep: flow networks (a batch of catchments with its recorded schedule) and
// the two ways to run an operation on one:
//   - record: round 0, then rounds over the tiles queued by the previous round, then after a fixed number
//     of rounds the chain-streaming tail. The rounds' tile lists and the tail's chains depend on the ldd
//     only; they are kept in the network. Host synchronisation after each round (queue length).
//   - replay: the same launches from the recorded tile lists and chains, without queueing and without
//     any host synchronisation (the whole operation is a fixed sequence of kernels on one stream).
// Per-call entry points record into a temporary network; network entry points replay.
#pragma once

#include "routing_common.hpp"
#include "sweep.hpp"
#include "tail.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace lue_amd {

    // Last error message of this thread (C ABI lue_amd_last_error).
    std::string& last_error_message();

    inline int fail(int status, std::string message)
    {
        last_error_message() = std::move(message);
        return status;
    }

    inline int hip_fail(hipError_t e, char const* what)
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

    // Tile shapes (TW x TH) the kernels are instantiated for.
    enum TileShape : int { tile_32x32 = 0, tile_64x32 = 1, tile_64x64 = 2, tile_shape_count = 3 };

    inline void tile_extent(int shape, int& tw, int& th)
    {
        tw = shape == tile_32x32 ? 32 : 64;
        th = shape == tile_64x64 ? 64 : 32;
    }

    // Timing of the sweeps with HIP events (lue_amd_profile_*).
    struct Profile
    {
        bool enabled{false};
        double first_round_ms{0.0};
        double next_rounds_ms{0.0};
        long long calls{0};
        long long rounds{0};
        long long cells{0};
    };
    Profile& profile();

    // Stream-ordered device buffer.
    template<typename P>
    struct DeviceArray
    {
        P* ptr{};
        hipStream_t stream{};
        std::size_t count{};

        DeviceArray() = default;
        DeviceArray(DeviceArray const&) = delete;
        auto operator=(DeviceArray const&) -> DeviceArray& = delete;
        ~DeviceArray() { reset(); }

        void reset()
        {
            if (ptr)
                (void)hipFreeAsync(ptr, stream);
            ptr = nullptr;
            count = 0;
        }

        hipError_t allocate(std::size_t n, hipStream_t s)
        {
            reset();
            stream = s;
            count = n;
            void* raw = nullptr;
            hipError_t const e = hipMallocAsync(&raw, std::max<std::size_t>(n, 1) * sizeof(P), s);
            ptr = static_cast<P*>(raw);
            return e;
        }
    };

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

    // Pinned host staging for small per-call tables (the per-catchment input/output pointers): a ring of
    // slots per device. A slot is reused once the copy that read it has completed (an event recorded
    // after the copy), so a call returns without waiting for the stream.
    class HostStaging
    {
    public:
        static constexpr int slot_count = 16;

        // Copy `bytes` from `src` to the device buffer `dst` in stream order.
        hipError_t copy_to_device(void* dst, void const* src, std::size_t bytes, hipStream_t stream)
        {
            int device = 0;
            hipError_t e = hipGetDevice(&device);
            if (e != hipSuccess)
                return e;
            std::lock_guard<std::mutex> lock(mutex_);
            auto& ring = rings_[device];
            Slot& s = ring.slots[ring.next];
            ring.next = (ring.next + 1) % slot_count;
            if (s.done)
            {
                if ((e = hipEventSynchronize(s.done)) != hipSuccess)
                    return e;
            }
            else if ((e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming)) != hipSuccess)
                return e;
            if (s.capacity < bytes)
            {
                if (s.host)
                    (void)hipHostFree(s.host);
                s.host = nullptr;
                s.capacity = 0;
                std::size_t const cap = std::max<std::size_t>(bytes, 4096);
                if ((e = hipHostMalloc(&s.host, cap, hipHostMallocDefault)) != hipSuccess)
                    return e;
                s.capacity = cap;
            }
            std::memcpy(s.host, src, bytes);
            if ((e = hipMemcpyAsync(dst, s.host, bytes, hipMemcpyHostToDevice, stream)) != hipSuccess)
                return e;
            return hipEventRecord(s.done, stream);
        }

    private:
        struct Slot
        {
            void* host{};
            std::size_t capacity{};
            hipEvent_t done{};
        };
        struct Ring
        {
            Slot slots[slot_count];
            int next{0};
        };
        std::mutex mutex_;
        std::map<int, Ring> rings_;
    };
    HostStaging& host_staging();

    inline unsigned blocks_for(std::uint64_t n, unsigned threads = 256)
    {
        return static_cast<unsigned>(std::max<std::uint64_t>((n + threads - 1) / threads, 1));
    }

    __global__ inline void fill_u32_kernel(std::uint32_t* p, std::uint32_t value, std::uint64_t n)
    {
        std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (i < n)
            p[i] = value;
    }

    __global__ inline void copy_i32_kernel(std::int32_t* dst, std::int32_t const* src, std::uint64_t n)
    {
        std::uint64_t const i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (i < n)
            dst[i] = src[i];
    }

    // A batch of catchments with its recorded schedule.
    struct Network
    {
        int device{0};
        int tile_shape{tile_32x32};
        int tail_after{8};  // rounds before the tail takes over (< 0: never)
        std::vector<Catchment> cats_host;
        std::int64_t ntiles{0};
        std::int64_t ncells{0};
        std::int64_t nodata_cells{0};
        DeviceArray<Catchment> cats;
        DeviceArray<std::uint32_t> tile_cat;

        bool recorded{false};
        std::vector<std::uint64_t> round_offset;  // tiles of round r (r >= 1): [round_offset[r-1], round_offset[r])
        DeviceArray<std::uint32_t> round_tiles;

        bool has_tail{false};
        std::uint32_t tail_cells{0};
        std::uint32_t nchains{0};
        std::uint32_t nready{0};
        std::uint32_t longest_chain{0};
        DeviceArray<std::uint32_t> chain_cells;
        DeviceArray<std::uint32_t> chain_base;
        DeviceArray<std::uint32_t> chain_length;
        DeviceArray<std::uint32_t> chain_cat;
        DeviceArray<std::int32_t> chain_downstream;
        DeviceArray<std::int32_t> chain_arrivals;
        DeviceArray<std::int32_t> chain_ready;

        auto rounds() const -> int { return static_cast<int>(round_offset.empty() ? 0 : round_offset.size() - 1); }
    };

    // Describe a batch: tiles are numbered catchment after catchment.
    inline int network_init(Network& net, int n, std::uint8_t const* const* ldd, std::int64_t const* rows,
                            std::int64_t const* cols, int tile_shape, int tail_after, hipStream_t stream)
    {
        if (n <= 0 || !ldd || !rows || !cols)
            return fail(status_invalid_argument, "a network needs at least one catchment");
        if (tile_shape < 0 || tile_shape >= tile_shape_count)
            return fail(status_invalid_argument, "unknown tile shape");
        int tw, th;
        tile_extent(tile_shape, tw, th);
        LUE_AMD_HIP(hipGetDevice(&net.device));
        net.tile_shape = tile_shape;
        net.tail_after = tail_after;
        net.cats_host.clear();
        std::int64_t tiles = 0;
        std::int64_t cells = 0;
        std::vector<std::uint32_t> tile_cat;
        for (int c = 0; c < n; ++c)
        {
            if (!ldd[c])
                return fail(status_invalid_argument, "ldd must not be null");
            if (rows[c] <= 0 || cols[c] <= 0)
                return fail(status_invalid_argument, "array shape must not be empty");
            if (rows[c] > (1ll << 31) || cols[c] > (1ll << 31))
                return fail(status_invalid_argument, "array extent too large");
            Catchment cat{};
            cat.ldd = ldd[c];
            cat.rows = rows[c];
            cat.cols = cols[c];
            cat.tiles_x = static_cast<int>((cols[c] + tw - 1) / tw);
            cat.tiles_y = static_cast<int>((rows[c] + th - 1) / th);
            cat.tile_base = tiles;
            cat.cell_base = cells;
            std::int64_t const ct = static_cast<std::int64_t>(cat.tiles_x) * cat.tiles_y;
            tile_cat.insert(tile_cat.end(), static_cast<std::size_t>(ct), static_cast<std::uint32_t>(c));
            tiles += ct;
            cells += rows[c] * cols[c];
            net.cats_host.push_back(cat);
        }
        if (tiles >= (1ll << 31) || cells >= (1ll << 32))
            return fail(status_invalid_argument, "batch too large (2^31 tiles, 2^32 cells)");
        net.ntiles = tiles;
        net.ncells = cells;
        LUE_AMD_HIP(net.cats.allocate(static_cast<std::size_t>(n), stream));
        LUE_AMD_HIP(net.tile_cat.allocate(tile_cat.size(), stream));
        LUE_AMD_HIP(hipMemcpyAsync(net.cats.ptr, net.cats_host.data(), sizeof(Catchment) * n, hipMemcpyHostToDevice,
                                   stream));
        LUE_AMD_HIP(hipMemcpyAsync(net.tile_cat.ptr, tile_cat.data(), sizeof(std::uint32_t) * tile_cat.size(),
                                   hipMemcpyHostToDevice, stream));
        LUE_AMD_HIP(hipStreamSynchronize(stream));  // the host vectors go out of scope
        return status_success;
    }

    // Scratch of one operation on a network.
    struct Workspace
    {
        DeviceArray<std::uint32_t> pending_bits;
        DeviceArray<std::uint32_t> border_stamp;
        DeviceArray<float> out_lo;
        DeviceArray<std::uint32_t> tile_flag;
        DeviceArray<std::uint32_t> worklist;
        DeviceArray<std::uint32_t> work_count;
        DeviceArray<unsigned long long> counters;
        DeviceArray<std::int32_t> arrivals;
        DeviceArray<std::uint32_t> cursor;
        DeviceArray<unsigned long long> tail_solved;
        DeviceArray<unsigned char> io;
    };

    template<typename T, int TW, int TH>
    int workspace_init(Network const& net, Workspace& w, SweepArgs<T>& a, bool recording, hipStream_t stream)
    {
        using G = Tile<TW, TH>;
        LUE_AMD_HIP(w.pending_bits.allocate(static_cast<std::size_t>(net.ntiles * G::words), stream));
        LUE_AMD_HIP(w.border_stamp.allocate(static_cast<std::size_t>(net.ntiles * G::perimeter), stream));
        if constexpr (std::is_same_v<T, float>)
        {
            LUE_AMD_HIP(w.out_lo.allocate(static_cast<std::size_t>(net.ncells), stream));
        }
        if (recording)
        {
            LUE_AMD_HIP(w.tile_flag.allocate(static_cast<std::size_t>(3 * net.ntiles), stream));
            LUE_AMD_HIP(w.worklist.allocate(static_cast<std::size_t>(3 * net.ntiles), stream));
            LUE_AMD_HIP(w.work_count.allocate(4, stream));
            LUE_AMD_HIP(w.counters.allocate(2, stream));
            LUE_AMD_HIP(hipMemsetAsync(w.tile_flag.ptr, 0, 3 * net.ntiles * sizeof(std::uint32_t), stream));
            LUE_AMD_HIP(hipMemsetAsync(w.work_count.ptr, 0, 4 * sizeof(std::uint32_t), stream));
            LUE_AMD_HIP(hipMemsetAsync(w.counters.ptr, 0, 2 * sizeof(unsigned long long), stream));
        }
        a.cats = net.cats.ptr;
        a.tile_cat = net.tile_cat.ptr;
        a.ncats = static_cast<int>(net.cats_host.size());
        a.ntiles = net.ntiles;
        a.out_lo = w.out_lo.ptr;
        a.pending_bits = w.pending_bits.ptr;
        a.border_stamp = w.border_stamp.ptr;
        a.tile_flag = w.tile_flag.ptr;
        a.worklist = w.worklist.ptr;
        a.work_count = w.work_count.ptr;
        a.counters = w.counters.ptr;
        a.record = recording ? 1 : 0;
        return status_success;
    }

    // Copy the per-catchment inputs and outputs to the device (the host table may go out of scope after).
    template<typename T>
    int upload_io(Workspace& w, std::vector<CatchmentIO<T>> const& io, SweepArgs<T>& a, hipStream_t stream)
    {
        LUE_AMD_HIP(w.io.allocate(sizeof(CatchmentIO<T>) * io.size(), stream));
        LUE_AMD_HIP(host_staging().copy_to_device(w.io.ptr, io.data(), sizeof(CatchmentIO<T>) * io.size(), stream));
        a.io = reinterpret_cast<CatchmentIO<T> const*>(w.io.ptr);
        return status_success;
    }

    // Build the tail's chains from the pending bitmaps and keep them in the network.
    template<int TW, int TH>
    int build_tail(Network& net, std::uint32_t const* pending_bits, hipStream_t stream)
    {
        std::int64_t const ntiles = net.ntiles;
        DeviceArray<std::uint32_t> counts, offsets;
        LUE_AMD_HIP(counts.allocate(static_cast<std::size_t>(ntiles + 1), stream));
        LUE_AMD_HIP(offsets.allocate(static_cast<std::size_t>(ntiles + 1), stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, counts.ptr + ntiles, 0u, 1ull);
        hipLaunchKernelGGL((tail_count_kernel<TW, TH>), dim3(blocks_for(ntiles)), dim3(256), 0, stream, pending_bits,
                           ntiles, counts.ptr);
        LUE_AMD_HIP(hipGetLastError());
        LUE_AMD_HIP(exclusive_scan(counts.ptr, offsets.ptr, static_cast<std::size_t>(ntiles + 1), stream));
        std::uint32_t n = 0;
        LUE_AMD_HIP(hipMemcpyAsync(&n, offsets.ptr + ntiles, sizeof(n), hipMemcpyDeviceToHost, stream));
        LUE_AMD_HIP(hipStreamSynchronize(stream));
        net.tail_cells = n;
        net.has_tail = n > 0;
        if (n == 0)
            return status_success;

        DeviceArray<std::uint32_t> cells, index_of, is_last, chain_of_last;
        DeviceArray<std::int32_t> next, upstream_count, succ[2], rank[2], last[2], first;
        LUE_AMD_HIP(cells.allocate(n, stream));
        LUE_AMD_HIP(index_of.allocate(static_cast<std::size_t>(net.ncells), stream));
        LUE_AMD_HIP(next.allocate(n, stream));
        LUE_AMD_HIP(upstream_count.allocate(n, stream));
        for (int b = 0; b < 2; ++b)
        {
            LUE_AMD_HIP(succ[b].allocate(n, stream));
            LUE_AMD_HIP(rank[b].allocate(n, stream));
            LUE_AMD_HIP(last[b].allocate(n, stream));
        }
        LUE_AMD_HIP(is_last.allocate(n + 1, stream));
        LUE_AMD_HIP(chain_of_last.allocate(n + 1, stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(n)), dim3(256), 0, stream,
                           reinterpret_cast<std::uint32_t*>(upstream_count.ptr), 0u, static_cast<std::uint64_t>(n));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, is_last.ptr + n, 0u, 1ull);

        int const ncats = static_cast<int>(net.cats_host.size());
        hipLaunchKernelGGL((tail_compact_kernel<TW, TH>), dim3(static_cast<unsigned>(ntiles)), dim3(64), 0, stream,
                           pending_bits, net.cats.ptr, net.tile_cat.ptr, offsets.ptr, cells.ptr, index_of.ptr);
        hipLaunchKernelGGL((tail_links_kernel<TW, TH>), dim3(blocks_for(n)), dim3(256), 0, stream, net.cats.ptr, ncats,
                           pending_bits, cells.ptr, index_of.ptr, n, next.ptr, upstream_count.ptr);
        hipLaunchKernelGGL(tail_rank_init_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, next.ptr,
                           upstream_count.ptr, n, succ[0].ptr, rank[0].ptr, last[0].ptr);
        int cur = 0;
        for (std::uint32_t reach = 1; reach < n; reach <<= 1)
        {
            hipLaunchKernelGGL(tail_rank_step_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, succ[cur].ptr,
                               rank[cur].ptr, last[cur].ptr, n, succ[1 - cur].ptr, rank[1 - cur].ptr, last[1 - cur].ptr);
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
            return fail(status_invalid_network, "invalid flow direction network: cells on a cycle");
        net.nchains = nchains;

        DeviceArray<std::uint32_t> ready_count;
        LUE_AMD_HIP(net.chain_length.allocate(nchains + 1, stream));
        LUE_AMD_HIP(net.chain_base.allocate(nchains + 1, stream));
        LUE_AMD_HIP(net.chain_cat.allocate(nchains, stream));
        LUE_AMD_HIP(net.chain_downstream.allocate(nchains, stream));
        LUE_AMD_HIP(net.chain_arrivals.allocate(nchains, stream));
        LUE_AMD_HIP(net.chain_ready.allocate(nchains, stream));
        LUE_AMD_HIP(net.chain_cells.allocate(n, stream));
        LUE_AMD_HIP(first.allocate(nchains, stream));
        LUE_AMD_HIP(ready_count.allocate(1, stream));
        // a chain whose first cell is never found (a cycle without entry) keeps length 0 and never gets ready
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(nchains + 1)), dim3(256), 0, stream,
                           net.chain_length.ptr, 0u, static_cast<std::uint64_t>(nchains + 1));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream,
                           reinterpret_cast<std::uint32_t*>(net.chain_arrivals.ptr), 0xFFFFFFFFu,
                           static_cast<std::uint64_t>(nchains));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream, net.chain_cat.ptr, 0u,
                           static_cast<std::uint64_t>(nchains));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, ready_count.ptr, 0u, 1ull);
        hipLaunchKernelGGL(tail_chain_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, next.ptr, upstream_count.ptr,
                           rank[cur].ptr, last[cur].ptr, chain_of_last.ptr, cells.ptr, net.cats.ptr, ncats, n,
                           net.chain_length.ptr, first.ptr, net.chain_arrivals.ptr, net.chain_cat.ptr);
        LUE_AMD_HIP(exclusive_scan(net.chain_length.ptr, net.chain_base.ptr, static_cast<std::size_t>(nchains + 1),
                                   stream));
        hipLaunchKernelGGL(tail_downstream_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, next.ptr, last[cur].ptr,
                           chain_of_last.ptr, is_last.ptr, n, net.chain_downstream.ptr);
        hipLaunchKernelGGL(tail_layout_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, cells.ptr, rank[cur].ptr,
                           last[cur].ptr, is_last.ptr, chain_of_last.ptr, net.chain_base.ptr, net.chain_length.ptr, n,
                           net.chain_cells.ptr);
        hipLaunchKernelGGL(tail_ready_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream, net.chain_arrivals.ptr,
                           nchains, net.chain_ready.ptr, ready_count.ptr);
        LUE_AMD_HIP(hipGetLastError());
        LUE_AMD_HIP(hipMemcpyAsync(&net.nready, ready_count.ptr, sizeof(std::uint32_t), hipMemcpyDeviceToHost, stream));
        LUE_AMD_HIP(hipStreamSynchronize(stream));
        return status_success;
    }

    // Run the tail of one operation from the network's chains.
    template<int OP, typename T>
    int launch_tail(Network const& net, Workspace& w, SweepArgs<T> const& a, hipStream_t stream, bool count_solved)
    {
        if (!net.has_tail || net.nready == 0)
            return status_success;
        LUE_AMD_HIP(w.arrivals.allocate(net.nchains, stream));
        LUE_AMD_HIP(w.cursor.allocate(1, stream));
        LUE_AMD_HIP(w.tail_solved.allocate(1, stream));
        hipLaunchKernelGGL(copy_i32_kernel, dim3(blocks_for(net.nchains)), dim3(256), 0, stream, w.arrivals.ptr,
                           net.chain_arrivals.ptr, static_cast<std::uint64_t>(net.nchains));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, w.cursor.ptr, 0u, 1ull);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream,
                           reinterpret_cast<std::uint32_t*>(w.tail_solved.ptr), 0u, 2ull);
        TailChains t{net.chain_cells.ptr, net.chain_cat.ptr, net.chain_base.ptr, net.chain_length.ptr,
                     net.chain_downstream.ptr, w.arrivals.ptr, net.chain_ready.ptr, net.nready, w.cursor.ptr,
                     w.tail_solved.ptr};
        (void)count_solved;
        // one wavefront per chain, 4 per workgroup
        unsigned const grid = static_cast<unsigned>(std::min<std::uint64_t>((net.nready + 3) / 4, 2048));
        hipLaunchKernelGGL((tail_execute_kernel<OP, T>), dim3(std::max(grid, 1u)), dim3(256), 0, stream, a, t);
        LUE_AMD_HIP(hipGetLastError());
        return status_success;
    }

    // Record: run the operation while recording the schedule into `net`.
    template<int OP, typename T, int TW, int TH>
    int record_run(Network& net, SweepArgs<T> a, std::vector<CatchmentIO<T>> const& io, hipStream_t stream)
    {
        Workspace w;
        if (int const st = workspace_init<T, TW, TH>(net, w, a, true, stream))
            return st;
        if (int const st = upload_io<T>(w, io, a, stream))
            return st;
        hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
        if (profile().enabled)
        {
            for (auto& e : ev)
                LUE_AMD_HIP(hipEventCreate(&e));
            LUE_AMD_HIP(hipEventRecord(ev[0], stream));
        }
        hipLaunchKernelGGL((sweep_first_round<OP, T, TW, TH>), dim3(static_cast<unsigned>(net.ntiles)), dim3(256), 0,
                           stream, a);
        LUE_AMD_HIP(hipGetLastError());
        if (profile().enabled)
            LUE_AMD_HIP(hipEventRecord(ev[1], stream));

        net.round_offset.assign(1, 0);
        std::vector<std::uint32_t> tiles_of_rounds;
        unsigned const grid = static_cast<unsigned>(std::min<std::int64_t>(net.ntiles, 2048));
        std::uint32_t round = 1;
        bool tail = false;
        while (true)
        {
            // the tiles queued for this round
            std::uint32_t const slot = round % 3;
            std::uint32_t count = 0;
            LUE_AMD_HIP(hipMemcpyAsync(&count, w.work_count.ptr + slot, sizeof(count), hipMemcpyDeviceToHost, stream));
            LUE_AMD_HIP(hipStreamSynchronize(stream));
            if (count == 0)
                break;
            if (net.tail_after >= 0 && static_cast<int>(round) > net.tail_after)
            {
                tail = true;
                break;
            }
            std::size_t const at = tiles_of_rounds.size();
            tiles_of_rounds.resize(at + count);
            LUE_AMD_HIP(hipMemcpyAsync(tiles_of_rounds.data() + at, w.worklist.ptr + static_cast<std::int64_t>(slot) * net.ntiles,
                                       sizeof(std::uint32_t) * count, hipMemcpyDeviceToHost, stream));
            hipLaunchKernelGGL((sweep_next_round<OP, T, TW, TH>), dim3(std::min<unsigned>(grid, count)), dim3(256), 0,
                               stream, a, round);
            LUE_AMD_HIP(hipGetLastError());
            net.round_offset.push_back(tiles_of_rounds.size());
            ++round;
            if (round > 0xFFFF0000u)
                return fail(status_invalid_network, "routing did not converge");
        }
        LUE_AMD_HIP(hipStreamSynchronize(stream));
        LUE_AMD_HIP(net.round_tiles.allocate(tiles_of_rounds.size(), stream));
        if (!tiles_of_rounds.empty())
        {
            LUE_AMD_HIP(hipMemcpyAsync(net.round_tiles.ptr, tiles_of_rounds.data(),
                                       sizeof(std::uint32_t) * tiles_of_rounds.size(), hipMemcpyHostToDevice, stream));
        }
        net.has_tail = false;
        net.nready = 0;
        if (tail)
        {
            if (int const st = build_tail<TW, TH>(net, w.pending_bits.ptr, stream))
                return st;
            if (int const st = launch_tail<OP, T>(net, w, a, stream, true))
                return st;
        }
        if (profile().enabled)
            LUE_AMD_HIP(hipEventRecord(ev[2], stream));
        unsigned long long counters[2] = {0, 0};
        unsigned long long tail_solved = 0;
        LUE_AMD_HIP(hipMemcpyAsync(counters, w.counters.ptr, sizeof(counters), hipMemcpyDeviceToHost, stream));
        if (net.has_tail && net.nready > 0)
            LUE_AMD_HIP(hipMemcpyAsync(&tail_solved, w.tail_solved.ptr, sizeof(tail_solved), hipMemcpyDeviceToHost, stream));
        LUE_AMD_HIP(hipStreamSynchronize(stream));
        if (profile().enabled)
        {
            float first = 0.0f, rest = 0.0f;
            LUE_AMD_HIP(hipEventElapsedTime(&first, ev[0], ev[1]));
            LUE_AMD_HIP(hipEventElapsedTime(&rest, ev[1], ev[2]));
            profile().first_round_ms += first;
            profile().next_rounds_ms += rest;
            profile().calls += 1;
            profile().rounds += round;
            profile().cells += net.ncells;
            for (auto& e : ev)
                (void)hipEventDestroy(e);
        }
        net.nodata_cells = static_cast<std::int64_t>(counters[1]);
        std::uint64_t const total = counters[0] + counters[1] + tail_solved;
        if (total != static_cast<std::uint64_t>(net.ncells))
        {
            return fail(status_invalid_network,
                        "invalid flow direction network: " + std::to_string(net.ncells - static_cast<std::int64_t>(total)) +
                            " cells lie on or downstream of a cycle");
        }
        net.recorded = true;
        return status_success;
    }

    // Replay: the operation as a fixed launch sequence from the recorded schedule (no host synchronisation).
    template<int OP, typename T, int TW, int TH>
    int replay_run(Network const& net, SweepArgs<T> a, std::vector<CatchmentIO<T>> const& io, hipStream_t stream)
    {
        if (!net.recorded)
            return fail(status_invalid_argument, "network has no recorded schedule");
        Workspace w;
        if (int const st = workspace_init<T, TW, TH>(net, w, a, false, stream))
            return st;
        if (int const st = upload_io<T>(w, io, a, stream))
            return st;
        hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
        if (profile().enabled)
        {
            for (auto& e : ev)
                LUE_AMD_HIP(hipEventCreate(&e));
            LUE_AMD_HIP(hipEventRecord(ev[0], stream));
        }
        hipLaunchKernelGGL((sweep_first_round<OP, T, TW, TH>), dim3(static_cast<unsigned>(net.ntiles)), dim3(256), 0,
                           stream, a);
        if (profile().enabled)
            LUE_AMD_HIP(hipEventRecord(ev[1], stream));
        for (int r = 1; r <= net.rounds(); ++r)
        {
            std::uint64_t const begin = net.round_offset[r - 1];
            std::uint32_t const count = static_cast<std::uint32_t>(net.round_offset[r] - begin);
            unsigned const grid = std::min<unsigned>(count, 2048);
            hipLaunchKernelGGL((sweep_list_round<OP, T, TW, TH>), dim3(grid), dim3(256), 0, stream, a,
                               net.round_tiles.ptr + begin, count, static_cast<std::uint32_t>(r));
        }
        LUE_AMD_HIP(hipGetLastError());
        if (int const st = launch_tail<OP, T>(net, w, a, stream, false))
            return st;
        if (profile().enabled)
        {
            LUE_AMD_HIP(hipEventRecord(ev[2], stream));
            LUE_AMD_HIP(hipEventSynchronize(ev[2]));
            float first = 0.0f, rest = 0.0f;
            LUE_AMD_HIP(hipEventElapsedTime(&first, ev[0], ev[1]));
            LUE_AMD_HIP(hipEventElapsedTime(&rest, ev[1], ev[2]));
            profile().first_round_ms += first;
            profile().next_rounds_ms += rest;
            profile().calls += 1;
            profile().rounds += net.rounds() + 1;
            profile().cells += net.ncells;
            for (auto& e : ev)
                (void)hipEventDestroy(e);
        }
        return status_success;
    }

}  // namespace lue_amd
