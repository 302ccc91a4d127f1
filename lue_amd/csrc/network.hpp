// Host side of the routing sweep: flow networks (a batch of catchments with its recorded schedule) and
// the two ways to run an operation on one:
//   - record: round 0 of the tiled sweep, then, depending on the schedule,
//       chains (default): the heavy-path chain layout of the cells round 0 leaves pending (chains.hpp);
//       rounds: rounds over the tiles queued by the previous round until no tile is left;
//     both depend on the ldd only and are kept in the network. Host synchronisation while recording.
//   - replay: the same launches from the recorded schedule, without any host synchronisation (the whole
//     operation is a fixed sequence of kernels on one stream).
// Per-call entry points record into a temporary network; network entry points replay.
#pragma once

#include "routing_common.hpp"
#include "sweep.hpp"
#include "tile_pass.hpp"
#include "chains.hpp"
#include "chains_wide.hpp"

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
    enum TileShape : int { tile_32x32 = 0, tile_64x32 = 1, tile_64x64 = 2, tile_32x64 = 3, tile_shape_count = 4 };

    inline void tile_extent(int shape, int& tw, int& th)
    {
        tw = shape == tile_32x32 || shape == tile_32x64 ? 32 : 64;
        th = shape == tile_64x64 || shape == tile_32x64 ? 64 : 32;
    }

    // Timing of the sweeps with HIP events (lue_amd_profile_*). A profiled operation records three events on
    // its stream (start, end of round 0, end) without waiting for them; reading the profile waits for the
    // events recorded so far and sums their times, so profiling adds no host synchronisation to the calls.
    class Profile
    {
    public:
        struct Totals
        {
            double first_round_ms{0.0};
            double next_rounds_ms{0.0};
            double first_round_min_ms{0.0};
            double first_round_max_ms{0.0};
            long long calls{0};
            long long rounds{0};
            long long cells{0};
        };

        bool enabled()
        {
            std::lock_guard<std::mutex> lock(mutex_);
            return enabled_;
        }

        // Reset the totals and start (or stop) collecting.
        hipError_t reset(bool enable)
        {
            std::lock_guard<std::mutex> lock(mutex_);
            hipError_t const e = resolve_locked();
            totals_ = Totals{};
            for (auto& t : by_kind_)
                t = Totals{};
            enabled_ = enable;
            return e;
        }

        // Three events for one operation (from the pool).
        hipError_t take(hipEvent_t (&ev)[3])
        {
            std::lock_guard<std::mutex> lock(mutex_);
            for (auto& e : ev)
            {
                if (!pool_.empty())
                {
                    e = pool_.back();
                    pool_.pop_back();
                }
                else if (hipError_t const err = hipEventCreate(&e); err != hipSuccess)
                {
                    e = nullptr;
                    return err;
                }
            }
            return hipSuccess;
        }

        void give_back(hipEvent_t (&ev)[3])
        {
            std::lock_guard<std::mutex> lock(mutex_);
            for (auto& e : ev)
            {
                if (e)
                    pool_.push_back(e);
                e = nullptr;
            }
        }

        // The events of one operation (kind = operation * 2 + 1 for FP64), all recorded.
        void finish(hipEvent_t (&ev)[3], int kind, long long rounds, std::int64_t cells)
        {
            std::lock_guard<std::mutex> lock(mutex_);
            pending_.push_back(Pending{{ev[0], ev[1], ev[2]}, kind, rounds, cells});
            for (auto& e : ev)
                e = nullptr;
        }

        // Totals of all operations (kind < 0) or of one kind.
        hipError_t read(int kind, Totals& out)
        {
            std::lock_guard<std::mutex> lock(mutex_);
            hipError_t const e = resolve_locked();
            out = kind < 0 ? totals_ : (kind < kinds ? by_kind_[kind] : Totals{});
            return e;
        }

        static constexpr int kinds = 12;

    private:
        struct Pending
        {
            hipEvent_t ev[3];
            int kind;
            long long rounds;
            std::int64_t cells;
        };

        static void add(Totals& t, float first, float rest, Pending const& p)
        {
            t.first_round_min_ms = t.calls == 0 ? first : std::min<double>(t.first_round_min_ms, first);
            t.first_round_max_ms = std::max<double>(t.first_round_max_ms, first);
            t.first_round_ms += first;
            t.next_rounds_ms += rest;
            t.calls += 1;
            t.rounds += p.rounds;
            t.cells += p.cells;
        }

        hipError_t resolve_locked()
        {
            hipError_t result = hipSuccess;
            for (auto& p : pending_)
            {
                float first = 0.0f, rest = 0.0f;
                hipError_t e = hipEventSynchronize(p.ev[2]);
                if (e == hipSuccess)
                    e = hipEventElapsedTime(&first, p.ev[0], p.ev[1]);
                if (e == hipSuccess)
                    e = hipEventElapsedTime(&rest, p.ev[1], p.ev[2]);
                if (e == hipSuccess)
                {
                    add(totals_, first, rest, p);
                    if (p.kind >= 0 && p.kind < kinds)
                        add(by_kind_[p.kind], first, rest, p);
                }
                else if (result == hipSuccess)
                {
                    result = e;
                }
                for (auto& ev : p.ev)
                    pool_.push_back(ev);
            }
            pending_.clear();
            return result;
        }

        std::mutex mutex_;
        bool enabled_{false};
        Totals totals_;
        Totals by_kind_[kinds];
        std::vector<Pending> pending_;
        std::vector<hipEvent_t> pool_;
    };
    Profile& profile();

    // Frees of stream-ordered buffers held back and issued together later: the frees of the chain layout's
    // temporaries that end it (about 13 ms of host time at 16384^2) are issued while the in-tile layout kernel
    // runs instead of before it is launched. While a thread's `deferred_frees` is set, DeviceArray::reset
    // queues its buffer there.
    struct DeferredFrees
    {
        std::vector<std::pair<void*, hipStream_t>> list;

        DeferredFrees() = default;
        DeferredFrees(DeferredFrees const&) = delete;
        auto operator=(DeferredFrees const&) -> DeferredFrees& = delete;
        ~DeferredFrees() { flush(); }

        void flush()
        {
            for (auto const& [p, s] : list)
                (void)hipFreeAsync(p, s);
            list.clear();
        }
    };

    inline thread_local DeferredFrees* deferred_frees = nullptr;

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
            {
                if (deferred_frees)
                    deferred_frees->list.emplace_back(ptr, stream);
                else
                    (void)hipFreeAsync(ptr, stream);
            }
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

    // Device blocks for the per-call workspaces, kept for reuse by later calls instead of a stream-ordered
    // malloc and free per call (hipFreeAsync has taken tens of milliseconds at times, stalling a caller that
    // serves many small calls). A released block records an event on the releasing stream; it is handed
    // out again on that stream at once (stream order), on another stream once the event has completed.
    // Best fit among the free blocks of at most twice (plus 4 MiB) the request; new blocks come from
    // hipMalloc, rounded up to 256 KiB. Blocks stay allocated until trim() (an allocation that fails frees the
    // idle blocks and tries again).
    class ScratchCache
    {
    public:
        struct Block
        {
            void* ptr{};
            std::size_t bytes{};
            int device{};
            hipStream_t stream{};
            hipEvent_t done{};
            bool busy{};
        };

        hipError_t acquire(std::size_t bytes, hipStream_t stream, Block*& out)
        {
            out = nullptr;
            int device = 0;
            hipError_t e = hipGetDevice(&device);
            if (e != hipSuccess)
                return e;
            bytes = std::max<std::size_t>(bytes, 1);
            std::size_t const limit = 2 * bytes + (std::size_t{4} << 20);
            std::lock_guard<std::mutex> lock(mutex_);
            Block* best = nullptr;
            for (auto& b : blocks_)
            {
                if (b->busy || b->device != device || b->bytes < bytes || b->bytes > limit)
                    continue;
                if (best && b->bytes >= best->bytes)
                    continue;
                if (b->stream != stream && hipEventQuery(b->done) != hipSuccess)
                    continue;
                best = b.get();
            }
            if (!best)
            {
                auto b = std::make_unique<Block>();
                b->bytes = (bytes + block_granule - 1) / block_granule * block_granule;
                b->device = device;
                e = hipMalloc(&b->ptr, b->bytes);
                if (e == hipErrorOutOfMemory)
                {
                    (void)hipGetLastError();
                    if ((e = trim_locked(device)) != hipSuccess)
                        return e;
                    e = hipMalloc(&b->ptr, b->bytes);
                }
                if (e != hipSuccess)
                    return e;
                if ((e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming)) != hipSuccess)
                {
                    (void)hipFree(b->ptr);
                    return e;
                }
                best = b.get();
                blocks_.push_back(std::move(b));
            }
            best->busy = true;
            best->stream = stream;
            out = best;
            return hipSuccess;
        }

        void release(Block* b, hipStream_t stream)
        {
            std::lock_guard<std::mutex> lock(mutex_);
            int current = 0;
            bool const switched = hipGetDevice(&current) == hipSuccess && current != b->device &&
                                  hipSetDevice(b->device) == hipSuccess;
            (void)hipEventRecord(b->done, stream);
            if (switched)
                (void)hipSetDevice(current);
            b->stream = stream;
            b->busy = false;
        }

        // Free the idle blocks of the current device (waits for their last use).
        hipError_t trim()
        {
            int device = 0;
            hipError_t const e = hipGetDevice(&device);
            if (e != hipSuccess)
                return e;
            std::lock_guard<std::mutex> lock(mutex_);
            return trim_locked(device);
        }

    private:
        static constexpr std::size_t block_granule = std::size_t{256} << 10;

        hipError_t trim_locked(int device)
        {
            hipError_t result = hipSuccess;
            std::vector<std::unique_ptr<Block>> kept;
            for (auto& b : blocks_)
            {
                if (b->busy || b->device != device)
                {
                    kept.push_back(std::move(b));
                    continue;
                }
                hipError_t e = hipEventSynchronize(b->done);
                if (e == hipSuccess)
                    e = hipEventDestroy(b->done);
                if (e == hipSuccess)
                    e = hipFree(b->ptr);
                if (e != hipSuccess && result == hipSuccess)
                    result = e;
            }
            blocks_ = std::move(kept);
            return result;
        }

        std::mutex mutex_;
        std::vector<std::unique_ptr<Block>> blocks_;
    };
    ScratchCache& scratch_cache();

    // Per-call device buffer from the scratch cache (same interface as DeviceArray).
    template<typename P>
    struct ScratchArray
    {
        P* ptr{};
        hipStream_t stream{};
        std::size_t count{};
        ScratchCache::Block* block{};

        ScratchArray() = default;
        ScratchArray(ScratchArray const&) = delete;
        auto operator=(ScratchArray const&) -> ScratchArray& = delete;
        ~ScratchArray() { reset(); }

        void reset()
        {
            if (block)
                scratch_cache().release(block, stream);
            block = nullptr;
            ptr = nullptr;
            count = 0;
        }

        hipError_t allocate(std::size_t n, hipStream_t s)
        {
            reset();
            stream = s;
            hipError_t const e = scratch_cache().acquire(std::max<std::size_t>(n, 1) * sizeof(P), s, block);
            if (e != hipSuccess)
                return e;
            count = n;
            ptr = static_cast<P*>(block->ptr);
            return hipSuccess;
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
        static constexpr std::size_t initial_capacity = 4096;

        // Copy `bytes` from `src` to the device buffer `dst` in stream order.
        hipError_t copy_to_device(void* dst, void const* src, std::size_t bytes, hipStream_t stream)
        {
            int device = 0;
            hipError_t e = hipGetDevice(&device);
            if (e != hipSuccess)
                return e;
            std::lock_guard<std::mutex> lock(mutex_);
            auto& ring = rings_[device];
            if (!ring.ready)
            {
                // every slot at once on first use: pinned allocations and event creation stay out of the
                // calls that follow (a pinned allocation can take milliseconds)
                for (Slot& x : ring.slots)
                {
                    if (!x.done && (e = hipEventCreateWithFlags(&x.done, hipEventDisableTiming)) != hipSuccess)
                        return e;
                    if (!x.host)
                    {
                        if ((e = hipHostMalloc(&x.host, initial_capacity, hipHostMallocDefault)) != hipSuccess)
                            return e;
                        x.capacity = initial_capacity;
                    }
                }
                ring.ready = true;
            }
            Slot& s = ring.slots[ring.next];
            ring.next = (ring.next + 1) % slot_count;
            if (s.used)
            {
                if ((e = hipEventSynchronize(s.done)) != hipSuccess)
                    return e;
            }
            s.used = true;
            if (s.capacity < bytes)
            {
                if (s.host)
                    (void)hipHostFree(s.host);
                s.host = nullptr;
                s.capacity = 0;
                std::size_t const cap = std::max<std::size_t>(bytes, initial_capacity);
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
            bool used{false};  // `done` marks a copy that read this slot
        };
        struct Ring
        {
            Slot slots[slot_count];
            int next{0};
            bool ready{false};
        };
        std::mutex mutex_;
        std::map<int, Ring> rings_;
    };
    HostStaging& host_staging();

    template<typename In, typename Out>
    hipError_t inclusive_scan(In const* in, Out* out, std::size_t n, hipStream_t stream)
    {
        std::size_t bytes = 0;
        hipError_t e = hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out, n, stream);
        if (e != hipSuccess)
            return e;
        void* temp = nullptr;
        e = hipMallocAsync(&temp, std::max<std::size_t>(bytes, 1), stream);
        if (e != hipSuccess)
            return e;
        e = hipcub::DeviceScan::InclusiveSum(temp, bytes, in, out, n, stream);
        hipError_t const f = hipFreeAsync(temp, stream);
        return e != hipSuccess ? e : f;
    }

    // Stable radix sort of (key, value) pairs on the low end_bit bits of the keys.
    template<typename Key, typename Value>
    hipError_t sort_pairs(Key const* keys_in, Key* keys_out, Value const* values_in, Value* values_out, std::size_t n,
                          int end_bit, hipStream_t stream)
    {
        std::size_t bytes = 0;
        hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out, values_in, values_out,
                                                          static_cast<int>(n), 0, end_bit, stream);
        if (e != hipSuccess)
            return e;
        void* temp = nullptr;
        e = hipMallocAsync(&temp, std::max<std::size_t>(bytes, 1), stream);
        if (e != hipSuccess)
            return e;
        e = hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys_in, keys_out, values_in, values_out,
                                               static_cast<int>(n), 0, end_bit, stream);
        hipError_t const f = hipFreeAsync(temp, stream);
        return e != hipSuccess ? e : f;
    }

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

    // Schedules of the sweep after round 0.
    enum Schedule : int {
        schedule_chains = 0,  // round 0, then the pending cells as heavy-path chains, level after level
        schedule_rounds = 1,  // rounds over the tiles that received flow, until none is left
        schedule_count = 2
    };

    // A batch of catchments with its recorded schedule.
    struct Network
    {
        int device{0};
        int tile_shape{tile_32x32};
        int schedule{schedule_chains};
        std::vector<Catchment> cats_host;
        std::int64_t ntiles{0};
        std::int64_t ncells{0};
        std::int64_t nodata_cells{0};
        DeviceArray<Catchment> cats;
        DeviceArray<std::uint32_t> tile_cat;

        bool recorded{false};

        // rounds schedule: tiles of round r (r >= 1) are [round_offset[r-1], round_offset[r])
        std::vector<std::uint64_t> round_offset;
        DeviceArray<std::uint32_t> round_tiles;

        // chains schedule
        std::uint32_t pending{0};         // cells left pending by round 0
        std::uint32_t nchains{0};
        std::uint32_t longest_chain{0};
        std::uint64_t nrefs{0};
        // chain levels whose longest work item has at least this many positions run on the wide path (0: none)
        std::uint32_t wide_min{wide_min_positions};
        // work items of a chain level with at most this many positions run one per lane (0: none)
        std::uint32_t lane_max{lane_max_positions};
        // the tile pass as a persistent launch fed by a tile work queue (tile_pass_queue_kernel)
        bool tile_queue{false};
        int cus{1};  // compute units of the device
        std::uint64_t tile_blocks{0};           // 64-position blocks of the tile pass (in-tile layout)
        std::uint64_t tile_critical_blocks{0};  // of them on the tiles' critical paths (per level the most)
        std::vector<std::uint32_t> level_items;  // work items of level l: [level_items[l], level_items[l+1])
        std::vector<std::uint32_t> level_longest;  // positions of the longest work item of level l
        std::vector<std::uint32_t> item_count_host;  // positions per work item (a copy of item_count)
        std::vector<std::uint32_t> level_pos;    // positions of level l: [level_pos[l], level_pos[l+1])
        DeviceArray<std::uint32_t> tile_off;     // per tile (+1): first compact index of its pending cells
        DeviceArray<std::uint32_t> pos_of;       // per compact index: chain position (recording only)
        DeviceArray<std::uint8_t> seg_start;     // per position
        DeviceArray<std::uint32_t> push_slot;    // per position: the side slot its outflow feeds (no_slot: none)
        DeviceArray<std::uint32_t> ref_pos;      // the positions that have side slots, level after level
        DeviceArray<std::uint32_t> ref_first;    // per entry of ref_pos (+1): its first side slot
        std::vector<std::uint32_t> level_ref;    // those of level l: ref_pos[level_ref[l] .. level_ref[l+1])
        DeviceArray<std::uint32_t> item_start;
        DeviceArray<std::uint32_t> item_count;
        // in-tile chain layout of round 0 (tile_layout.hpp)
        DeviceArray<std::uint16_t> tile_order;
        DeviceArray<std::uint8_t> tile_flags;
        DeviceArray<std::uint16_t> tile_split;
        DeviceArray<std::uint8_t> tile_nlevels;
        // the rest of the tiles' replay programs (tile_program_kernel)
        DeviceArray<std::uint32_t> pending_bits;  // per tile: cells left pending by round 0
        DeviceArray<std::uint32_t> lo_off;        // per tile (+1): its local cells that feed side slots
        DeviceArray<std::uint64_t> lo_entry;      // per such cell: side slot << 16 | local cell
        DeviceArray<std::uint32_t> stage_cell;    // per compact index: local cell | solved upstream mask << 16
                                                  // (recording only: replays read the sorted entries)
        DeviceArray<std::uint32_t> sorted_pos;    // per tile from tile_off: its pending cells' chain positions, sorted
        DeviceArray<std::uint32_t> sorted_cell;   //   and their staging entries in that order
        DeviceArray<std::uint32_t> scatter_entry; // per compact index: local cell | place in the sorted order << 16

        Network() = default;
        Network(Network const&) = delete;
        auto operator=(Network const&) -> Network& = delete;

        // A replay enqueued on any stream may still read the buffers: wait for the network's device before
        // freeing them (hipFreeAsync on the null stream, which is not ordered with non-blocking streams).
        ~Network()
        {
            int current = 0;
            bool const switched = hipGetDevice(&current) == hipSuccess && current != device &&
                                  hipSetDevice(device) == hipSuccess;
            (void)hipDeviceSynchronize();
            detach_stream();
            cats.reset();
            tile_cat.reset();
            round_tiles.reset();
            tile_off.reset();
            pos_of.reset();
            seg_start.reset();
            push_slot.reset();
            ref_pos.reset();
            ref_first.reset();
            item_start.reset();
            item_count.reset();
            tile_order.reset();
            tile_flags.reset();
            tile_split.reset();
            tile_nlevels.reset();
            pending_bits.reset();
            lo_off.reset();
            lo_entry.reset();
            stage_cell.reset();
            sorted_pos.reset();
            sorted_cell.reset();
            scatter_entry.reset();
            if (switched)
                (void)hipSetDevice(current);
        }

        auto rounds() const -> int { return static_cast<int>(round_offset.empty() ? 0 : round_offset.size() - 1); }
        auto levels() const -> int { return static_cast<int>(level_items.empty() ? 0 : level_items.size() - 1); }

        // First work item of [begin, end) (one level, longest first) with at most lane_max positions (end: none).
        auto lane_split(std::uint32_t begin, std::uint32_t end) const -> std::uint32_t
        {
            if (lane_max == 0u || end > item_count_host.size())
                return end;
            auto const first = item_count_host.begin() + begin;
            auto const last = item_count_host.begin() + end;
            auto const it = std::partition_point(first, last, [&](std::uint32_t n) { return n > lane_max; });
            return begin + static_cast<std::uint32_t>(it - first);
        }

        // Free the buffers on the null stream from now on (the recording stream may not outlive the network;
        // the owner synchronises the device before destroying it).
        void detach_stream()
        {
            cats.stream = nullptr;
            tile_cat.stream = nullptr;
            round_tiles.stream = nullptr;
            tile_off.stream = nullptr;
            pos_of.stream = nullptr;
            seg_start.stream = nullptr;
            push_slot.stream = nullptr;
            ref_pos.stream = nullptr;
            ref_first.stream = nullptr;
            item_start.stream = nullptr;
            item_count.stream = nullptr;
            tile_order.stream = nullptr;
            tile_flags.stream = nullptr;
            tile_split.stream = nullptr;
            tile_nlevels.stream = nullptr;
            pending_bits.stream = nullptr;
            lo_off.stream = nullptr;
            lo_entry.stream = nullptr;
            stage_cell.stream = nullptr;
            sorted_pos.stream = nullptr;
            sorted_cell.stream = nullptr;
            scatter_entry.stream = nullptr;
        }
    };

    // 128-bit fingerprint of n bytes: two sums over the 32-bit words, of splitmix64(word + index * golden ratio)
    // and of murmur3's fmix64(word ^ index * another odd constant) (order-sensitive, reduction-friendly; two
    // independent 64-bit sums, so that two different ldds share a fingerprint with odds of about 2^-128); word i
    // holds bytes 4i .. 4i + 3 (little endian), the tail zero-padded. Whole-word loads where `bytes` is 4-byte
    // aligned, byte loads otherwise. out[0], out[1]: the two sums (cleared before the launch).
    __global__ inline void fingerprint_kernel(std::uint8_t const* bytes, std::uint64_t n, unsigned long long* out)
    {
        std::uint64_t const words = (n + 3) / 4;
        bool const aligned = (reinterpret_cast<std::uintptr_t>(bytes) & 3u) == 0;
        std::uint64_t sum = 0;
        std::uint64_t sum2 = 0;
        for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < words;
             i += static_cast<std::uint64_t>(gridDim.x) * blockDim.x)
        {
            std::uint32_t w;
            if (aligned && 4 * i + 4 <= n)
            {
                w = as_global(reinterpret_cast<std::uint32_t const*>(bytes))[i];
            }
            else
            {
                w = 0;
                for (std::uint64_t j = 4 * i; j < n && j < 4 * i + 4; ++j)
                    w |= static_cast<std::uint32_t>(as_global(bytes)[j]) << (8 * (j - 4 * i));
            }
            std::uint64_t z = static_cast<std::uint64_t>(w) + (i + 1) * 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            sum += z ^ (z >> 31);
            std::uint64_t y = static_cast<std::uint64_t>(w) ^ ((i + 1) * 0xD6E8FEB86659FD93ull);
            y = (y ^ (y >> 33)) * 0xFF51AFD7ED558CCDull;
            y = (y ^ (y >> 33)) * 0xC4CEB9FE1A85EC53ull;
            sum2 += y ^ (y >> 33);
        }
        for (int offset = 32; offset > 0; offset >>= 1)
        {
            sum += __shfl_down(sum, offset, 64);
            sum2 += __shfl_down(sum2, offset, 64);
        }
        if ((threadIdx.x & 63) == 0)
        {
            atomicAdd(out, static_cast<unsigned long long>(sum));
            atomicAdd(out + 1, static_cast<unsigned long long>(sum2));
        }
    }

    // Keep memory freed by the operations' stream-ordered buffers in the device's pool (instead of returning
    // it to the system at every synchronisation), so repeated operations do not map their buffers anew.
    inline hipError_t keep_pool_memory(int device)
    {
        static std::mutex mutex;
        static std::vector<int> done;
        std::lock_guard<std::mutex> lock(mutex);
        if (std::find(done.begin(), done.end(), device) != done.end())
            return hipSuccess;
        hipMemPool_t pool = nullptr;
        hipError_t e = hipDeviceGetDefaultMemPool(&pool, device);
        if (e != hipSuccess)
            return e;
        std::uint64_t threshold = ~std::uint64_t{0};
        e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
        if (e == hipSuccess)
            done.push_back(device);
        return e;
    }

    // Describe a batch: tiles are numbered catchment after catchment.
    inline int network_init(Network& net, int n, std::uint8_t const* const* ldd, std::int64_t const* rows,
                            std::int64_t const* cols, int tile_shape, int schedule, hipStream_t stream)
    {
        if (n <= 0 || !ldd || !rows || !cols)
            return fail(status_invalid_argument, "a network needs at least one catchment");
        if (tile_shape < 0 || tile_shape >= tile_shape_count)
            return fail(status_invalid_argument, "unknown tile shape");
        if (schedule < 0 || schedule >= schedule_count)
            return fail(status_invalid_argument, "unknown schedule");
        int tw, th;
        tile_extent(tile_shape, tw, th);
        LUE_AMD_HIP(hipGetDevice(&net.device));
        LUE_AMD_HIP(keep_pool_memory(net.device));
        LUE_AMD_HIP(hipDeviceGetAttribute(&net.cus, hipDeviceAttributeMultiprocessorCount, net.device));
        net.tile_shape = tile_shape;
        net.schedule = schedule;
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
            if (rows[c] > (1ll << 31) || cols[c] > (1ll << 31) || rows[c] * cols[c] > (1ll << 32))
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

    // Per-catchment input/output tables of up to io_inline_max catchments travel to the device as a kernel
    // argument (no host staging buffer, so the launch sequence needs no host memory that outlives the call).
    constexpr int io_inline_max = 32;

    template<typename T>
    struct IoTable
    {
        CatchmentIO<T> entries[io_inline_max];
        int n;
    };

    template<typename T>
    __global__ void write_io_kernel(IoTable<T> table, CatchmentIO<T>* out)
    {
        int const i = static_cast<int>(threadIdx.x);
        if (i < table.n)
            out[i] = table.entries[i];
    }

    // Write the IO table of `n` catchments to `dst` in stream order.
    template<typename T>
    int write_io(CatchmentIO<T> const* io, int n, CatchmentIO<T>* dst, hipStream_t stream)
    {
        if (n <= io_inline_max)
        {
            IoTable<T> table{};
            for (int c = 0; c < n; ++c)
                table.entries[c] = io[c];
            table.n = n;
            hipLaunchKernelGGL((write_io_kernel<T>), dim3(1), dim3(64), 0, stream, table, dst);
            LUE_AMD_HIP(hipGetLastError());
            return status_success;
        }
        LUE_AMD_HIP(host_staging().copy_to_device(dst, io, sizeof(CatchmentIO<T>) * n, stream));
        return status_success;
    }

    // A network of one catchment passes the catchment and its rasters to the kernels as arguments.
    template<typename T>
    void set_one_catchment(Network const& net, CatchmentIO<T> const* io_host, SweepArgs<T>& a)
    {
        a.one_catchment = net.cats_host.size() == 1 ? 1 : 0;
        a.cat0 = a.one_catchment ? net.cats_host[0] : Catchment{};
        a.io0 = a.one_catchment ? io_host[0] : CatchmentIO<T>{};
    }

    // Scratch of recording an operation on a network (round 0 with walks, and the rounds of the rounds
    // schedule).
    struct RecordWorkspace
    {
        ScratchArray<std::uint32_t> pending_bits;
        ScratchArray<std::uint32_t> border_stamp;
        ScratchArray<float> out_lo;
        ScratchArray<std::uint32_t> tile_flag;
        ScratchArray<std::uint32_t> worklist;
        ScratchArray<std::uint32_t> work_count;
        ScratchArray<unsigned long long> counters;
        ScratchArray<unsigned char> io;
    };

    template<typename T, int TW, int TH>
    int record_workspace_init(Network const& net, RecordWorkspace& w, std::vector<CatchmentIO<T>> const& io,
                              SweepArgs<T>& a, hipStream_t stream)
    {
        using G = Tile<TW, TH>;
        LUE_AMD_HIP(w.pending_bits.allocate(static_cast<std::size_t>(net.ntiles * G::words), stream));
        LUE_AMD_HIP(w.border_stamp.allocate(static_cast<std::size_t>(net.ntiles * G::perimeter), stream));
        if constexpr (std::is_same_v<T, float>)
        {
            LUE_AMD_HIP(w.out_lo.allocate(static_cast<std::size_t>(net.ncells), stream));
        }
        LUE_AMD_HIP(w.tile_flag.allocate(static_cast<std::size_t>(3 * net.ntiles), stream));
        LUE_AMD_HIP(w.worklist.allocate(static_cast<std::size_t>(3 * net.ntiles), stream));
        LUE_AMD_HIP(w.work_count.allocate(4, stream));
        LUE_AMD_HIP(w.counters.allocate(2, stream));
        LUE_AMD_HIP(hipMemsetAsync(w.tile_flag.ptr, 0, 3 * net.ntiles * sizeof(std::uint32_t), stream));
        LUE_AMD_HIP(hipMemsetAsync(w.work_count.ptr, 0, 4 * sizeof(std::uint32_t), stream));
        LUE_AMD_HIP(hipMemsetAsync(w.counters.ptr, 0, 2 * sizeof(unsigned long long), stream));
        LUE_AMD_HIP(w.io.allocate(sizeof(CatchmentIO<T>) * io.size(), stream));
        if (int const st = write_io<T>(io.data(), static_cast<int>(io.size()),
                                       reinterpret_cast<CatchmentIO<T>*>(w.io.ptr), stream))
            return st;
        set_one_catchment(net, io.data(), a);
        a.cats = net.cats.ptr;
        a.tile_cat = net.tile_cat.ptr;
        a.ncats = static_cast<int>(net.cats_host.size());
        a.ntiles = net.ntiles;
        a.io = reinterpret_cast<CatchmentIO<T> const*>(w.io.ptr);
        a.out_lo = w.out_lo.ptr;
        a.pending_bits = w.pending_bits.ptr;
        a.border_stamp = w.border_stamp.ptr;
        a.tile_flag = w.tile_flag.ptr;
        a.worklist = w.worklist.ptr;
        a.work_count = w.work_count.ptr;
        a.counters = w.counters.ptr;
        a.record = 1;
        a.stage_off = nullptr;
        a.stage_pos = nullptr;
        a.staged = nullptr;
        a.stage_side = nullptr;
        a.tile_order = nullptr;
        a.tile_flags = nullptr;
        a.tile_split = nullptr;
        a.tile_nlevels = nullptr;
        a.stage_cell = nullptr;
        a.lo_off = nullptr;
        a.lo_entry = nullptr;
        a.slot = nullptr;
        return status_success;
    }

    template<typename P>
    hipError_t copy_to_host(P* dst, P const* src, std::size_t n, hipStream_t stream)
    {
        hipError_t e = hipMemcpyAsync(dst, src, sizeof(P) * n, hipMemcpyDeviceToHost, stream);
        return e != hipSuccess ? e : hipStreamSynchronize(stream);
    }

    // List ranking of the tour elements by a sparse ruling set (chains.hpp: rank_* kernels): value[e] = number
    // of enter elements from e to the end of its list (inclusive), last[e] = the list's final element. Lists
    // that do not end (cycles in the flow network) leave `cyclic` set. The rulers (about n2 / 64 plus the list
    // heads) are ranked by pointer jumping with a fixed number of steps and one readback at the end.
    inline int rank_lists(std::int32_t const* succ, std::uint32_t n2, DeviceArray<std::int32_t>& value,
                          DeviceArray<std::int32_t>& last, bool& cyclic, hipStream_t stream)
    {
        DeviceArray<std::uint32_t> has_pred, is_ruler, ruler_rank;
        DeviceArray<std::uint64_t> ownpre;
        DeviceArray<std::uint32_t> bad;
        LUE_AMD_HIP(has_pred.allocate(n2, stream));
        LUE_AMD_HIP(is_ruler.allocate(n2 + 1, stream));
        LUE_AMD_HIP(ruler_rank.allocate(n2 + 1, stream));
        LUE_AMD_HIP(ownpre.allocate(n2, stream));
        LUE_AMD_HIP(bad.allocate(1, stream));
        LUE_AMD_HIP(value.allocate(n2, stream));
        LUE_AMD_HIP(last.allocate(n2, stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(n2)), dim3(256), 0, stream, has_pred.ptr, 0u,
                           static_cast<std::uint64_t>(n2));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, bad.ptr, 0u, 1ull);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, is_ruler.ptr + n2, 0u, 1ull);
        hipLaunchKernelGGL(rank_mark_successors_kernel, dim3(blocks_for(n2)), dim3(256), 0, stream, succ, n2,
                           has_pred.ptr);
        hipLaunchKernelGGL(rank_rulers_packed_kernel, dim3(blocks_for(n2)), dim3(256), 0, stream, has_pred.ptr, n2,
                           is_ruler.ptr, ownpre.ptr);
        LUE_AMD_HIP(hipGetLastError());
        LUE_AMD_HIP(exclusive_scan(is_ruler.ptr, ruler_rank.ptr, static_cast<std::size_t>(n2) + 1, stream));
        std::uint32_t nrulers = 0;
        LUE_AMD_HIP(copy_to_host(&nrulers, ruler_rank.ptr + n2, 1, stream));
        if (nrulers == 0)
        {
            cyclic = true;  // every list circles
            return status_success;
        }
        DeviceArray<std::uint32_t> ruler_elem;
        DeviceArray<std::int32_t> rnext[2], rvalue[2], rlast[2];  // [0] starts as the walks' results
        LUE_AMD_HIP(ruler_elem.allocate(nrulers, stream));
        for (int b = 0; b < 2; ++b)
        {
            LUE_AMD_HIP(rnext[b].allocate(nrulers, stream));
            LUE_AMD_HIP(rvalue[b].allocate(nrulers, stream));
            LUE_AMD_HIP(rlast[b].allocate(nrulers, stream));
        }
        hipLaunchKernelGGL(rank_ruler_elements_kernel, dim3(blocks_for(n2)), dim3(256), 0, stream, is_ruler.ptr,
                           ruler_rank.ptr, n2, ruler_elem.ptr);
        hipLaunchKernelGGL(rank_walk_kernel, dim3(blocks_for(nrulers)), dim3(256), 0, stream, succ, ruler_rank.ptr,
                           ruler_elem.ptr, nrulers, ownpre.ptr, rvalue[0].ptr, rnext[0].ptr, rlast[0].ptr, bad.ptr);
        LUE_AMD_HIP(hipGetLastError());
        // the rulers' list ranked by a second ruling set: super-rulers (the rulers no ruler links to and a
        // pseudo-random 1/64 of the others) walk their sublists of rulers, the super-rulers are ranked by
        // pointer jumping (value = sum of the sublists from the super-ruler to the list's end, last = the
        // list's last element, known at the list's final super-ruler and carried back), and every ruler's
        // rank follows from its super-ruler's. rvalue[1] / rlast[1]: the rulers' results.
        DeviceArray<std::uint32_t> has_pred2, is_super, super_rank, super_elem;
        DeviceArray<std::int32_t> owner2, prefix2;
        LUE_AMD_HIP(has_pred2.allocate(nrulers, stream));
        LUE_AMD_HIP(is_super.allocate(nrulers + 1, stream));
        LUE_AMD_HIP(super_rank.allocate(nrulers + 1, stream));
        LUE_AMD_HIP(owner2.allocate(nrulers, stream));
        LUE_AMD_HIP(prefix2.allocate(nrulers, stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(nrulers)), dim3(256), 0, stream, has_pred2.ptr, 0u,
                           static_cast<std::uint64_t>(nrulers));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, is_super.ptr + nrulers, 0u, 1ull);
        hipLaunchKernelGGL(rank_mark_successors_kernel, dim3(blocks_for(nrulers)), dim3(256), 0, stream,
                           rnext[0].ptr, nrulers, has_pred2.ptr);
        hipLaunchKernelGGL(rank_rulers_kernel, dim3(blocks_for(nrulers)), dim3(256), 0, stream, has_pred2.ptr,
                           nrulers, is_super.ptr, owner2.ptr);
        LUE_AMD_HIP(hipGetLastError());
        LUE_AMD_HIP(exclusive_scan(is_super.ptr, super_rank.ptr, static_cast<std::size_t>(nrulers) + 1, stream));
        std::uint32_t nsupers = 0;
        LUE_AMD_HIP(copy_to_host(&nsupers, super_rank.ptr + nrulers, 1, stream));
        if (nsupers == 0)
        {
            cyclic = true;  // every list of rulers circles
            return status_success;
        }
        DeviceArray<std::int32_t> qnext[2], qvalue[2], qlast[2];
        LUE_AMD_HIP(super_elem.allocate(nsupers, stream));
        for (int b = 0; b < 2; ++b)
        {
            LUE_AMD_HIP(qnext[b].allocate(nsupers, stream));
            LUE_AMD_HIP(qvalue[b].allocate(nsupers, stream));
            LUE_AMD_HIP(qlast[b].allocate(nsupers, stream));
        }
        hipLaunchKernelGGL(rank_ruler_elements_kernel, dim3(blocks_for(nrulers)), dim3(256), 0, stream, is_super.ptr,
                           super_rank.ptr, nrulers, super_elem.ptr);
        hipLaunchKernelGGL(rank2_walk_kernel, dim3(blocks_for(nsupers)), dim3(256), 0, stream, rnext[0].ptr,
                           rvalue[0].ptr, rlast[0].ptr, is_super.ptr, super_rank.ptr, super_elem.ptr, nsupers,
                           owner2.ptr, prefix2.ptr, qvalue[0].ptr, qnext[0].ptr, qlast[0].ptr, bad.ptr);
        LUE_AMD_HIP(hipGetLastError());
        int cur = 0;
        int steps = 2;
        for (std::uint64_t span = 1; span < nsupers; span <<= 1)
            ++steps;
        DeviceArray<std::uint32_t> linked;
        LUE_AMD_HIP(linked.allocate(1, stream));
        for (int step = 0; step < steps; ++step)
        {
            hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, linked.ptr, 0u, 1ull);
            hipLaunchKernelGGL(chains_rank_step_kernel, dim3(blocks_for(nsupers)), dim3(256), 0, stream,
                               qnext[cur].ptr, qvalue[cur].ptr, qlast[cur].ptr, nsupers, qnext[1 - cur].ptr,
                               qvalue[1 - cur].ptr, qlast[1 - cur].ptr, linked.ptr);
            cur = 1 - cur;
        }
        LUE_AMD_HIP(hipGetLastError());
        hipLaunchKernelGGL(rank2_rulers_kernel, dim3(blocks_for(nrulers)), dim3(256), 0, stream, owner2.ptr,
                           prefix2.ptr, rvalue[0].ptr, qvalue[cur].ptr, qlast[cur].ptr, nrulers, rvalue[1].ptr,
                           rlast[1].ptr, bad.ptr);
        LUE_AMD_HIP(hipGetLastError());
        cur = 1;
        hipLaunchKernelGGL(rank_elements_kernel, dim3(blocks_for(n2)), dim3(256), 0, stream, ownpre.ptr,
                           rvalue[cur].ptr, rlast[cur].ptr, n2, value.ptr, last.ptr, bad.ptr);
        LUE_AMD_HIP(hipGetLastError());
        std::uint32_t still = 0;
        std::uint32_t nbad = 0;
        LUE_AMD_HIP(copy_to_host(&still, linked.ptr, 1, stream));
        LUE_AMD_HIP(copy_to_host(&nbad, bad.ptr, 1, stream));
        cyclic = still > 0 || nbad > 0;
        return status_success;
    }

    // Whether perm[0, n) is a permutation of [0, n); recording stops with an error instead of indexing out
    // of bounds if it is not.
    inline int check_permutation(std::uint32_t const* perm, std::uint32_t n, char const* what, hipStream_t stream)
    {
        DeviceArray<std::uint32_t> seen, bad;
        LUE_AMD_HIP(seen.allocate(n, stream));
        LUE_AMD_HIP(bad.allocate(1, stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, seen.ptr, 0u,
                           static_cast<std::uint64_t>(n));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, bad.ptr, 0u, 1ull);
        hipLaunchKernelGGL(chains_check_permutation_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, perm, n,
                           seen.ptr, bad.ptr);
        LUE_AMD_HIP(hipGetLastError());
        std::uint32_t nbad = 0;
        LUE_AMD_HIP(copy_to_host(&nbad, bad.ptr, 1, stream));
        if (nbad != 0)
            return fail(status_hip_error, std::string("chain layout: ") + what + " is not a permutation");
        return status_success;
    }

    // Record the chain schedule from the pending bitmaps of round 0.
    template<int TW, int TH>
    int build_chains(Network& net, std::uint32_t const* pending_bits, std::uint32_t* lo_slot, hipStream_t stream,
                     DeferredFrees* later = nullptr)
    {
        std::int64_t const ntiles = net.ntiles;
        int const ncats = static_cast<int>(net.cats_host.size());
        {
            DeviceArray<std::uint32_t> counts;
            LUE_AMD_HIP(counts.allocate(static_cast<std::size_t>(ntiles + 1), stream));
            LUE_AMD_HIP(net.tile_off.allocate(static_cast<std::size_t>(ntiles + 1), stream));
            hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, counts.ptr + ntiles, 0u, 1ull);
            hipLaunchKernelGGL((chains_count_kernel<TW, TH>), dim3(blocks_for(ntiles)), dim3(256), 0, stream,
                               pending_bits, ntiles, counts.ptr);
            LUE_AMD_HIP(hipGetLastError());
            LUE_AMD_HIP(exclusive_scan(counts.ptr, net.tile_off.ptr, static_cast<std::size_t>(ntiles + 1), stream));
        }
        std::uint32_t n = 0;
        LUE_AMD_HIP(copy_to_host(&n, net.tile_off.ptr + ntiles, 1, stream));
        net.pending = n;
        net.level_items.assign(1, 0);
        net.level_pos.assign(1, 0);
        if (n == 0)
            return status_success;
        if (n >= (1u << 30))
            return fail(status_invalid_argument, "too many cells left for the chain schedule (2^30)");

        DeviceArray<std::uint32_t> cells, index_of;
        DeviceArray<std::int32_t> parent, heavy, succ;
        LUE_AMD_HIP(cells.allocate(n, stream));
        LUE_AMD_HIP(index_of.allocate(static_cast<std::size_t>(net.ncells), stream));
        LUE_AMD_HIP(parent.allocate(n, stream));
        LUE_AMD_HIP(heavy.allocate(n, stream));
        LUE_AMD_HIP(succ.allocate(2 * static_cast<std::size_t>(n), stream));
        hipLaunchKernelGGL((chains_compact_kernel<TW, TH>), dim3(static_cast<unsigned>(ntiles)), dim3(64), 0, stream,
                           pending_bits, net.cats.ptr, net.tile_cat.ptr, net.tile_off.ptr, cells.ptr, index_of.ptr);
        hipLaunchKernelGGL((chains_parent_kernel<TW, TH>), dim3(blocks_for(n)), dim3(256), 0, stream, net.cats.ptr,
                           ncats, pending_bits, cells.ptr, index_of.ptr, n, parent.ptr);
        LUE_AMD_HIP(hipGetLastError());

        // subtree sizes (Euler tour, children in neighbour order), heavy children
        DeviceArray<std::int32_t> value, last;
        bool cyclic = false;
        hipLaunchKernelGGL((chains_tour_kernel<TW, TH>), dim3(blocks_for(n)), dim3(256), 0, stream, net.cats.ptr, ncats,
                           pending_bits, cells.ptr, index_of.ptr, parent.ptr, nullptr, n, succ.ptr);
        LUE_AMD_HIP(hipGetLastError());
        if (int const st = rank_lists(succ.ptr, 2 * n, value, last, cyclic, stream))
            return st;
        if (cyclic)
            return fail(status_invalid_network, "invalid flow direction network: cells on or downstream of a cycle");
        hipLaunchKernelGGL((chains_heavy_kernel<TW, TH>), dim3(blocks_for(n)), dim3(256), 0, stream, net.cats.ptr,
                           ncats, pending_bits, cells.ptr, index_of.ptr, value.ptr, n, heavy.ptr);
        LUE_AMD_HIP(hipGetLastError());

        // preorder with heavy children first: every chain is a contiguous run
        hipLaunchKernelGGL((chains_tour_kernel<TW, TH>), dim3(blocks_for(n)), dim3(256), 0, stream, net.cats.ptr, ncats,
                           pending_bits, cells.ptr, index_of.ptr, parent.ptr, heavy.ptr, n, succ.ptr);
        LUE_AMD_HIP(hipGetLastError());
        if (int const st = rank_lists(succ.ptr, 2 * n, value, last, cyclic, stream))
            return st;
        if (cyclic)
            return fail(status_invalid_network, "invalid flow direction network: cells on or downstream of a cycle");
        DeviceArray<std::uint32_t> root_flag, root_index, tree_size, tree_base, pre, head_flag, chain_incl, at_pre;
        LUE_AMD_HIP(root_flag.allocate(n + 1, stream));
        LUE_AMD_HIP(root_index.allocate(n + 1, stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, root_flag.ptr + n, 0u, 1ull);
        hipLaunchKernelGGL(chains_root_flag_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, parent.ptr, n,
                           root_flag.ptr);
        LUE_AMD_HIP(exclusive_scan(root_flag.ptr, root_index.ptr, n + 1, stream));
        std::uint32_t nroots = 0;
        LUE_AMD_HIP(copy_to_host(&nroots, root_index.ptr + n, 1, stream));
        LUE_AMD_HIP(tree_size.allocate(nroots + 1, stream));
        LUE_AMD_HIP(tree_base.allocate(nroots + 1, stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, tree_size.ptr + nroots, 0u, 1ull);
        hipLaunchKernelGGL(chains_tree_size_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, parent.ptr,
                           value.ptr, root_index.ptr, n, tree_size.ptr);
        LUE_AMD_HIP(exclusive_scan(tree_size.ptr, tree_base.ptr, nroots + 1, stream));
        LUE_AMD_HIP(pre.allocate(n, stream));
        LUE_AMD_HIP(head_flag.allocate(n, stream));
        LUE_AMD_HIP(chain_incl.allocate(n, stream));
        LUE_AMD_HIP(at_pre.allocate(n, stream));
        hipLaunchKernelGGL(chains_preorder_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, parent.ptr, heavy.ptr,
                           value.ptr, last.ptr, root_index.ptr, tree_base.ptr, n, pre.ptr, head_flag.ptr, at_pre.ptr);
        LUE_AMD_HIP(hipGetLastError());
        if (int const st = check_permutation(pre.ptr, n, "preorder", stream))
            return st;
        LUE_AMD_HIP(inclusive_scan(head_flag.ptr, chain_incl.ptr, n, stream));
        std::uint32_t nchains = 0;
        LUE_AMD_HIP(copy_to_host(&nchains, chain_incl.ptr + n - 1, 1, stream));
        net.nchains = nchains;

        // chains: preorder of their heads, lengths, levels
        DeviceArray<std::uint32_t> chain_pre, length, level, base, changed;
        LUE_AMD_HIP(chain_pre.allocate(nchains, stream));
        LUE_AMD_HIP(length.allocate(nchains, stream));
        LUE_AMD_HIP(level.allocate(nchains, stream));
        LUE_AMD_HIP(base.allocate(nchains, stream));
        LUE_AMD_HIP(changed.allocate(1, stream));
        hipLaunchKernelGGL(chains_head_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, head_flag.ptr, chain_incl.ptr,
                           n, chain_pre.ptr);
        hipLaunchKernelGGL(chains_length_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream, chain_pre.ptr, nchains,
                           n, length.ptr);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream, level.ptr, 0u,
                           static_cast<std::uint64_t>(nchains));
        DeviceArray<std::uint32_t> parent_chain;
        LUE_AMD_HIP(parent_chain.allocate(nchains, stream));
        hipLaunchKernelGGL(chains_parent_chain_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream, parent.ptr,
                           pre.ptr, at_pre.ptr, chain_pre.ptr, chain_incl.ptr, nchains, parent_chain.ptr);
        for (int iteration = 0; iteration < chains_max_levels; ++iteration)
        {
            std::uint32_t count = 0;
            hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, changed.ptr, 0u, 1ull);
            hipLaunchKernelGGL(chains_level_by_chain_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream,
                               parent_chain.ptr, nchains, level.ptr, changed.ptr);
            LUE_AMD_HIP(hipGetLastError());
            LUE_AMD_HIP(copy_to_host(&count, changed.ptr, 1, stream));
            if (count == 0)
                break;
            if (iteration == chains_max_levels - 1)
                return fail(status_invalid_network, "chain levels did not converge");
        }

        // layout (device): chains level after level, in preorder of their heads inside a level; work items are
        // one long chain, or consecutive short chains of one level starting in one 64-position window; inside
        // a level the longest items come first
        {
            DeviceArray<std::uint32_t> order_in, order, level_sorted, len_sorted, base_sorted, stats, item_flag,
                item_incl;
            LUE_AMD_HIP(order_in.allocate(nchains, stream));
            LUE_AMD_HIP(order.allocate(nchains, stream));
            LUE_AMD_HIP(level_sorted.allocate(nchains, stream));
            LUE_AMD_HIP(len_sorted.allocate(nchains, stream));
            LUE_AMD_HIP(base_sorted.allocate(nchains, stream));
            LUE_AMD_HIP(item_flag.allocate(nchains, stream));
            LUE_AMD_HIP(item_incl.allocate(nchains, stream));
            // stats: [0, 64) first position per level, [64, 128) first item per level (~0: no such level),
            // [128] longest chain; level_first: the first sorted chain of each level
            constexpr int nstats = 2 * chains_max_levels + 1;
            DeviceArray<std::uint32_t> level_first;
            LUE_AMD_HIP(stats.allocate(nstats, stream));
            LUE_AMD_HIP(level_first.allocate(chains_max_levels, stream));
            hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(256), 0, stream, stats.ptr, 0u,
                               static_cast<std::uint64_t>(nstats));
            hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, level_first.ptr, ~0u,
                               static_cast<std::uint64_t>(chains_max_levels));
            hipLaunchKernelGGL(iota_u32_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream, order_in.ptr, nchains);
            LUE_AMD_HIP(sort_pairs(level.ptr, level_sorted.ptr, order_in.ptr, order.ptr, nchains, 7, stream));
            hipLaunchKernelGGL(chains_sorted_lengths_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream, order.ptr,
                               length.ptr, level_sorted.ptr, nchains, len_sorted.ptr, level_first.ptr,
                               stats.ptr + 2 * chains_max_levels);
            LUE_AMD_HIP(exclusive_scan(len_sorted.ptr, base_sorted.ptr, nchains, stream));
            hipLaunchKernelGGL(chains_items_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream, order.ptr,
                               level_sorted.ptr, len_sorted.ptr, base_sorted.ptr, nchains, base.ptr, item_flag.ptr);
            LUE_AMD_HIP(inclusive_scan(item_flag.ptr, item_incl.ptr, nchains, stream));
            std::uint32_t nitems = 0;
            LUE_AMD_HIP(copy_to_host(&nitems, item_incl.ptr + nchains - 1, 1, stream));
            DeviceArray<std::uint32_t> item_begin, item_end, item_level, index_in, index;
            DeviceArray<std::uint64_t> key_in, key;
            LUE_AMD_HIP(item_begin.allocate(nitems, stream));
            LUE_AMD_HIP(item_end.allocate(nitems, stream));
            LUE_AMD_HIP(item_level.allocate(nitems, stream));
            LUE_AMD_HIP(index_in.allocate(nitems, stream));
            LUE_AMD_HIP(index.allocate(nitems, stream));
            LUE_AMD_HIP(key_in.allocate(nitems, stream));
            LUE_AMD_HIP(key.allocate(nitems, stream));
            hipLaunchKernelGGL(chains_item_fill_kernel, dim3(blocks_for(nchains)), dim3(256), 0, stream,
                               level_sorted.ptr, len_sorted.ptr, base_sorted.ptr, item_flag.ptr, item_incl.ptr, nchains,
                               item_begin.ptr, item_end.ptr, item_level.ptr);
            hipLaunchKernelGGL(chains_level_stats_kernel, dim3(1), dim3(chains_max_levels), 0, stream, level_first.ptr,
                               base_sorted.ptr, item_incl.ptr, stats.ptr);
            hipLaunchKernelGGL(chains_item_key_kernel, dim3(blocks_for(nitems)), dim3(256), 0, stream, item_begin.ptr,
                               item_end.ptr, item_level.ptr, nitems, key_in.ptr, index_in.ptr);
            LUE_AMD_HIP(sort_pairs(key_in.ptr, key.ptr, index_in.ptr, index.ptr, nitems, 32 + 7, stream));
            LUE_AMD_HIP(net.item_start.allocate(nitems, stream));
            LUE_AMD_HIP(net.item_count.allocate(nitems, stream));
            hipLaunchKernelGGL(chains_item_gather_kernel, dim3(blocks_for(nitems)), dim3(256), 0, stream, index.ptr,
                               item_begin.ptr, item_end.ptr, nitems, net.item_start.ptr, net.item_count.ptr);
            LUE_AMD_HIP(hipGetLastError());
            std::uint32_t h_stats[nstats];
            LUE_AMD_HIP(copy_to_host(h_stats, stats.ptr, nstats, stream));
            // levels are 0 .. nlevels - 1 without gaps (a chain's level is one more than one of its side chains')
            std::uint32_t nlevels = 0;
            while (nlevels < static_cast<std::uint32_t>(chains_max_levels) && h_stats[nlevels] != ~0u)
                ++nlevels;
            if (nlevels == 0 || h_stats[0] != 0 || h_stats[chains_max_levels] != 0)
                return fail(status_invalid_network, "chain layout: levels not accounted for");
            net.level_pos.assign(1, 0);
            net.level_items.assign(1, 0);
            for (std::uint32_t l = 1; l <= nlevels; ++l)
            {
                net.level_pos.push_back(l < nlevels ? h_stats[l] : n);
                net.level_items.push_back(l < nlevels ? h_stats[chains_max_levels + l] : nitems);
            }
            net.longest_chain = h_stats[2 * chains_max_levels];
            // inside a level the longest work item comes first
            net.item_count_host.resize(nitems);
            LUE_AMD_HIP(copy_to_host(net.item_count_host.data(), net.item_count.ptr, nitems, stream));
            net.level_longest.assign(nlevels, 0u);
            for (std::uint32_t l = 0; l < nlevels; ++l)
            {
                if (net.level_items[l + 1] > net.level_items[l])
                    net.level_longest[l] = net.item_count_host[net.level_items[l]];
            }
            LUE_AMD_HIP(hipStreamSynchronize(stream));  // before the temporaries go
        }

        // positions, chain starts, side references
        LUE_AMD_HIP(net.pos_of.allocate(n, stream));
        LUE_AMD_HIP(net.seg_start.allocate(n, stream));
        hipLaunchKernelGGL(chains_position_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, pre.ptr, chain_incl.ptr,
                           chain_pre.ptr, length.ptr, base.ptr, n, net.pos_of.ptr, net.seg_start.ptr);
        LUE_AMD_HIP(hipGetLastError());
        if (int const st = check_permutation(net.pos_of.ptr, n, "chain positions", stream))
            return st;
        DeviceArray<std::uint32_t> ref_count, ref_off;
        LUE_AMD_HIP(ref_count.allocate(n + 1, stream));
        LUE_AMD_HIP(ref_off.allocate(n + 1, stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, ref_count.ptr + n, 0u, 1ull);
        hipLaunchKernelGGL((chains_refs_kernel<TW, TH>), dim3(blocks_for(n)), dim3(256), 0, stream, net.cats.ptr, ncats,
                           pending_bits, cells.ptr, index_of.ptr, heavy.ptr, net.pos_of.ptr, n, false, ref_count.ptr,
                           nullptr, nullptr, nullptr);
        LUE_AMD_HIP(hipGetLastError());
        LUE_AMD_HIP(exclusive_scan(ref_count.ptr, ref_off.ptr, n + 1, stream));
        std::uint32_t nrefs = 0;
        LUE_AMD_HIP(copy_to_host(&nrefs, ref_off.ptr + n, 1, stream));
        net.nrefs = nrefs;
        // side slots: each source (a light chain's last position, a raster cell of round 0) feeds one
        LUE_AMD_HIP(net.push_slot.allocate(n, stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, net.push_slot.ptr, no_slot,
                           static_cast<std::uint64_t>(n));
        hipLaunchKernelGGL((chains_refs_kernel<TW, TH>), dim3(blocks_for(n)), dim3(256), 0, stream, net.cats.ptr, ncats,
                           pending_bits, cells.ptr, index_of.ptr, heavy.ptr, net.pos_of.ptr, n, true, nullptr,
                           ref_off.ptr, net.push_slot.ptr, lo_slot);
        LUE_AMD_HIP(hipGetLastError());

        // the positions with side references, in position order (so level after level): the gather of a level
        // visits only those
        {
            DeviceArray<std::uint32_t> has_ref;
            DeviceArray<std::uint32_t> ref_rank;
            LUE_AMD_HIP(has_ref.allocate(n + 1, stream));
            LUE_AMD_HIP(ref_rank.allocate(n + 1, stream));
            hipLaunchKernelGGL(chains_has_ref_kernel, dim3(blocks_for(n + 1)), dim3(256), 0, stream, ref_off.ptr, n,
                               has_ref.ptr);
            LUE_AMD_HIP(hipGetLastError());
            LUE_AMD_HIP(exclusive_scan(has_ref.ptr, ref_rank.ptr, n + 1, stream));
            std::uint32_t nref_pos = 0;
            LUE_AMD_HIP(copy_to_host(&nref_pos, ref_rank.ptr + n, 1, stream));
            LUE_AMD_HIP(net.ref_pos.allocate(std::max<std::uint32_t>(nref_pos, 1), stream));
            hipLaunchKernelGGL(chains_ref_pos_kernel, dim3(blocks_for(n)), dim3(256), 0, stream, has_ref.ptr,
                               ref_rank.ptr, n, net.ref_pos.ptr);
            LUE_AMD_HIP(hipGetLastError());
            LUE_AMD_HIP(net.ref_first.allocate(nref_pos + 1, stream));
            hipLaunchKernelGGL(chains_ref_first_kernel, dim3(blocks_for(nref_pos + 1)), dim3(256), 0, stream,
                               net.ref_pos.ptr, ref_off.ptr, nref_pos, n, net.ref_first.ptr);
            LUE_AMD_HIP(hipGetLastError());
            net.level_ref.assign(net.level_pos.size(), 0);
            for (std::size_t l = 0; l < net.level_pos.size(); ++l)
                LUE_AMD_HIP(copy_to_host(&net.level_ref[l], ref_rank.ptr + net.level_pos[l], 1, stream));
            LUE_AMD_HIP(hipStreamSynchronize(stream));  // before the temporaries go
        }
        LUE_AMD_HIP(hipStreamSynchronize(stream));  // host vectors go out of scope
        deferred_frees = later;  // the function's temporaries, freed on return, go to `later`
        return status_success;
    }

    // Record the in-tile chain layout of round 0 from the pending bitmaps.
    template<int TW, int TH>
    int build_tile_layout(Network& net, std::uint32_t const* pending_bits, std::uint32_t const* lo_slot,
                          hipStream_t stream, DeferredFrees* later = nullptr)
    {
        using G = Tile<TW, TH>;
        std::size_t const ntiles = static_cast<std::size_t>(net.ntiles);
        LUE_AMD_HIP(net.tile_order.allocate(ntiles * G::n, stream));
        LUE_AMD_HIP(net.tile_flags.allocate(ntiles * G::n, stream));
        LUE_AMD_HIP(net.tile_split.allocate(ntiles * tile_split_entries, stream));
        LUE_AMD_HIP(net.tile_nlevels.allocate(ntiles, stream));
        DeviceArray<std::uint32_t> error;
        LUE_AMD_HIP(error.allocate(1, stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, error.ptr, 0u, 1ull);
        hipLaunchKernelGGL((tile_layout_kernel<TW, TH>), dim3(static_cast<unsigned>(ntiles)), dim3(256), 0, stream,
                           net.cats.ptr, net.tile_cat.ptr, pending_bits, net.tile_order.ptr, net.tile_flags.ptr,
                           net.tile_split.ptr, net.tile_nlevels.ptr, error.ptr);
        LUE_AMD_HIP(hipGetLastError());
        if (later)
            later->flush();  // while the layout kernel runs
        // the pending bitmaps stay with the network (the tile pass reads them), beside the staging entries of
        // the pending cells and the list of the local cells that feed side slots
        LUE_AMD_HIP(net.pending_bits.allocate(ntiles * G::words, stream));
        LUE_AMD_HIP(net.stage_cell.allocate(std::max<std::size_t>(net.pending, 1), stream));
        LUE_AMD_HIP(net.sorted_pos.allocate(std::max<std::size_t>(net.pending, 1), stream));
        LUE_AMD_HIP(net.sorted_cell.allocate(std::max<std::size_t>(net.pending, 1), stream));
        LUE_AMD_HIP(net.scatter_entry.allocate(std::max<std::size_t>(net.pending, 1), stream));
        LUE_AMD_HIP(hipMemcpyAsync(net.pending_bits.ptr, pending_bits, ntiles * G::words * sizeof(std::uint32_t),
                                   hipMemcpyDeviceToDevice, stream));
        DeviceArray<std::uint32_t> lo_bits, lo_count;
        LUE_AMD_HIP(lo_bits.allocate(ntiles * G::words, stream));
        LUE_AMD_HIP(lo_count.allocate(ntiles + 1, stream));
        LUE_AMD_HIP(net.lo_off.allocate(ntiles + 1, stream));
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, lo_count.ptr + ntiles, 0u, 1ull);
        hipLaunchKernelGGL((tile_program_kernel<TW, TH>), dim3(static_cast<unsigned>(ntiles)), dim3(256), 0, stream,
                           net.cats.ptr, net.tile_cat.ptr, pending_bits, net.tile_off.ptr, lo_bits.ptr, lo_count.ptr,
                           net.stage_cell.ptr);
        LUE_AMD_HIP(hipGetLastError());
        if (net.pending > 0)
        {
            hipLaunchKernelGGL((tile_sort_kernel<TW, TH>), dim3(static_cast<unsigned>(ntiles)), dim3(256), 0, stream,
                               net.tile_off.ptr, net.pos_of.ptr, net.stage_cell.ptr, net.sorted_pos.ptr,
                               net.sorted_cell.ptr, net.scatter_entry.ptr);
            LUE_AMD_HIP(hipGetLastError());
        }
        net.pos_of.reset();  // replays use the sorted positions and entries
        net.stage_cell.reset();
        LUE_AMD_HIP(exclusive_scan(lo_count.ptr, net.lo_off.ptr, ntiles + 1, stream));
        std::uint32_t nlo = 0;
        LUE_AMD_HIP(copy_to_host(&nlo, net.lo_off.ptr + ntiles, 1, stream));
        if (nlo > net.nrefs)
            return fail(status_invalid_network, "tile programs: more side-slot sources than side slots");
        LUE_AMD_HIP(net.lo_entry.allocate(std::max<std::uint32_t>(nlo, 1), stream));
        if (nlo > 0)
        {
            DeviceArray<std::uint32_t> unassigned;
            LUE_AMD_HIP(unassigned.allocate(1, stream));
            hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, unassigned.ptr, 0u, 1ull);
            hipLaunchKernelGGL((tile_lo_list_kernel<TW, TH>), dim3(static_cast<unsigned>(ntiles)), dim3(256), 0, stream,
                               net.cats.ptr, net.tile_cat.ptr, lo_bits.ptr, net.lo_off.ptr, lo_slot,
                               static_cast<std::uint32_t>(net.nrefs), net.lo_entry.ptr, unassigned.ptr);
            LUE_AMD_HIP(hipGetLastError());
            std::uint32_t bad = 0;
            LUE_AMD_HIP(copy_to_host(&bad, unassigned.ptr, 1, stream));
            if (bad != 0)
                return fail(status_invalid_network, "tile programs: a raster cell feeds no side slot");
        }
        LUE_AMD_HIP(hipStreamSynchronize(stream));  // before the temporaries go
        std::uint32_t errors = 0;
        LUE_AMD_HIP(copy_to_host(&errors, error.ptr, 1, stream));
        if (errors != 0)
            return fail(status_hip_error, "in-tile chain layout: too many levels");
        {
            DeviceArray<unsigned long long> counts;
            LUE_AMD_HIP(counts.allocate(2, stream));
            LUE_AMD_HIP(hipMemsetAsync(counts.ptr, 0, 2 * sizeof(unsigned long long), stream));
            hipLaunchKernelGGL(tile_block_stats_kernel, dim3(blocks_for(ntiles)), dim3(256), 0, stream,
                               net.tile_split.ptr, net.tile_nlevels.ptr, static_cast<std::int64_t>(ntiles), counts.ptr);
            LUE_AMD_HIP(hipGetLastError());
            unsigned long long h[2] = {0, 0};
            LUE_AMD_HIP(copy_to_host(h, counts.ptr, 2, stream));
            net.tile_blocks = h[0];
            net.tile_critical_blocks = h[1];
        }
        return status_success;
    }

    // Levels and the final scatter of the chain schedule.
    template<int OP, typename T, int TW, int TH>
    int launch_chains(Network const& net, SweepArgs<T> const& a, ChainRun<OP, T> const& c, hipStream_t stream)
    {
        if (net.pending == 0)
            return status_success;
        for (int l = 0; l < net.levels(); ++l)
        {
            std::uint32_t const ref_begin = net.level_ref[l];
            std::uint32_t const ref_end = net.level_ref[l + 1];
            if (ref_end > ref_begin)
            {
                std::uint32_t const per_block = 256u * gather_entries_per_thread;
                hipLaunchKernelGGL((chains_gather_kernel<OP, T>), dim3((ref_end - ref_begin + per_block - 1) / per_block),
                                   dim3(256), 0, stream, c, net.ref_pos.ptr, net.ref_first.ptr, ref_begin, ref_end);
            }
            std::uint32_t const begin = net.level_items[l];
            std::uint32_t const end = net.level_items[l + 1];
            if (end == begin)
                continue;
            bool const wide = has_wide_path<OP>() && net.wide_min > 0 &&
                              static_cast<std::size_t>(l) < net.level_longest.size() &&
                              net.level_longest[static_cast<std::size_t>(l)] >= net.wide_min;
            // whether an item is walked by a lane depends on its own length only (up to lane_max positions),
            // not on the other items of its level, which a batch of catchments changes
            std::uint32_t const split = net.lane_split(begin, end);
            if (wide)
            {
                if constexpr (has_wide_path<OP>())
                {
                    if (split > begin)
                        hipLaunchKernelGGL((chains_run_wide_kernel<OP, T>), dim3(split - begin), dim3(wide_threads), 0,
                                           stream, a, c, begin);
                    if (end > split)
                        hipLaunchKernelGGL((chains_run_lanes_kernel<OP, T>), dim3((end - split + 63) / 64), dim3(64), 0,
                                           stream, a, c, split, end);
                }
            }
            else
            {
                if (split >= end)
                    hipLaunchKernelGGL((chains_run_kernel<OP, T>), dim3((end - begin + 3) / 4), dim3(256), 0, stream, a,
                                       c, begin, end);
                else if (split == begin)
                    hipLaunchKernelGGL((chains_run_lanes_kernel<OP, T>), dim3((end - begin + 63) / 64), dim3(64), 0,
                                       stream, a, c, begin, end);
                else
                    hipLaunchKernelGGL((chains_run_mixed_kernel<OP, T>), dim3((split - begin) + (end - split + 63) / 64),
                                       dim3(64), 0, stream, a, c, begin, split, end);
            }
        }
        constexpr int tiles_per_block = scatter_tiles_per_block<OP, T, TW, TH>();
        hipLaunchKernelGGL((chains_scatter_kernel<OP, T, TW, TH>),
                           dim3(static_cast<unsigned>((net.ntiles + tiles_per_block - 1) / tiles_per_block)),
                           dim3(64 * tiles_per_block), 0, stream, a, net.tile_off.ptr, net.scatter_entry.ptr,
                           net.sorted_pos.ptr, c.res, c.res_rem, net.ntiles);
        LUE_AMD_HIP(hipGetLastError());
        return status_success;
    }

    // Timing of one operation (lue_amd_profile_*): events around round 0 and the rest, queued in the
    // profile without waiting.
    struct OpTimer
    {
        hipEvent_t ev[3]{nullptr, nullptr, nullptr};
        bool on{false};
        int kind{-1};  // operation * 2 + 1 for FP64

        int start(hipStream_t stream)
        {
            on = profile().enabled();
            if (!on)
                return status_success;
            LUE_AMD_HIP(profile().take(ev));
            LUE_AMD_HIP(hipEventRecord(ev[0], stream));
            return status_success;
        }

        int mark(hipStream_t stream)
        {
            if (on)
                LUE_AMD_HIP(hipEventRecord(ev[1], stream));
            return status_success;
        }

        int stop(hipStream_t stream, long long rounds, std::int64_t cells)
        {
            if (!on)
                return status_success;
            LUE_AMD_HIP(hipEventRecord(ev[2], stream));
            profile().finish(ev, kind, rounds, cells);
            return status_success;
        }

        ~OpTimer()
        {
            if (on)
                profile().give_back(ev);
        }
    };

    // Per-call device buffers of a replay, carved out of one block at these byte offsets (each 256-byte
    // aligned). The sizes depend on the network and the operation only.
    struct ReplayLayout
    {
        std::size_t pending_bits{};  // rounds schedule, per tile: pending bitmap (written by each round)
        std::size_t border_stamp{};  // per tile perimeter cell (rounds schedule only)
        std::size_t out_lo{};        // rounds schedule, FP32 path: compensation term per cell
        std::size_t io{};            // per catchment: CatchmentIO<T>
        std::size_t staged{};        // chain schedule, per pending cell: CellInput<OP, T>
        std::size_t side{};          //   FP64 side sum
        std::size_t res{};           //   result in the output type
        std::size_t res_rem{};       //   remainder (two-output operations)
        std::size_t slot{};          //   per side slot: the source's FP64 outflow
        std::size_t next_tile{};     //   the tile work queue's counter (tile_queue)
        std::size_t total{};
    };

    template<int OP, typename T, int TW, int TH>
    ReplayLayout replay_layout(Network const& net)
    {
        using G = Tile<TW, TH>;
        ReplayLayout l;
        std::size_t at = 0;
        auto take = [&at](std::size_t bytes)
        {
            std::size_t const offset = at;
            at += (std::max<std::size_t>(bytes, 1) + 255) / 256 * 256;
            return offset;
        };
        std::size_t const ntiles = static_cast<std::size_t>(net.ntiles);
        bool const chains = net.schedule == schedule_chains;
        std::size_t const pending = chains ? net.pending : 0;
        l.pending_bits = take(chains ? 0 : ntiles * G::words * sizeof(std::uint32_t));  // chains: the network's
        l.border_stamp = take(chains ? 0 : ntiles * G::perimeter * sizeof(std::uint32_t));
        l.out_lo = take(!chains && std::is_same_v<T, float> ? static_cast<std::size_t>(net.ncells) * sizeof(float) : 0);
        l.io = take(net.cats_host.size() * sizeof(CatchmentIO<T>));
        l.staged = take(OP == op_accu ? 0 : pending * sizeof(CellInput<OP, T>));  // accu: in the side sums
        l.side = take(pending * sizeof(double));
        l.res = take(pending * sizeof(T));
        l.res_rem = take(is_split_op(OP) ? pending * sizeof(T) : 0);
        l.slot = take(chains ? static_cast<std::size_t>(net.nrefs) * sizeof(double) : 0);
        l.next_tile = take(chains && net.tile_queue ? sizeof(std::uint32_t) : 0);
        l.total = at;
        return l;
    }

    // The launches of a replay on the per-call buffers at `base` (laid out by replay_layout; the IO table
    // already written). No device allocation and no host synchronisation; with `timed` false not even the
    // profiler's events are taken.
    template<int OP, typename T, int TW, int TH>
    int replay_launch(Network const& net, SweepArgs<T> a, CatchmentIO<T> const* io_host, unsigned char* base,
                      ReplayLayout const& l, hipStream_t stream, bool timed)
    {
        bool const chains = net.schedule == schedule_chains;
        set_one_catchment(net, io_host, a);
        a.cats = net.cats.ptr;
        a.tile_cat = net.tile_cat.ptr;
        a.ncats = static_cast<int>(net.cats_host.size());
        a.ntiles = net.ntiles;
        a.io = reinterpret_cast<CatchmentIO<T> const*>(base + l.io);
        a.out_lo = !chains && std::is_same_v<T, float> ? reinterpret_cast<float*>(base + l.out_lo) : nullptr;
        a.pending_bits = reinterpret_cast<std::uint32_t*>(base + l.pending_bits);
        a.border_stamp = chains ? nullptr : reinterpret_cast<std::uint32_t*>(base + l.border_stamp);
        a.tile_flag = nullptr;
        a.worklist = nullptr;
        a.work_count = nullptr;
        a.counters = nullptr;
        a.record = 0;
        a.stage_off = nullptr;
        a.stage_pos = nullptr;
        a.staged = nullptr;
        a.stage_side = nullptr;
        a.compact_stage = 0;
        if constexpr (OP == op_kinematic_wave && std::is_same_v<T, float>)
        {
            a.compact_stage = 1;
            for (std::size_t k = 0; k < net.cats_host.size(); ++k)
                a.compact_stage = a.compact_stage && io_host[k].beta == nullptr && io_host[k].dx == nullptr;
        }
        a.tile_order = nullptr;
        a.tile_flags = nullptr;
        a.tile_split = nullptr;
        a.tile_nlevels = nullptr;
        a.stage_cell = nullptr;
        a.lo_off = nullptr;
        a.lo_entry = nullptr;
        a.slot = nullptr;
        ChainRun<OP, T> c{};
        if (chains)
        {
            a.tile_order = net.tile_order.ptr;
            a.tile_flags = net.tile_flags.ptr;
            a.tile_split = net.tile_split.ptr;
            a.tile_nlevels = net.tile_nlevels.ptr;
            a.pending_bits = net.pending_bits.ptr;
            a.stage_cell = net.sorted_cell.ptr;  // the tile pass stages in chain-position order
            a.lo_off = net.lo_off.ptr;
            a.lo_entry = net.lo_entry.ptr;
            a.slot = reinterpret_cast<double*>(base + l.slot);
            a.stage_off = net.tile_off.ptr;
            a.stage_pos = net.sorted_pos.ptr;
            if (net.pending > 0)
            {
                a.staged = base + l.staged;
                a.stage_side = reinterpret_cast<double*>(base + l.side);
                c = ChainRun<OP, T>{net.item_start.ptr,
                                    net.item_count.ptr,
                                    net.seg_start.ptr,
                                    net.push_slot.ptr,
                                    reinterpret_cast<double*>(base + l.slot),
                                    reinterpret_cast<CellInput<OP, T> const*>(base + l.staged),
                                    reinterpret_cast<double*>(base + l.side),
                                    reinterpret_cast<T*>(base + l.res),
                                    is_split_op(OP) ? reinterpret_cast<T*>(base + l.res_rem) : nullptr};
            }
        }
        OpTimer timer;
        timer.kind = OP * 2 + (std::is_same_v<T, double> ? 1 : 0);
        if (timed)
        {
            if (int const st = timer.start(stream))
                return st;
        }
        if (chains && net.tile_queue)
        {
            // as many workgroups as the device holds at once, the counter cleared by a kernel (graph capture)
            std::uint32_t* const next_tile = reinterpret_cast<std::uint32_t*>(base + l.next_tile);
            hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, next_tile, 0u, 1ull);
            std::int64_t const resident =
                static_cast<std::int64_t>(std::max(net.cus, 1)) * tile_pass_workgroups_per_cu<OP, T, TW, TH>();
            hipLaunchKernelGGL((tile_pass_queue_kernel<OP, T, TW, TH>),
                               dim3(static_cast<unsigned>(std::min<std::int64_t>(net.ntiles, resident))),
                               dim3(sweep_threads), 0, stream, a, next_tile);
        }
        else if (chains)
            hipLaunchKernelGGL((tile_pass_kernel<OP, T, TW, TH>), dim3(static_cast<unsigned>(net.ntiles)),
                               dim3(sweep_threads), 0, stream, a);
        else
            hipLaunchKernelGGL((sweep_first_round<OP, T, TW, TH>), dim3(static_cast<unsigned>(net.ntiles)),
                               dim3(sweep_threads), 0, stream, a);
        LUE_AMD_HIP(hipGetLastError());
        if (int const st = timer.mark(stream))
            return st;
        if (chains)
        {
            if (int const st = launch_chains<OP, T, TW, TH>(net, a, c, stream))
                return st;
        }
        else
        {
            for (int r = 1; r <= net.rounds(); ++r)
            {
                std::uint64_t const begin = net.round_offset[r - 1];
                std::uint32_t const count = static_cast<std::uint32_t>(net.round_offset[r] - begin);
                unsigned const grid = std::min<unsigned>(count, 2048);
                hipLaunchKernelGGL((sweep_list_round<OP, T, TW, TH>), dim3(grid), dim3(sweep_threads), 0, stream, a,
                                   net.round_tiles.ptr + begin, count, static_cast<std::uint32_t>(r));
            }
            LUE_AMD_HIP(hipGetLastError());
        }
        return timer.stop(stream, chains ? net.levels() + 1 : net.rounds() + 1, net.ncells);
    }

    // A captured operation on a network: its own per-call buffers (replay_layout) with the IO table written
    // once, and the HIP graph of its launch sequence. A launch replays the graph on the caller's stream; the
    // arguments (raster pointers and scalars) are those given at capture. The network and the rasters must
    // outlive the plan.
    struct Plan
    {
        Network const* net{};
        unsigned char* buffer{};
        hipGraph_t graph{};
        hipGraphExec_t exec{};
        std::size_t nodes{};
        std::size_t bytes{};

        Plan() = default;
        Plan(Plan const&) = delete;
        auto operator=(Plan const&) -> Plan& = delete;

        // a launch may still use the buffers: wait for the device first
        ~Plan()
        {
            int current = 0;
            bool const switched = net && hipGetDevice(&current) == hipSuccess && current != net->device &&
                                  hipSetDevice(net->device) == hipSuccess;
            (void)hipDeviceSynchronize();
            if (exec)
                (void)hipGraphExecDestroy(exec);
            if (graph)
                (void)hipGraphDestroy(graph);
            if (buffer)
                (void)hipFree(buffer);
            if (switched)
                (void)hipSetDevice(current);
        }
    };

    template<int OP, typename T, int TW, int TH>
    int plan_capture(Network const& net, SweepArgs<T> a, std::vector<CatchmentIO<T>> const& io, Plan& plan)
    {
        if (!net.recorded)
            return fail(status_invalid_argument, "network has no recorded schedule");
        if (io.size() != net.cats_host.size())
            return fail(status_invalid_argument, "one input/output set per catchment");
        ReplayLayout const l = replay_layout<OP, T, TW, TH>(net);
        plan.net = &net;
        plan.bytes = l.total;
        LUE_AMD_HIP(hipMalloc(reinterpret_cast<void**>(&plan.buffer), l.total));
        LUE_AMD_HIP(hipMemcpy(plan.buffer + l.io, io.data(), sizeof(CatchmentIO<T>) * io.size(), hipMemcpyHostToDevice));
        hipStream_t s = nullptr;
        LUE_AMD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        int st = status_success;
        if (e == hipSuccess)
        {
            st = replay_launch<OP, T, TW, TH>(net, a, io.data(), plan.buffer, l, s, false);
            e = hipStreamEndCapture(s, &plan.graph);  // ends the capture whatever the launches returned
        }
        (void)hipStreamDestroy(s);
        if (st != status_success)
            return st;
        LUE_AMD_HIP(e);
        LUE_AMD_HIP(hipGraphGetNodes(plan.graph, nullptr, &plan.nodes));
        LUE_AMD_HIP(hipGraphInstantiate(&plan.exec, plan.graph, nullptr, nullptr, 0));
        return status_success;
    }

    // Replay: the operation as a fixed launch sequence from the recorded schedule, its per-call buffers one
    // block of the scratch cache (no host synchronisation).
    template<int OP, typename T, int TW, int TH>
    int replay_run(Network const& net, SweepArgs<T> a, std::vector<CatchmentIO<T>> const& io, hipStream_t stream)
    {
        if (!net.recorded)
            return fail(status_invalid_argument, "network has no recorded schedule");
        if (io.size() != net.cats_host.size())
            return fail(status_invalid_argument, "one input/output set per catchment");
        ReplayLayout const l = replay_layout<OP, T, TW, TH>(net);
        ScratchArray<unsigned char> block;
        LUE_AMD_HIP(block.allocate(l.total, stream));
        if (int const st = write_io<T>(io.data(), static_cast<int>(io.size()),
                                       reinterpret_cast<CatchmentIO<T>*>(block.ptr + l.io), stream))
            return st;
        return replay_launch<OP, T, TW, TH>(net, a, io.data(), block.ptr, l, stream, true);
    }

    // Record: run round 0 (and, for the rounds schedule, the later rounds) while recording the schedule into
    // `net`; the chain schedule is then built and the operation replayed.
    template<int OP, typename T, int TW, int TH>
    int record_run(Network& net, SweepArgs<T> a, std::vector<CatchmentIO<T>> const& io, hipStream_t stream,
                   bool produce_output = true)
    {
        RecordWorkspace w;
        if (int const st = record_workspace_init<T, TW, TH>(net, w, io, a, stream))
            return st;
        OpTimer timer;
        timer.kind = OP * 2 + (std::is_same_v<T, double> ? 1 : 0);
        bool const chains = net.schedule == schedule_chains;
        if (!chains)
        {
            if (int const st = timer.start(stream))
                return st;
        }
        hipLaunchKernelGGL((sweep_first_round<OP, T, TW, TH>),
                           dim3(static_cast<unsigned>(net.ntiles)), dim3(sweep_threads), 0,
                           stream, a);
        LUE_AMD_HIP(hipGetLastError());
        if (int const st = timer.mark(stream))
            return st;

        net.round_offset.assign(1, 0);
        std::uint32_t round = 1;
        if (!chains)
        {
            std::vector<std::uint32_t> tiles_of_rounds;
            unsigned const grid = static_cast<unsigned>(std::min<std::int64_t>(net.ntiles, 2048));
            while (true)
            {
                // the tiles queued for this round
                std::uint32_t const slot = round % 3;
                std::uint32_t count = 0;
                LUE_AMD_HIP(copy_to_host(&count, w.work_count.ptr + slot, 1, stream));
                if (count == 0)
                    break;
                std::size_t const at = tiles_of_rounds.size();
                tiles_of_rounds.resize(at + count);
                LUE_AMD_HIP(hipMemcpyAsync(tiles_of_rounds.data() + at,
                                           w.worklist.ptr + static_cast<std::int64_t>(slot) * net.ntiles,
                                           sizeof(std::uint32_t) * count, hipMemcpyDeviceToHost, stream));
                hipLaunchKernelGGL((sweep_next_round<OP, T, TW, TH>), dim3(std::min<unsigned>(grid, count)), dim3(sweep_threads),
                                   0, stream, a, round);
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
                                           sizeof(std::uint32_t) * tiles_of_rounds.size(), hipMemcpyHostToDevice,
                                           stream));
            }
            if (int const st = timer.stop(stream, round, net.ncells))
                return st;
        }
        unsigned long long counters[2] = {0, 0};
        LUE_AMD_HIP(copy_to_host(counters, w.counters.ptr, 2, stream));
        net.nodata_cells = static_cast<std::int64_t>(counters[1]);
        if (chains)
        {
            // side slot of every raster cell that feeds one (by cell index in the batch), from the chain layout
            // to the tile programs
            DeviceArray<std::uint32_t> lo_slot;
            LUE_AMD_HIP(lo_slot.allocate(static_cast<std::size_t>(net.ncells), stream));
            hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(static_cast<std::uint64_t>(net.ncells))), dim3(256), 0,
                               stream, lo_slot.ptr, no_slot, static_cast<std::uint64_t>(net.ncells));
            DeferredFrees later;
            int const chained = build_chains<TW, TH>(net, w.pending_bits.ptr, lo_slot.ptr, stream, &later);
            deferred_frees = nullptr;
            if (chained != status_success)
                return chained;
            if (int const st = build_tile_layout<TW, TH>(net, w.pending_bits.ptr, lo_slot.ptr, stream, &later))
                return st;
            if (counters[0] + counters[1] + net.pending != static_cast<std::uint64_t>(net.ncells))
                return fail(status_invalid_network, "invalid flow direction network: cells not accounted for");
            net.recorded = true;
            // round 0 of the recording did not stage the pending cells: run the operation on the schedule
            return produce_output ? replay_run<OP, T, TW, TH>(net, a, io, stream) : status_success;
        }
        std::uint64_t const total = counters[0] + counters[1];
        if (total != static_cast<std::uint64_t>(net.ncells))
        {
            return fail(status_invalid_network,
                        "invalid flow direction network: " + std::to_string(net.ncells - static_cast<std::int64_t>(total)) +
                            " cells lie on or downstream of a cycle");
        }
        net.recorded = true;
        return status_success;
    }

}  // namespace lue_amd
