// C++ front end of the MI355X-native flow-routing path.
//
// Mirrors the part of LUE's C++ API that the routing path uses, over device-resident storage:
//   - lue::PartitionedArray<E, 2>  (reference source/framework/partitioned_array/include/lue/framework/
//     partitioned_array_decl.hpp:23-98): move-only, shape, partition shape, partitions whose extents follow
//     the last-partition-absorbs-the-remainder rule (source/framework/core/include/lue/framework/core/
//     domain_decomposition.hpp:91-183). Here the partitions are contiguous sub-blocks of one row-major
//     raster in HBM;
//   - lue::Scalar<E>  (source/framework/algorithm/include/lue/framework/algorithm/scalar.hpp:20);
//   - lue::create_partitioned_array(shape, partition_shape, fill_value);
//   - lue::value_policies::{accu, accu_threshold, accu_fraction, accu_capacity, accu_trigger, kinematic_wave,
//     d8_flow_direction, upstream, downstream, accu_info} with the reference's overload sets
//     (ALG/value_policies/*.hpp), the cell classes of accu_info (ALG/accu_info.hpp:22-63) and
//     lue::inflow_count<std::uint8_t>(ldd) and lue::downstream_distance(ldd, cell_size)
//     (ALG/inflow_count.hpp:29-42), ALG = source/framework/algorithm/include/lue/framework/algorithm.
// Each operation calls the C ABI (include/lue_amd/routing.h). Shape and partitioning mismatches throw
// std::runtime_error as detail::verify_compatible does (ALG/detail/verify_compatible.hpp:22-41); other
// failures throw std::runtime_error with the library's message.
//
// Operations are asynchronous, like LUE's, whose partitions are hpx::future-backed components
// (partitioned_array_decl.hpp:23-98): an operation makes the front end's stream (lue::set_stream, default the
// null stream) wait for the completion events of its input arrays, enqueues its kernels, records a new
// completion event into its output arrays and returns. to_host() waits for the array's event; ready() tells
// whether it has completed. The routing operations run on the flow network (recorded schedule) of their ldd
// array, which the array keeps after its first routing operation and drops when its contents change
// (from_host, set_partitions, mutable_data); recording it is the only step that waits for the device.
#pragma once

#include "lue_amd/routing.h"

#include <array>
#include <cmath>
#include <concepts>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace lue {

    using Index = std::ptrdiff_t;
    using Count = std::ptrdiff_t;
    using Rank = std::size_t;

    template<typename Element>
    using Shape2 = std::array<Count, 2>;

    using FlowDirectionElement = std::uint8_t;

    namespace detail {

        inline void check(int status)
        {
            if (status == LUE_AMD_SUCCESS)
            {
                return;
            }
            std::string message = std::string(lue_amd_status_string(status)) + ": " + lue_amd_last_error();
            throw std::runtime_error(message);
        }

        inline auto shape_string(std::array<Count, 2> const& shape) -> std::string
        {
            return "(" + std::to_string(shape[0]) + ", " + std::to_string(shape[1]) + ")";
        }

        // The stream the front end enqueues its operations on (per thread; null: the default stream).
        inline auto front_end_stream() -> void*&
        {
            thread_local void* stream = nullptr;
            return stream;
        }

        // Completion event of an array's contents (a hipEvent_t behind the C ABI).
        class Event
        {
            public:

                Event() { check(lue_amd_event_create(&_event)); }

                Event(Event const&) = delete;
                auto operator=(Event const&) -> Event& = delete;

                ~Event() { (void)lue_amd_event_destroy(_event); }

                auto handle() const -> void* { return _event; }

            private:

                void* _event{nullptr};
        };

        // A recorded flow network (lue_amd_network).
        struct NetworkHandle
        {
            lue_amd_network network{nullptr};

            NetworkHandle() = default;
            NetworkHandle(NetworkHandle const&) = delete;
            auto operator=(NetworkHandle const&) -> NetworkHandle& = delete;

            ~NetworkHandle()
            {
                if (network)
                {
                    (void)lue_amd_network_destroy(network);
                }
            }
        };

    }  // namespace detail

    // Stream of the front end's operations on this thread (a hipStream_t; nullptr: the default stream).
    inline void set_stream(void* stream)
    {
        detail::front_end_stream() = stream;
    }

    inline auto stream() -> void*
    {
        return detail::front_end_stream();
    }

    // A value broadcast to every cell (LUE's Scalar<E>).
    template<typename Element>
    class Scalar
    {
        public:

            Scalar() = default;

            Scalar(Element value):
                _value{value}
            {
            }

            auto value() const -> Element { return _value; }

        private:

            Element _value{};
    };

    template<typename Element>
    Scalar(Element) -> Scalar<Element>;

    template<typename Element, Rank rank>
    class PartitionedArray;

    // Device-resident partitioned raster (rank 2 only on this path).
    template<typename Element>
    class PartitionedArray<Element, 2>
    {
        public:

            using Shape = std::array<Count, 2>;

            PartitionedArray() = default;

            PartitionedArray(Shape const& shape, Shape const& partition_shape):
                _shape{shape},
                _partition_shape{partition_shape}
            {
                if (shape[0] <= 0 || shape[1] <= 0)
                {
                    throw std::runtime_error("Array shape must not be empty");
                }
                if (partition_shape[0] <= 0 || partition_shape[1] <= 0)
                {
                    throw std::runtime_error("Partition shape must not be empty");
                }
                if (partition_shape[0] > shape[0] || partition_shape[1] > shape[1])
                {
                    throw std::runtime_error(
                        "Partition size (" + std::to_string(partition_shape[0] * partition_shape[1]) +
                        ") must be smaller than or equal to array size (" + std::to_string(shape[0] * shape[1]) +
                        ")");
                }
                void* ptr = nullptr;
                detail::check(lue_amd_device_malloc(&ptr, sizeof(Element) * nr_elements()));
                _data = std::shared_ptr<Element>(static_cast<Element*>(ptr), [](Element* p) { lue_amd_device_free(p); });
            }

            PartitionedArray(PartitionedArray const&) = delete;
            PartitionedArray(PartitionedArray&&) noexcept = default;
            auto operator=(PartitionedArray const&) -> PartitionedArray& = delete;
            auto operator=(PartitionedArray&&) noexcept -> PartitionedArray& = default;
            ~PartitionedArray() = default;

            auto shape() const -> Shape const& { return _shape; }

            auto partition_shape() const -> Shape const& { return _partition_shape; }

            auto nr_elements() const -> Count { return _shape[0] * _shape[1]; }

            auto shape_in_partitions() const -> Shape
            {
                return {_shape[0] / _partition_shape[0], _shape[1] / _partition_shape[1]};
            }

            auto nr_partitions() const -> Count
            {
                auto const s = shape_in_partitions();
                return s[0] * s[1];
            }

            // Extents of partition (p0, p1): the last partition along a dimension absorbs the remainder.
            auto partition_extent(Index p0, Index p1) const -> Shape
            {
                auto const s = shape_in_partitions();
                Count const e0 = p0 == s[0] - 1 ? _shape[0] - p0 * _partition_shape[0] : _partition_shape[0];
                Count const e1 = p1 == s[1] - 1 ? _shape[1] - p1 * _partition_shape[1] : _partition_shape[1];
                return {e0, e1};
            }

            // Device pointer to read the raster (row-major); mutable_data() to write it.
            auto data() const -> Element const* { return _data.get(); }

            // Device pointer to write the raster: the array's contents count as changed (a flow network kept
            // for it is recorded anew by the next routing operation).
            auto mutable_data() -> Element*
            {
                ++_version;
                return _data.get();
            }

            auto valid() const -> bool { return static_cast<bool>(_data); }

            // Whether the operation that wrote the array has completed.
            auto ready() const -> bool
            {
                if (!_ready)
                {
                    return true;
                }
                int complete = 0;
                detail::check(lue_amd_event_query(_ready->handle(), &complete));
                return complete != 0;
            }

            // Wait (host) for the operation that wrote the array.
            void wait() const
            {
                if (_ready)
                {
                    detail::check(lue_amd_event_synchronize(_ready->handle()));
                }
            }

            // Make `stream` wait for the operation that wrote the array.
            void order_before(void* stream) const
            {
                if (_ready)
                {
                    detail::check(lue_amd_stream_wait_event(stream, _ready->handle()));
                }
            }

            // The array was written by work enqueued on `stream` up to now.
            void written_on(void* stream)
            {
                auto event = std::make_shared<detail::Event>();
                detail::check(lue_amd_event_record(event->handle(), stream));
                _ready = std::move(event);
            }

            // Contents version: changes with every write through from_host, set_partitions or mutable_data.
            auto version() const -> std::uint64_t { return _version; }

            // Whole raster, row-major (waits for the operation that wrote it).
            auto to_host() const -> std::vector<Element>
            {
                std::vector<Element> host(static_cast<std::size_t>(nr_elements()));
                void* const s = lue::stream();
                order_before(s);
                detail::check(lue_amd_copy_to_host(host.data(), _data.get(), sizeof(Element) * host.size(), s));
                return host;
            }

            // Overwrite the raster (in the front end's stream order, after the operations enqueued before).
            void from_host(std::vector<Element> const& host)
            {
                if (static_cast<Count>(host.size()) != nr_elements())
                {
                    throw std::runtime_error("host buffer size differs from the array size");
                }
                void* const s = lue::stream();
                order_before(s);
                detail::check(
                    lue_amd_copy_to_device(mutable_data(), host.data(), sizeof(Element) * host.size(), s));
                _ready.reset();
            }

            // The flow network of this (ldd) array with tiles of `tile_shape` (LUE_AMD_TILE_*) for the routing
            // operations, recorded at the first call that asks for that shape.
            auto network(int tile_shape) const -> lue_amd_network
            {
                if (tile_shape < 0 || tile_shape >= static_cast<int>(_networks.size()))
                {
                    throw std::invalid_argument("unknown tile shape");
                }
                auto& cached = _networks[static_cast<std::size_t>(tile_shape)];
                if (!cached.first || cached.second != _version)
                {
                    void* const s = lue::stream();
                    order_before(s);
                    auto handle = std::make_shared<detail::NetworkHandle>();
                    std::uint8_t const* ldd = reinterpret_cast<std::uint8_t const*>(_data.get());
                    std::int64_t const rows = _shape[0];
                    std::int64_t const cols = _shape[1];
                    detail::check(lue_amd_network_create(1, &ldd, &rows, &cols, tile_shape, LUE_AMD_SCHEDULE_CHAINS,
                                                         s, &handle->network));
                    cached = {std::move(handle), _version};
                }
                return cached.first->network;
            }

            // Set partition by partition from per-partition row-major values, partitions in row-major
            // order (the layout lue::test::create_partitioned_array takes, test/array.hpp:76-105).
            void set_partitions(std::vector<std::vector<Element>> const& partitions)
            {
                if (static_cast<Count>(partitions.size()) != nr_partitions())
                {
                    throw std::runtime_error("number of partitions differs");
                }
                std::vector<Element> host(static_cast<std::size_t>(nr_elements()));
                auto const s = shape_in_partitions();
                std::size_t k = 0;
                for (Index p0 = 0; p0 < s[0]; ++p0)
                {
                    for (Index p1 = 0; p1 < s[1]; ++p1, ++k)
                    {
                        auto const e = partition_extent(p0, p1);
                        if (static_cast<Count>(partitions[k].size()) != e[0] * e[1])
                        {
                            throw std::runtime_error("partition " + std::to_string(k) + " has the wrong size");
                        }
                        for (Index i = 0; i < e[0]; ++i)
                        {
                            for (Index j = 0; j < e[1]; ++j)
                            {
                                host[static_cast<std::size_t>((p0 * _partition_shape[0] + i) * _shape[1] +
                                                              p1 * _partition_shape[1] + j)] =
                                    partitions[k][static_cast<std::size_t>(i * e[1] + j)];
                            }
                        }
                    }
                }
                from_host(host);
            }

        private:

            Shape _shape{0, 0};
            Shape _partition_shape{1, 1};
            std::shared_ptr<Element> _data;
            std::shared_ptr<detail::Event> _ready;  // completion of the operation that wrote the array
            std::uint64_t _version{0};
            // ldd arrays: recorded flow networks per tile shape, with the version of the array they were recorded
            // from
            mutable std::array<std::pair<std::shared_ptr<detail::NetworkHandle>, std::uint64_t>, 4> _networks{};
    };

    template<typename Element>
    auto create_partitioned_array(
        std::array<Count, 2> const& shape, std::array<Count, 2> const& partition_shape, Element fill_value)
        -> PartitionedArray<Element, 2>
    {
        PartitionedArray<Element, 2> array{shape, partition_shape};
        array.from_host(std::vector<Element>(static_cast<std::size_t>(array.nr_elements()), fill_value));
        return array;
    }

    namespace detail {

        // verify_compatible (ALG/detail/verify_compatible.hpp:22-41)
        template<typename E1, typename E2>
        void verify_compatible(PartitionedArray<E1, 2> const& a, PartitionedArray<E2, 2> const& b)
        {
            if (a.shape() != b.shape())
            {
                throw std::runtime_error("Shapes differ: " + shape_string(a.shape()) + " != " + shape_string(b.shape()));
            }
            if (a.shape_in_partitions() != b.shape_in_partitions())
            {
                throw std::runtime_error("Shapes in partitions differ: " + shape_string(a.shape_in_partitions()) +
                                         " != " + shape_string(b.shape_in_partitions()));
            }
        }

        template<typename E1, typename E2>
        void verify_compatible(PartitionedArray<E1, 2> const&, Scalar<E2> const&)
        {
        }

        template<typename E>
        auto pointer(PartitionedArray<E, 2> const& a) -> E const*
        {
            return a.data();
        }

        template<typename E>
        auto pointer(Scalar<E> const&) -> E const*
        {
            return nullptr;
        }

        template<typename E>
        auto scalar(PartitionedArray<E, 2> const&) -> E
        {
            return E{0};
        }

        template<typename E>
        auto scalar(Scalar<E> const& s) -> E
        {
            return s.value();
        }

        template<typename E>
        auto output_like(PartitionedArray<FlowDirectionElement, 2> const& ldd) -> PartitionedArray<E, 2>
        {
            return PartitionedArray<E, 2>{ldd.shape(), ldd.partition_shape()};
        }

        // Make the stream wait for an input array (a Scalar needs nothing).
        template<typename E>
        void order_before(PartitionedArray<E, 2> const& a, void* stream)
        {
            a.order_before(stream);
        }

        template<typename E>
        void order_before(Scalar<E> const&, void*)
        {
        }

        template<typename T>
        concept RasterOrScalar = requires(T const& t) { pointer(t); };

    }  // namespace detail

    // lue::inflow_count<CountElement>(flow_direction) for CountElement = uint8 (ALG/inflow_count.hpp:29-42)
    template<typename CountElement>
        requires std::same_as<CountElement, std::uint8_t>
    auto inflow_count(PartitionedArray<FlowDirectionElement, 2> const& flow_direction)
        -> PartitionedArray<CountElement, 2>
    {
        auto result = detail::output_like<CountElement>(flow_direction);
        void* const s = lue::stream();
        flow_direction.order_before(s);
        detail::check(lue_amd_inflow_count_u8(flow_direction.data(), flow_direction.shape()[0],
                                              flow_direction.shape()[1], result.mutable_data(), s));
        result.written_on(s);
        return result;
    }

    // Cell classes of accu_info (ALG/accu_info.hpp:22-63)
    using CellClass = std::uint8_t;
    inline constexpr CellClass ridge_cell = 11;
    inline constexpr CellClass intra_partition_stream_cell = 12;
    inline constexpr CellClass partition_output_cell = 13;
    inline constexpr CellClass partition_input_cell = 14;
    inline constexpr CellClass partition_input_output_cell = 15;
    inline constexpr CellClass inter_partition_stream_cell = 16;
    inline constexpr CellClass sink_cell = 17;
    inline constexpr CellClass confluence_cell = 18;

    namespace value_policies {

        // lue::value_policies::accu_info(flow_direction) (ALG/value_policies/accu_info.hpp:19-28): the
        // class the accumulating router leaves in each cell for the array's partition shape
        inline auto accu_info(PartitionedArray<FlowDirectionElement, 2> const& flow_direction)
            -> PartitionedArray<CellClass, 2>
        {
            auto result = detail::output_like<CellClass>(flow_direction);
            void* const s = lue::stream();
            flow_direction.order_before(s);
            detail::check(lue_amd_accu_info_u8(flow_direction.data(), flow_direction.shape()[0],
                                               flow_direction.shape()[1], flow_direction.partition_shape()[0],
                                               flow_direction.partition_shape()[1], result.mutable_data(), s));
            result.written_on(s);
            return result;
        }

    }  // namespace value_policies

    namespace detail {

        // the C ABI's upstream / downstream per material element type (LUE's default material elements,
        // environment/cmake/LueConfiguration.cmake:171-228)
        template<typename E>
        struct FlowNeighbourOps;

#define LUE_AMD_FLOW_NEIGHBOUR_OPS(T, SUFFIX)                                                                 \
    template<>                                                                                               \
    struct FlowNeighbourOps<T>                                                                               \
    {                                                                                                        \
        static constexpr auto upstream = &lue_amd_upstream_##SUFFIX;                                        \
        static constexpr auto downstream = &lue_amd_downstream_##SUFFIX;                                    \
    };
        LUE_AMD_FLOW_NEIGHBOUR_OPS(std::uint8_t, u8)
        LUE_AMD_FLOW_NEIGHBOUR_OPS(std::uint32_t, u32)
        LUE_AMD_FLOW_NEIGHBOUR_OPS(std::uint64_t, u64)
        LUE_AMD_FLOW_NEIGHBOUR_OPS(std::int8_t, i8)
        LUE_AMD_FLOW_NEIGHBOUR_OPS(std::int32_t, i32)
        LUE_AMD_FLOW_NEIGHBOUR_OPS(std::int64_t, i64)
        LUE_AMD_FLOW_NEIGHBOUR_OPS(float, f32)
        LUE_AMD_FLOW_NEIGHBOUR_OPS(double, f64)
#undef LUE_AMD_FLOW_NEIGHBOUR_OPS

        template<typename E>
        concept MaterialElement = requires { FlowNeighbourOps<E>::upstream; };

        template<bool upstream, typename E>
        auto flow_neighbour(PartitionedArray<FlowDirectionElement, 2> const& flow_direction,
                            PartitionedArray<E, 2> const& material) -> PartitionedArray<E, 2>
        {
            verify_compatible(flow_direction, material);
            auto result = output_like<E>(flow_direction);
            auto const [rows, cols] = flow_direction.shape();
            auto const fn = upstream ? FlowNeighbourOps<E>::upstream : FlowNeighbourOps<E>::downstream;
            void* const s = lue::stream();
            flow_direction.order_before(s);
            material.order_before(s);
            check(fn(flow_direction.data(), material.data(), rows, cols, result.mutable_data(), s));
            result.written_on(s);
            return result;
        }

    }  // namespace detail

    // lue::downstream_distance(flow_direction, cell_size) (ALG/downstream_distance.hpp:31-39, default policies:
    // a no-data ldd is an unknown code, distance 0; ALG/definition/downstream_distance.hpp:13-75): per cell,
    // the distance to its downstream cell
    template<std::integral FlowDirection, std::floating_point Distance>
        requires std::same_as<FlowDirection, FlowDirectionElement>
    auto downstream_distance(PartitionedArray<FlowDirection, 2> const& flow_direction, Distance const cell_size)
        -> PartitionedArray<Distance, 2>
    {
        auto result = detail::output_like<Distance>(flow_direction);
        auto const [rows, cols] = flow_direction.shape();
        void* const s = lue::stream();
        flow_direction.order_before(s);
        if constexpr (std::is_same_v<Distance, float>)
            detail::check(lue_amd_downstream_distance_f32(flow_direction.data(), rows, cols, cell_size, 0,
                                                          result.mutable_data(), s));
        else
            detail::check(lue_amd_downstream_distance_f64(flow_direction.data(), rows, cols, cell_size, 0,
                                                          result.mutable_data(), s));
        result.written_on(s);
        return result;
    }

    namespace value_policies {

        // lue::value_policies::upstream(flow_direction, material) (ALG/value_policies/upstream.hpp:16-30,
        // ALG/definition/upstream.hpp:36-190): per cell, the sum of the material of the neighbours that
        // drain into it
        template<std::integral FlowDirection, lue::detail::MaterialElement E>
            requires std::same_as<FlowDirection, FlowDirectionElement>
        auto upstream(PartitionedArray<FlowDirection, 2> const& flow_direction, PartitionedArray<E, 2> const& material)
            -> PartitionedArray<E, 2>
        {
            return lue::detail::flow_neighbour<true>(flow_direction, material);
        }

        // lue::value_policies::downstream(flow_direction, material) (ALG/value_policies/downstream.hpp:16-30,
        // ALG/definition/downstream.hpp:36-186): per cell, the material of its downstream neighbour
        template<std::integral FlowDirection, lue::detail::MaterialElement E>
            requires std::same_as<FlowDirection, FlowDirectionElement>
        auto downstream(PartitionedArray<FlowDirection, 2> const& flow_direction,
                        PartitionedArray<E, 2> const& material) -> PartitionedArray<E, 2>
        {
            return lue::detail::flow_neighbour<false>(flow_direction, material);
        }

        // lue::value_policies::d8_flow_direction<FlowDirection>(elevation)
        // (ALG/value_policies/d8_flow_direction.hpp:46-55): the ldd of an elevation raster, NaN = no-data
        template<std::integral FlowDirection, std::floating_point E>
            requires std::same_as<FlowDirection, FlowDirectionElement>
        auto d8_flow_direction(PartitionedArray<E, 2> const& elevation) -> PartitionedArray<FlowDirection, 2>
        {
            PartitionedArray<FlowDirection, 2> result{elevation.shape(), elevation.partition_shape()};
            auto const [rows, cols] = elevation.shape();
            void* const s = lue::stream();
            elevation.order_before(s);
            if constexpr (std::is_same_v<E, float>)
                detail::check(lue_amd_d8_flow_direction_f32(elevation.data(), rows, cols, result.mutable_data(), s));
            else
                detail::check(lue_amd_d8_flow_direction_f64(elevation.data(), rows, cols, result.mutable_data(), s));
            result.written_on(s);
            return result;
        }

        // ALG/value_policies/accu.hpp:33-55
        template<std::integral FlowDirection, std::floating_point E, typename Inflow>
            requires std::same_as<FlowDirection, FlowDirectionElement> &&
                     (std::same_as<Inflow, PartitionedArray<E, 2>> || std::same_as<Inflow, Scalar<E>>)
        auto accu_impl(PartitionedArray<FlowDirection, 2> const& flow_direction, Inflow const& inflow)
            -> PartitionedArray<E, 2>
        {
            detail::verify_compatible(flow_direction, inflow);
            auto outflow = detail::output_like<E>(flow_direction);
            lue_amd_network const network = flow_direction.network(LUE_AMD_TILE_64X32);
            void* const s = lue::stream();
            detail::order_before(inflow, s);
            E const* const in[1] = {detail::pointer(inflow)};
            E* const out[1] = {outflow.mutable_data()};
            if constexpr (std::is_same_v<E, float>)
            {
                detail::check(lue_amd_network_accu_f32(network, in, detail::scalar(inflow), out, s));
            }
            else
            {
                detail::check(lue_amd_network_accu_f64(network, in, detail::scalar(inflow), out, s));
            }
            outflow.written_on(s);
            return outflow;
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu(PartitionedArray<FlowDirection, 2> const& flow_direction, PartitionedArray<E, 2> const& inflow)
            -> PartitionedArray<E, 2>
        {
            return accu_impl<FlowDirection, E>(flow_direction, inflow);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu(PartitionedArray<FlowDirection, 2> const& flow_direction, Scalar<E> const& inflow)
            -> PartitionedArray<E, 2>
        {
            return accu_impl<FlowDirection, E>(flow_direction, inflow);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu(PartitionedArray<FlowDirection, 2> const& flow_direction, E const inflow) -> PartitionedArray<E, 2>
        {
            return accu_impl<FlowDirection, E>(flow_direction, Scalar<E>{inflow});
        }

        namespace detail2 {

            // the two-output operations: (outflow, remainder)
            enum class Split { threshold, fraction, capacity, trigger };

            template<Split kind, std::floating_point E, typename Inflow, typename Param>
            auto split(PartitionedArray<FlowDirectionElement, 2> const& flow_direction, Inflow const& inflow,
                       Param const& param) -> std::tuple<PartitionedArray<E, 2>, PartitionedArray<E, 2>>
            {
                lue::detail::verify_compatible(flow_direction, inflow);
                lue::detail::verify_compatible(flow_direction, param);
                auto outflow = lue::detail::output_like<E>(flow_direction);
                auto remainder = lue::detail::output_like<E>(flow_direction);
                lue_amd_network const net = flow_direction.network(LUE_AMD_TILE_64X32);
                void* const s = lue::stream();
                lue::detail::order_before(inflow, s);
                lue::detail::order_before(param, s);
                E const* const p1[1] = {lue::detail::pointer(inflow)};
                auto const s1 = lue::detail::scalar(inflow);
                E const* const p2[1] = {lue::detail::pointer(param)};
                auto const s2 = lue::detail::scalar(param);
                E* const out[1] = {outflow.mutable_data()};
                E* const rem[1] = {remainder.mutable_data()};
                int status;
                if constexpr (std::is_same_v<E, float>)
                {
                    if constexpr (kind == Split::threshold)
                        status = lue_amd_network_accu_threshold_f32(net, p1, s1, p2, s2, out, rem, s);
                    else if constexpr (kind == Split::fraction)
                        status = lue_amd_network_accu_fraction_f32(net, p1, s1, p2, s2, out, rem, s);
                    else if constexpr (kind == Split::capacity)
                        status = lue_amd_network_accu_capacity_f32(net, p1, s1, p2, s2, out, rem, s);
                    else
                        status = lue_amd_network_accu_trigger_f32(net, p1, s1, p2, s2, out, rem, s);
                }
                else
                {
                    if constexpr (kind == Split::threshold)
                        status = lue_amd_network_accu_threshold_f64(net, p1, s1, p2, s2, out, rem, s);
                    else if constexpr (kind == Split::fraction)
                        status = lue_amd_network_accu_fraction_f64(net, p1, s1, p2, s2, out, rem, s);
                    else if constexpr (kind == Split::capacity)
                        status = lue_amd_network_accu_capacity_f64(net, p1, s1, p2, s2, out, rem, s);
                    else
                        status = lue_amd_network_accu_trigger_f64(net, p1, s1, p2, s2, out, rem, s);
                }
                lue::detail::check(status);
                outflow.written_on(s);
                remainder.written_on(s);
                return {std::move(outflow), std::move(remainder)};
            }

        }  // namespace detail2

        // ALG/value_policies/accu_threshold.hpp:34-87: raster|Scalar inflow x raster|Scalar threshold
        template<std::integral FlowDirection, std::floating_point E>
        auto accu_threshold(PartitionedArray<FlowDirection, 2> const& flow_direction, PartitionedArray<E, 2> const& inflow,
                     PartitionedArray<E, 2> const& threshold)
        {
            return detail2::split<detail2::Split::threshold, E>(flow_direction, inflow, threshold);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_threshold(PartitionedArray<FlowDirection, 2> const& flow_direction, PartitionedArray<E, 2> const& inflow,
                     Scalar<E> const& threshold)
        {
            return detail2::split<detail2::Split::threshold, E>(flow_direction, inflow, threshold);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_threshold(PartitionedArray<FlowDirection, 2> const& flow_direction, Scalar<E> const& inflow,
                     Scalar<E> const& threshold)
        {
            return detail2::split<detail2::Split::threshold, E>(flow_direction, inflow, threshold);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_threshold(PartitionedArray<FlowDirection, 2> const& flow_direction, Scalar<E> const& inflow,
                     PartitionedArray<E, 2> const& threshold)
        {
            return detail2::split<detail2::Split::threshold, E>(flow_direction, inflow, threshold);
        }

        // ALG/value_policies/accu_fraction.hpp: the same four combinations
        template<std::integral FlowDirection, std::floating_point E>
        auto accu_fraction(PartitionedArray<FlowDirection, 2> const& flow_direction, PartitionedArray<E, 2> const& inflow,
                     PartitionedArray<E, 2> const& fraction)
        {
            return detail2::split<detail2::Split::fraction, E>(flow_direction, inflow, fraction);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_fraction(PartitionedArray<FlowDirection, 2> const& flow_direction, PartitionedArray<E, 2> const& inflow,
                     Scalar<E> const& fraction)
        {
            return detail2::split<detail2::Split::fraction, E>(flow_direction, inflow, fraction);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_fraction(PartitionedArray<FlowDirection, 2> const& flow_direction, Scalar<E> const& inflow,
                     Scalar<E> const& fraction)
        {
            return detail2::split<detail2::Split::fraction, E>(flow_direction, inflow, fraction);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_fraction(PartitionedArray<FlowDirection, 2> const& flow_direction, Scalar<E> const& inflow,
                     PartitionedArray<E, 2> const& fraction)
        {
            return detail2::split<detail2::Split::fraction, E>(flow_direction, inflow, fraction);
        }

        // ALG/value_policies/accu_capacity.hpp:34-88: the same four combinations
        template<std::integral FlowDirection, std::floating_point E>
        auto accu_capacity(PartitionedArray<FlowDirection, 2> const& flow_direction, PartitionedArray<E, 2> const& inflow,
                     PartitionedArray<E, 2> const& capacity)
        {
            return detail2::split<detail2::Split::capacity, E>(flow_direction, inflow, capacity);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_capacity(PartitionedArray<FlowDirection, 2> const& flow_direction, PartitionedArray<E, 2> const& inflow,
                     Scalar<E> const& capacity)
        {
            return detail2::split<detail2::Split::capacity, E>(flow_direction, inflow, capacity);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_capacity(PartitionedArray<FlowDirection, 2> const& flow_direction, Scalar<E> const& inflow,
                     Scalar<E> const& capacity)
        {
            return detail2::split<detail2::Split::capacity, E>(flow_direction, inflow, capacity);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_capacity(PartitionedArray<FlowDirection, 2> const& flow_direction, Scalar<E> const& inflow,
                     PartitionedArray<E, 2> const& capacity)
        {
            return detail2::split<detail2::Split::capacity, E>(flow_direction, inflow, capacity);
        }

        // ALG/value_policies/accu_trigger.hpp:34-88: the same four combinations
        template<std::integral FlowDirection, std::floating_point E>
        auto accu_trigger(PartitionedArray<FlowDirection, 2> const& flow_direction, PartitionedArray<E, 2> const& inflow,
                     PartitionedArray<E, 2> const& trigger)
        {
            return detail2::split<detail2::Split::trigger, E>(flow_direction, inflow, trigger);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_trigger(PartitionedArray<FlowDirection, 2> const& flow_direction, PartitionedArray<E, 2> const& inflow,
                     Scalar<E> const& trigger)
        {
            return detail2::split<detail2::Split::trigger, E>(flow_direction, inflow, trigger);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_trigger(PartitionedArray<FlowDirection, 2> const& flow_direction, Scalar<E> const& inflow,
                     Scalar<E> const& trigger)
        {
            return detail2::split<detail2::Split::trigger, E>(flow_direction, inflow, trigger);
        }

        template<std::integral FlowDirection, std::floating_point E>
        auto accu_trigger(PartitionedArray<FlowDirection, 2> const& flow_direction, Scalar<E> const& inflow,
                     PartitionedArray<E, 2> const& trigger)
        {
            return detail2::split<detail2::Split::trigger, E>(flow_direction, inflow, trigger);
        }

        namespace detail2 {

            template<std::floating_point E, typename A, typename B, typename L>
            auto kinematic_wave(PartitionedArray<FlowDirectionElement, 2> const& flow_direction,
                                PartitionedArray<E, 2> const& current_outflow, PartitionedArray<E, 2> const& inflow,
                                A const& alpha, B const& beta, Scalar<E> const& time_step_duration,
                                L const& channel_length) -> PartitionedArray<E, 2>
            {
                lue::detail::verify_compatible(flow_direction, current_outflow);
                lue::detail::verify_compatible(flow_direction, inflow);
                lue::detail::verify_compatible(flow_direction, alpha);
                lue::detail::verify_compatible(flow_direction, beta);
                lue::detail::verify_compatible(flow_direction, channel_length);
                auto outflow = lue::detail::output_like<E>(flow_direction);
                lue_amd_network const net = flow_direction.network(LUE_AMD_TILE_32X32);
                void* const s = lue::stream();
                current_outflow.order_before(s);
                inflow.order_before(s);
                lue::detail::order_before(alpha, s);
                lue::detail::order_before(beta, s);
                lue::detail::order_before(channel_length, s);
                using lue::detail::pointer;
                using lue::detail::scalar;
                E const* const cur[1] = {current_outflow.data()};
                E const* const lat[1] = {inflow.data()};
                E const* const a[1] = {pointer(alpha)};
                E const* const b[1] = {pointer(beta)};
                E const* const dx[1] = {pointer(channel_length)};
                E* const out[1] = {outflow.mutable_data()};
                int status;
                if constexpr (std::is_same_v<E, float>)
                    status = lue_amd_network_kinematic_wave_f32(net, cur, lat, a, scalar(alpha), b, scalar(beta),
                                                                time_step_duration.value(), dx,
                                                                scalar(channel_length), out, s);
                else
                    status = lue_amd_network_kinematic_wave_f64(net, cur, lat, a, scalar(alpha), b, scalar(beta),
                                                                time_step_duration.value(), dx,
                                                                scalar(channel_length), out, s);
                lue::detail::check(status);
                outflow.written_on(s);
                return outflow;
            }

        }  // namespace detail2

        // ALG/value_policies/kinematic_wave.hpp:43-91: rasters alpha, beta, channel length
        template<std::integral FlowDirection, std::floating_point E>
        auto kinematic_wave(PartitionedArray<FlowDirection, 2> const& flow_direction,
                            PartitionedArray<E, 2> const& current_outflow, PartitionedArray<E, 2> const& inflow,
                            PartitionedArray<E, 2> const& alpha, PartitionedArray<E, 2> const& beta,
                            Scalar<E> const& time_step_duration, PartitionedArray<E, 2> const& channel_length)
            -> PartitionedArray<E, 2>
        {
            return detail2::kinematic_wave<E>(
                flow_direction, current_outflow, inflow, alpha, beta, time_step_duration, channel_length);
        }

        // ... and Scalars alpha, beta, time step duration, channel length
        template<std::integral FlowDirection, std::floating_point E>
        auto kinematic_wave(PartitionedArray<FlowDirection, 2> const& flow_direction,
                            PartitionedArray<E, 2> const& current_outflow, PartitionedArray<E, 2> const& inflow,
                            Scalar<E> const& alpha, Scalar<E> const& beta, Scalar<E> const& time_step_duration,
                            Scalar<E> const& channel_length) -> PartitionedArray<E, 2>
        {
            return detail2::kinematic_wave<E>(
                flow_direction, current_outflow, inflow, alpha, beta, time_step_duration, channel_length);
        }

    }  // namespace value_policies

}  // namespace lue
