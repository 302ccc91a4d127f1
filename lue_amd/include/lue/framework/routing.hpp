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
//     d8_flow_direction, upstream, downstream} with the reference's overload sets (ALG/value_policies/*.hpp) and
//     lue::inflow_count<std::uint8_t>(ldd) and lue::downstream_distance(ldd, cell_size)
//     (ALG/inflow_count.hpp:29-42), ALG = source/framework/algorithm/include/lue/framework/algorithm.
// Each operation calls one function of the C ABI (include/lue_amd/routing.h). Shape and partitioning
// mismatches throw std::runtime_error as detail::verify_compatible does (ALG/detail/verify_compatible.hpp:
// 22-41); other failures throw std::runtime_error with the library's message. Results are complete when a
// call returns.
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

    }  // namespace detail

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

            auto data() const -> Element* { return _data.get(); }

            auto valid() const -> bool { return static_cast<bool>(_data); }

            // Whole raster, row-major.
            auto to_host() const -> std::vector<Element>
            {
                std::vector<Element> host(static_cast<std::size_t>(nr_elements()));
                detail::check(lue_amd_copy_to_host(host.data(), _data.get(), sizeof(Element) * host.size(), nullptr));
                return host;
            }

            void from_host(std::vector<Element> const& host)
            {
                if (static_cast<Count>(host.size()) != nr_elements())
                {
                    throw std::runtime_error("host buffer size differs from the array size");
                }
                detail::check(
                    lue_amd_copy_to_device(_data.get(), host.data(), sizeof(Element) * host.size(), nullptr));
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
        detail::check(lue_amd_inflow_count_u8(
            flow_direction.data(), flow_direction.shape()[0], flow_direction.shape()[1], result.data(), nullptr));
        return result;
    }

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
            check(fn(flow_direction.data(), material.data(), rows, cols, result.data(), nullptr));
            check(lue_amd_stream_synchronize(nullptr));
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
        if constexpr (std::is_same_v<Distance, float>)
            detail::check(lue_amd_downstream_distance_f32(flow_direction.data(), rows, cols, cell_size, 0,
                                                          result.data(), nullptr));
        else
            detail::check(lue_amd_downstream_distance_f64(flow_direction.data(), rows, cols, cell_size, 0,
                                                          result.data(), nullptr));
        detail::check(lue_amd_stream_synchronize(nullptr));
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
            if constexpr (std::is_same_v<E, float>)
                detail::check(lue_amd_d8_flow_direction_f32(elevation.data(), rows, cols, result.data(), nullptr));
            else
                detail::check(lue_amd_d8_flow_direction_f64(elevation.data(), rows, cols, result.data(), nullptr));
            detail::check(lue_amd_stream_synchronize(nullptr));
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
            auto const [rows, cols] = flow_direction.shape();
            if constexpr (std::is_same_v<E, float>)
            {
                detail::check(lue_amd_accu_f32(flow_direction.data(), rows, cols, detail::pointer(inflow),
                                               detail::scalar(inflow), outflow.data(), nullptr));
            }
            else
            {
                detail::check(lue_amd_accu_f64(flow_direction.data(), rows, cols, detail::pointer(inflow),
                                               detail::scalar(inflow), outflow.data(), nullptr));
            }
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
                auto const [rows, cols] = flow_direction.shape();
                auto const* ldd = flow_direction.data();
                auto const* p1 = lue::detail::pointer(inflow);
                auto const s1 = lue::detail::scalar(inflow);
                auto const* p2 = lue::detail::pointer(param);
                auto const s2 = lue::detail::scalar(param);
                E* const out = outflow.data();
                E* const rem = remainder.data();
                int status;
                if constexpr (std::is_same_v<E, float>)
                {
                    if constexpr (kind == Split::threshold)
                        status = lue_amd_accu_threshold_f32(ldd, rows, cols, p1, s1, p2, s2, out, rem, nullptr);
                    else if constexpr (kind == Split::fraction)
                        status = lue_amd_accu_fraction_f32(ldd, rows, cols, p1, s1, p2, s2, out, rem, nullptr);
                    else if constexpr (kind == Split::capacity)
                        status = lue_amd_accu_capacity_f32(ldd, rows, cols, p1, s1, p2, s2, out, rem, nullptr);
                    else
                        status = lue_amd_accu_trigger_f32(ldd, rows, cols, p1, s1, p2, s2, out, rem, nullptr);
                }
                else
                {
                    if constexpr (kind == Split::threshold)
                        status = lue_amd_accu_threshold_f64(ldd, rows, cols, p1, s1, p2, s2, out, rem, nullptr);
                    else if constexpr (kind == Split::fraction)
                        status = lue_amd_accu_fraction_f64(ldd, rows, cols, p1, s1, p2, s2, out, rem, nullptr);
                    else if constexpr (kind == Split::capacity)
                        status = lue_amd_accu_capacity_f64(ldd, rows, cols, p1, s1, p2, s2, out, rem, nullptr);
                    else
                        status = lue_amd_accu_trigger_f64(ldd, rows, cols, p1, s1, p2, s2, out, rem, nullptr);
                }
                lue::detail::check(status);
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
                auto const [rows, cols] = flow_direction.shape();
                using lue::detail::pointer;
                using lue::detail::scalar;
                int status;
                if constexpr (std::is_same_v<E, float>)
                    status = lue_amd_kinematic_wave_f32(
                        flow_direction.data(), rows, cols, current_outflow.data(), inflow.data(), pointer(alpha),
                        scalar(alpha), pointer(beta), scalar(beta), time_step_duration.value(),
                        pointer(channel_length), scalar(channel_length), outflow.data(), nullptr);
                else
                    status = lue_amd_kinematic_wave_f64(
                        flow_direction.data(), rows, cols, current_outflow.data(), inflow.data(), pointer(alpha),
                        scalar(alpha), pointer(beta), scalar(beta), time_step_duration.value(),
                        pointer(channel_length), scalar(channel_length), outflow.data(), nullptr);
                lue::detail::check(status);
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
