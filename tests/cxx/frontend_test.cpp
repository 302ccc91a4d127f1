// Golden cases of LUE's routing tests, run through the C++ front end (lue::value_policies::* on
// lue::PartitionedArray) and the C ABI. The cases (golden_cases.inc) are generated by tests/test_cxx_frontend.py
// from tests/golden/*.json. Prints one line per case and exits non-zero if any case fails.
#include "lue/framework/routing.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <functional>
#include <limits>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace {

    using Shape = std::array<lue::Count, 2>;

    struct Arg
    {
        bool raster;
        double scalar;
        std::vector<double> values;
    };

    struct Case
    {
        char const* name;
        char const* kind;
        Shape shape;
        Shape partition_shape;
        std::vector<double> ldd;  // the elevation for d8_flow_direction
        std::vector<Arg> args;
        std::vector<std::vector<double>> expected;  // outflow [, remainder] | count
        bool close;                                  // Boost tolerance 1e-6 instead of exact
        bool all_positive;
    };

    double const nan = std::numeric_limits<double>::quiet_NaN();

#include "golden_cases.inc"

    template<typename E>
    auto make_raster(Case const& c, std::vector<double> const& values) -> lue::PartitionedArray<E, 2>
    {
        lue::PartitionedArray<E, 2> a{c.shape, c.partition_shape};
        std::vector<E> host(values.begin(), values.end());
        a.from_host(host);
        return a;
    }

    bool equal(double got, double want, bool close)
    {
        if (std::isnan(got) || std::isnan(want))
            return std::isnan(got) && std::isnan(want);
        if (!close)
            return got == want;
        if (got == 0 || want == 0)
            return std::fabs(got) < 1e-6 && std::fabs(want) < 1e-6;
        double const d = std::fabs(got - want);
        return d / std::fabs(got) <= 1e-6 && d / std::fabs(want) <= 1e-6;
    }

    template<typename E>
    bool compare(Case const& c, lue::PartitionedArray<E, 2> const& got, std::vector<double> const& want)
    {
        auto const host = got.to_host();
        for (std::size_t i = 0; i < host.size(); ++i)
        {
            if (!equal(static_cast<double>(host[i]), want[i], c.close))
            {
                std::printf("  %s: cell %zu got %.9g want %.9g\n", c.name, i, static_cast<double>(host[i]), want[i]);
                return false;
            }
        }
        return true;
    }

    // no-data of a material element: NaN, or the minimum (signed) / maximum (unsigned) integer; the cases
    // write no-data as NaN whatever the element type
    template<typename E>
    E no_data_of()
    {
        if constexpr (std::is_floating_point_v<E>)
            return std::numeric_limits<E>::quiet_NaN();
        else if constexpr (std::is_signed_v<E>)
            return std::numeric_limits<E>::min();
        else
            return std::numeric_limits<E>::max();
    }

    template<typename E>
    E material_value(double v)
    {
        return std::isnan(v) ? no_data_of<E>() : static_cast<E>(v);
    }

    // upstream / downstream for one material element type
    template<typename E>
    bool run_flow_neighbour(Case const& c)
    {
        using namespace lue;
        auto const ldd = make_raster<std::uint8_t>(c, c.ldd);
        PartitionedArray<E, 2> material{c.shape, c.partition_shape};
        std::vector<E> host;
        for (double v : c.args[0].values)
            host.push_back(material_value<E>(v));
        material.from_host(host);
        auto const got = std::string(c.kind) == "upstream" ? value_policies::upstream(ldd, material)
                                                           : value_policies::downstream(ldd, material);
        auto const values = got.to_host();
        for (std::size_t i = 0; i < values.size(); ++i)
        {
            E const want = material_value<E>(c.expected[0][i]);
            bool const same = std::is_floating_point_v<E>
                                  ? ((std::isnan(static_cast<double>(values[i])) &&
                                      std::isnan(static_cast<double>(want))) ||
                                     values[i] == want)
                                  : values[i] == want;
            if (!same)
            {
                std::printf("  %s: cell %zu got %.17g want %.17g\n", c.name, i, static_cast<double>(values[i]),
                            static_cast<double>(want));
                return false;
            }
        }
        return true;
    }

    template<typename E>
    bool run(Case const& c)
    {
        using namespace lue;
        std::string const kind = c.kind;
        if (kind == "d8_flow_direction")
        {
            // c.ldd holds the elevation here
            auto const elevation = make_raster<E>(c, c.ldd);
            return compare(c, value_policies::d8_flow_direction<std::uint8_t>(elevation), c.expected[0]);
        }
        auto const ldd = make_raster<std::uint8_t>(c, c.ldd);
        if (kind == "downstream_distance")
        {
            // the reference case holds no no-data ldd, so default and value policies agree. Its diagonal
            // distance q = sqrt(2) * cell_size is written in float; in E it is sqrt(E{2}) * cell_size
            E const cell_size = static_cast<E>(c.args[0].scalar);
            float const q_float = std::sqrt(2.0f) * static_cast<float>(c.args[0].scalar);
            auto const got = downstream_distance(ldd, cell_size).to_host();
            for (std::size_t i = 0; i < got.size(); ++i)
            {
                double const v = c.expected[0][i];
                E const want = static_cast<float>(v) == q_float ? std::sqrt(E{2}) * cell_size : static_cast<E>(v);
                if (got[i] != want)
                {
                    std::printf("  %s: cell %zu got %.9g want %.9g\n", c.name, i, static_cast<double>(got[i]),
                                static_cast<double>(want));
                    return false;
                }
            }
            return true;
        }
        if (kind == "inflow_count")
        {
            auto const count = inflow_count<std::uint8_t>(ldd);
            return compare(c, count, c.expected[0]);
        }
        if (kind == "accu_info")
        {
            PartitionedArray<CellClass, 2> const cell_class = value_policies::accu_info(ldd);
            return compare(c, cell_class, c.expected[0]);
        }
        auto raster = [&](std::size_t k) { return make_raster<E>(c, c.args[k].values); };
        auto scalar = [&](std::size_t k) { return Scalar<E>{static_cast<E>(c.args[k].scalar)}; };
        if (kind == "accu")
        {
            auto out = c.args[0].raster ? value_policies::accu(ldd, raster(0)) : value_policies::accu(ldd, scalar(0));
            return compare(c, out, c.expected[0]);
        }
        if (kind == "accu_threshold" || kind == "accu_fraction" || kind == "accu_capacity" || kind == "accu_trigger")
        {
            auto call = [&](auto const& a, auto const& b)
            {
                if (kind == "accu_threshold")
                    return value_policies::accu_threshold(ldd, a, b);
                if (kind == "accu_fraction")
                    return value_policies::accu_fraction(ldd, a, b);
                if (kind == "accu_capacity")
                    return value_policies::accu_capacity(ldd, a, b);
                return value_policies::accu_trigger(ldd, a, b);
            };
            std::tuple<PartitionedArray<E, 2>, PartitionedArray<E, 2>> result =
                c.args[0].raster ? (c.args[1].raster ? call(raster(0), raster(1)) : call(raster(0), scalar(1)))
                                 : (c.args[1].raster ? call(scalar(0), raster(1)) : call(scalar(0), scalar(1)));
            bool ok = true;
            for (std::size_t k = 0; k < c.expected.size(); ++k)
                ok = compare(c, k == 0 ? std::get<0>(result) : std::get<1>(result), c.expected[k]) && ok;
            return ok;
        }
        if (kind == "kinematic_wave")
        {
            auto full = [&](std::size_t k)
            {
                return c.args[k].raster ? raster(k)
                                        : create_partitioned_array<E>(c.shape, c.partition_shape,
                                                                      static_cast<E>(c.args[k].scalar));
            };
            auto const current = full(0);
            auto const inflow = full(1);
            bool const rasters = c.args[2].raster || c.args[3].raster || c.args[5].raster;
            auto out = rasters ? value_policies::kinematic_wave(ldd, current, inflow, full(2), full(3),
                                                                Scalar<E>{static_cast<E>(c.args[4].scalar)}, full(5))
                               : value_policies::kinematic_wave(ldd, current, inflow, scalar(2), scalar(3),
                                                                scalar(4), scalar(5));
            if (c.all_positive)
            {
                for (E v : out.to_host())
                    if (!(v > 0))
                        return false;
            }
            return c.expected.empty() || compare(c, out, c.expected[0]);
        }
        std::printf("  unknown kind %s\n", c.kind);
        return false;
    }

    bool shape_mismatch_throws()
    {
        auto const ldd = lue::create_partitioned_array<std::uint8_t>({8, 8}, {4, 4}, 2);
        auto const inflow = lue::create_partitioned_array<float>({8, 9}, {4, 4}, 1.0f);
        try
        {
            auto out = lue::value_policies::accu(ldd, inflow);
        }
        catch (std::runtime_error const& e)
        {
            return std::string(e.what()).rfind("Shapes differ", 0) == 0;
        }
        return false;
    }

    // Asynchrony: a kinematic-wave timestep returns before its kernels finish (its output is not ready yet),
    // and a second timestep on another stream, ordered only by the first one's completion event, gives the
    // same bits as the two timesteps run one after the other with the host waiting in between.
    bool asynchronous_timesteps()
    {
        std::int64_t const n = 8192;
        Shape const shape{n, n};
        Shape const ps{1024, 1024};
        lue::PartitionedArray<std::uint8_t, 2> ldd{shape, ps};
        lue::PartitionedArray<float, 2> q0{shape, ps};
        lue::PartitionedArray<float, 2> lateral{shape, ps};
        lue::detail::check(lue_amd_synthetic_ldd(20261209, n, n, ldd.mutable_data(), nullptr));
        lue::detail::check(lue_amd_synthetic_uniform_f32(20261209, 1, 0.0, 5.0, n * n, q0.mutable_data(), nullptr));
        lue::detail::check(
            lue_amd_synthetic_uniform_f32(20261209, 2, 0.0, 1e-4, n * n, lateral.mutable_data(), nullptr));
        lue::detail::check(lue_amd_stream_synchronize(nullptr));
        lue::Scalar<float> const alpha{2.0f}, beta{0.6f}, dt{3600.0f}, dx{100.0f};
        using lue::value_policies::kinematic_wave;

        auto r1 = kinematic_wave(ldd, q0, lateral, alpha, beta, dt, dx);  // records the ldd's network
        r1.wait();
        auto r2 = kinematic_wave(ldd, r1, lateral, alpha, beta, dt, dx);
        auto const want = r2.to_host();

        hipStream_t a = nullptr, b = nullptr;
        if (hipStreamCreateWithFlags(&a, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&b, hipStreamNonBlocking) != hipSuccess)
        {
            std::printf("  stream creation failed\n");
            return false;
        }
        lue::set_stream(a);
        auto q1 = kinematic_wave(ldd, q0, lateral, alpha, beta, dt, dx);
        bool const pending = !q1.ready();
        lue::set_stream(b);
        auto q2 = kinematic_wave(ldd, q1, lateral, alpha, beta, dt, dx);
        auto const got = q2.to_host();
        lue::set_stream(nullptr);
        (void)hipStreamDestroy(a);
        (void)hipStreamDestroy(b);
        bool const same = got == want;
        std::printf("  asynchronous_timesteps: returned before completion %d, same bits %d\n", pending ? 1 : 0,
                    same ? 1 : 0);
        return pending && same;
    }

}  // namespace

int main()
{
    int failed = 0;
    int passed = 0;
    for (Case const& c : cases())
    {
        std::string const kind = c.kind;
        if (kind == "upstream" || kind == "downstream")
        {
            // LUE's cases through every default material element type
            std::pair<char const*, std::function<bool(Case const&)>> const element_runs[] = {
                {"uint8", run_flow_neighbour<std::uint8_t>},   {"uint32", run_flow_neighbour<std::uint32_t>},
                {"uint64", run_flow_neighbour<std::uint64_t>}, {"int8", run_flow_neighbour<std::int8_t>},
                {"int32", run_flow_neighbour<std::int32_t>},   {"int64", run_flow_neighbour<std::int64_t>},
                {"float", run_flow_neighbour<float>},          {"double", run_flow_neighbour<double>}};
            for (auto const& [element, fn] : element_runs)
            {
                bool ok;
                try
                {
                    ok = fn(c);
                }
                catch (std::exception const& e)
                {
                    std::printf("  %s: exception %s\n", c.name, e.what());
                    ok = false;
                }
                std::printf("%s %s %s %s\n", ok ? "PASS" : "FAIL", c.kind, c.name, element);
                (ok ? passed : failed) += 1;
            }
            continue;
        }
        for (int precision = 0; precision < 2; ++precision)
        {
            if ((kind == "inflow_count" || kind == "accu_info") && precision == 1)
                continue;
            bool ok;
            try
            {
                ok = precision == 0 ? run<float>(c) : run<double>(c);
            }
            catch (std::exception const& e)
            {
                std::printf("  %s: exception %s\n", c.name, e.what());
                ok = false;
            }
            std::printf("%s %s %s %s\n", ok ? "PASS" : "FAIL", c.kind, c.name, precision == 0 ? "float" : "double");
            (ok ? passed : failed) += 1;
        }
    }
    bool const ok = shape_mismatch_throws();
    std::printf("%s shape_mismatch_throws\n", ok ? "PASS" : "FAIL");
    (ok ? passed : failed) += 1;
    bool async_ok = false;
    try
    {
        async_ok = asynchronous_timesteps();
    }
    catch (std::exception const& e)
    {
        std::printf("  asynchronous_timesteps: exception %s\n", e.what());
    }
    std::printf("%s asynchronous_timesteps\n", async_ok ? "PASS" : "FAIL");
    (async_ok ? passed : failed) += 1;
    std::printf("%d passed, %d failed\n", passed, failed);
    return failed == 0 ? 0 : 1;
}
