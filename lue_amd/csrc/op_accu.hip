// C ABI of accu (per call and on flow networks), of accu_info (which runs accu of a crossing indicator) and
// of flow-network creation, which records a network's schedule with accu(ldd, 1).
#include "accu_info.hpp"
#include "api.hpp"

using namespace lue_amd;

extern "C" {

int lue_amd_accu_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow, float inflow_scalar,
                     float* outflow, void* stream)
{
    return accu_impl<float>(ldd, rows, cols, inflow, inflow_scalar, outflow, stream);
}

int lue_amd_accu_info_u8(const uint8_t* ldd, int64_t rows, int64_t cols, int64_t partition_rows,
                         int64_t partition_cols, uint8_t* cell_class, void* stream)
{
    if (int const st = check_shape(ldd, rows, cols))
        return st;
    if (!cell_class)
        return fail(status_invalid_argument, "cell_class must not be null");
    if (partition_rows <= 0 || partition_cols <= 0 || partition_rows > rows || partition_cols > cols)
        return fail(status_invalid_argument, "partition shape must be positive and fit in the array shape");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    PartitionGrid const g{partition_rows, partition_cols, rows / partition_rows, cols / partition_cols};
    std::size_t const n = static_cast<std::size_t>(rows * cols);
    ScratchArray<float> crossing;
    ScratchArray<float> crossing_accu;
    LUE_AMD_HIP(crossing.allocate(n, s));
    LUE_AMD_HIP(crossing_accu.allocate(n, s));
    dim3 const grid(static_cast<unsigned>((cols + ai_tw - 1) / ai_tw), static_cast<unsigned>((rows + ai_th - 1) / ai_th));
    hipLaunchKernelGGL(accu_info_crossing_kernel, grid, dim3(256), 0, s, ldd, rows, cols, g, crossing.ptr);
    LUE_AMD_HIP(hipGetLastError());
    if (int const st = accu_impl<float>(ldd, rows, cols, crossing.ptr, 0.0f, crossing_accu.ptr, stream))
        return st;
    hipLaunchKernelGGL(accu_info_class_kernel, grid, dim3(256), 0, s, ldd, rows, cols, g, crossing_accu.ptr,
                       cell_class);
    LUE_AMD_HIP(hipGetLastError());
    return status_success;
}

int lue_amd_accu_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow, double inflow_scalar,
                     double* outflow, void* stream)
{
    return accu_impl<double>(ldd, rows, cols, inflow, inflow_scalar, outflow, stream);
}

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
    // ~Network waits for the device before it frees the buffers
    delete network;
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
                              static_cast<int64_t>(net.nrefs),
                              static_cast<int64_t>(net.tile_blocks),
                              static_cast<int64_t>(net.tile_critical_blocks),
                              static_cast<int64_t>(net.wide_min),
                              net.tile_queue ? 1 : 0,
                              static_cast<int64_t>(net.lane_max)};
    for (int i = 0; i < n && i < static_cast<int>(sizeof(values) / sizeof(values[0])); ++i)
        info[i] = values[i];
    return status_success;
}

}  // extern "C"

namespace {

    template<typename T>
    int network_accu(lue_amd_network network, T const* const* inflow, T inflow_scalar, T* const* outflow,
                     hipStream_t stream, Plan* plan)
    {
        if (!network || !outflow)
            return fail(status_invalid_argument, "null argument");
        Network const& net = network->net;
        Scalars<T> sc{};
        sc.in1_s = inflow_scalar;
        std::vector<CatchmentIO<T>> io;
        for (int c = 0; c < static_cast<int>(net.cats_host.size()); ++c)
        {
            if (!outflow[c])
                return fail(status_invalid_argument, "outflow must not be null");
            io.push_back(CatchmentIO<T>{entry(inflow, c), nullptr, nullptr, nullptr, nullptr, outflow[c], nullptr});
        }
        return plan ? plan_network<op_accu, T>(net, sc, io, *plan) : route_network<op_accu, T>(net, sc, io, stream);
    }

}  // namespace

extern "C" {

int lue_amd_network_accu_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                             float* const* outflow, void* stream)
{
    return network_accu<float>(network, inflow, inflow_scalar, outflow, static_cast<hipStream_t>(stream), nullptr);
}

int lue_amd_network_accu_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                             double* const* outflow, void* stream)
{
    return network_accu<double>(network, inflow, inflow_scalar, outflow, static_cast<hipStream_t>(stream), nullptr);
}

int lue_amd_plan_accu_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                          float* const* outflow, lue_amd_plan* plan)
{
    return make_plan(plan, [&](Plan& p) { return network_accu<float>(network, inflow, inflow_scalar, outflow, nullptr, &p); });
}

int lue_amd_plan_accu_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                          double* const* outflow, lue_amd_plan* plan)
{
    return make_plan(plan, [&](Plan& p) { return network_accu<double>(network, inflow, inflow_scalar, outflow, nullptr, &p); });
}

int lue_amd_network_set_wide_threshold(lue_amd_network network, int64_t positions)
{
    if (!network)
        return fail(status_invalid_argument, "null argument");
    if (positions < 0 || positions > INT32_MAX)
        return fail(status_invalid_argument, "wide threshold out of range");
    network->net.wide_min = static_cast<std::uint32_t>(positions);
    return status_success;
}

int lue_amd_network_set_lane_limit(lue_amd_network network, int64_t positions)
{
    if (!network)
        return fail(status_invalid_argument, "null argument");
    if (positions < 0 || positions > INT32_MAX)
        return fail(status_invalid_argument, "lane limit out of range");
    network->net.lane_max = static_cast<std::uint32_t>(positions);
    return status_success;
}

int lue_amd_network_set_tile_queue(lue_amd_network network, int on)
{
    if (!network)
        return fail(status_invalid_argument, "null argument");
    network->net.tile_queue = on != 0;
    return status_success;
}

int lue_amd_network_level_info(lue_amd_network network, int level, int64_t* info, int n)
{
    if (!network || !info || n < 0)
        return fail(status_invalid_argument, "null argument");
    Network const& net = network->net;
    if (level < 0 || level >= net.levels())
        return fail(status_invalid_argument, "no such chain level");
    std::size_t const l = static_cast<std::size_t>(level);
    int64_t const values[] = {static_cast<int64_t>(net.level_pos[l + 1] - net.level_pos[l]),
                              static_cast<int64_t>(net.level_items[l + 1] - net.level_items[l]),
                              l < net.level_longest.size() ? static_cast<int64_t>(net.level_longest[l]) : 0,
                              l + 1 < net.level_ref.size() ? static_cast<int64_t>(net.level_ref[l + 1] - net.level_ref[l])
                                                           : 0};
    for (int i = 0; i < n && i < static_cast<int>(sizeof(values) / sizeof(values[0])); ++i)
        info[i] = values[i];
    return status_success;
}

}  // extern "C"
