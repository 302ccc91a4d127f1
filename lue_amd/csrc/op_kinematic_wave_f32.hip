// C ABI of kinematic_wave (per call and on flow networks), float, and the network timestep that allocates
// nothing (its per-call buffers live in a fixed device array).
#include "api.hpp"

#include <array>
#include <mutex>

using namespace lue_amd;

// Per-call buffers of lue_amd_network_kinematic_wave_f32_noalloc: one array per device, held until the
// process ends. Every call on a device uses the same array.
constexpr std::size_t noalloc_scratch_bytes = std::size_t{64} << 20;
static __device__ __align__(256) unsigned char scratch[noalloc_scratch_bytes];

namespace {

    // The array's device address on each device, looked up once per device (with that device current) by
    // the first call there. The array's name is only a host-side stand-in for it.
    constexpr int noalloc_max_devices = 64;
    std::mutex scratch_mutex;
    std::array<unsigned char*, noalloc_max_devices> scratch_address{};

    int scratch_on_device(int device, unsigned char*& ptr)
    {
        if (device < 0 || device >= noalloc_max_devices)
            return fail(status_invalid_argument, "device index beyond the fixed scratch table");
        std::lock_guard<std::mutex> lock(scratch_mutex);
        if (!scratch_address[device])
        {
            void* address = nullptr;
            hipError_t const e = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch));
            if (e != hipSuccess)
                return hip_fail(e, "hipGetSymbolAddress(scratch)");
            scratch_address[device] = static_cast<unsigned char*>(address);
        }
        ptr = scratch_address[device];
        return status_success;
    }

}  // namespace

namespace {

    template<typename T>
    int network_kinematic_wave(lue_amd_network network, T const* const* current_outflow, T const* const* inflow,
                               T const* const* alpha, T alpha_scalar, T const* const* beta, T beta_scalar,
                               T time_step_duration, T const* const* channel_length, T channel_length_scalar,
                               T* const* outflow, hipStream_t stream, Plan* plan)
    {
        if (!network || !outflow || !current_outflow || !inflow)
            return fail(status_invalid_argument, "null argument");
        Network const& net = network->net;
        Scalars<T> sc{};
        sc.alpha_s = alpha_scalar;
        sc.beta_s = beta_scalar;
        sc.dx_s = channel_length_scalar;
        sc.dt = time_step_duration;
        std::vector<CatchmentIO<T>> io;
        for (int c = 0; c < static_cast<int>(net.cats_host.size()); ++c)
        {
            if (!current_outflow[c] || !inflow[c] || !outflow[c])
                return fail(status_invalid_argument, "null raster");
            io.push_back(CatchmentIO<T>{current_outflow[c], inflow[c], entry(alpha, c), entry(beta, c),
                                        entry(channel_length, c), outflow[c], nullptr});
        }
        return plan ? plan_network<op_kinematic_wave, T>(net, sc, io, *plan)
                    : route_network<op_kinematic_wave, T>(net, sc, io, stream);
    }

}  // namespace

extern "C" {

int lue_amd_kinematic_wave_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* current_outflow,
                               const float* inflow, const float* alpha, float alpha_scalar, const float* beta,
                               float beta_scalar, float time_step_duration, const float* channel_length,
                               float channel_length_scalar, float* outflow, void* stream)
{
    return kinematic_wave_impl<float>(ldd, rows, cols, current_outflow, inflow, alpha, alpha_scalar, beta,
                                     beta_scalar, time_step_duration, channel_length, channel_length_scalar,
                                     outflow, stream);
}

int lue_amd_network_kinematic_wave_f32(lue_amd_network network, const float* const* current_outflow,
                                       const float* const* inflow, const float* const* alpha, float alpha_scalar,
                                       const float* const* beta, float beta_scalar, float time_step_duration,
                                       const float* const* channel_length, float channel_length_scalar,
                                       float* const* outflow, void* stream)
{
    return network_kinematic_wave<float>(network, current_outflow, inflow, alpha, alpha_scalar, beta, beta_scalar,
                                      time_step_duration, channel_length, channel_length_scalar, outflow,
                                      static_cast<hipStream_t>(stream), nullptr);
}

int lue_amd_plan_kinematic_wave_f32(lue_amd_network network, const float* const* current_outflow,
                                    const float* const* inflow, const float* const* alpha, float alpha_scalar,
                                    const float* const* beta, float beta_scalar, float time_step_duration,
                                    const float* const* channel_length, float channel_length_scalar,
                                    float* const* outflow, lue_amd_plan* plan)
{
    return make_plan(plan, [&](Plan& p) {
        return network_kinematic_wave<float>(network, current_outflow, inflow, alpha, alpha_scalar, beta, beta_scalar,
                                          time_step_duration, channel_length, channel_length_scalar, outflow, nullptr,
                                          &p);
    });
}

int lue_amd_network_kinematic_wave_f32_noalloc(lue_amd_network network, const float* const* current_outflow,
                                               const float* const* inflow, const float* const* alpha,
                                               float alpha_scalar, const float* const* beta, float beta_scalar,
                                               float time_step_duration, const float* const* channel_length,
                                               float channel_length_scalar, float* const* outflow, void* stream)
{
    if (!network || !outflow || !current_outflow || !inflow)
        return fail(status_invalid_argument, "null argument");
    Network const& net = network->net;
    if (!net.recorded)
        return fail(status_invalid_argument, "network has no recorded schedule");
    int const ncats = static_cast<int>(net.cats_host.size());
    if (ncats > io_inline_max)
        return fail(status_invalid_argument, "the allocation-free timestep takes networks of at most " +
                                                 std::to_string(io_inline_max) + " catchments");
    CatchmentIO<float> io[io_inline_max];
    for (int c = 0; c < ncats; ++c)
    {
        if (!current_outflow[c] || !inflow[c] || !outflow[c])
            return fail(status_invalid_argument, "null raster");
        io[c] = CatchmentIO<float>{current_outflow[c], inflow[c], entry(alpha, c), entry(beta, c),
                                   entry(channel_length, c), outflow[c], nullptr};
    }
    Scalars<float> sc{};
    sc.alpha_s = alpha_scalar;
    sc.beta_s = beta_scalar;
    sc.dx_s = channel_length_scalar;
    sc.dt = time_step_duration;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    return with_tile_shape<op_kinematic_wave, float>(net.tile_shape, [&]<int TW, int TH>() -> int
    {
        ReplayLayout const l = replay_layout<op_kinematic_wave, float, TW, TH>(net);
        if (l.total > noalloc_scratch_bytes)
            return fail(status_invalid_argument,
                        "network needs " + std::to_string(l.total) + " bytes of per-call buffers; the allocation-free "
                        "timestep has " + std::to_string(noalloc_scratch_bytes));
        int device = 0;
        LUE_AMD_HIP(hipGetDevice(&device));
        if (device != net.device)
            return fail(status_invalid_argument, "the network lives on another device");
        unsigned char* base = nullptr;
        if (int const st = scratch_on_device(device, base))
            return st;
        if (int const st = write_io<float>(io, ncats, reinterpret_cast<CatchmentIO<float>*>(base + l.io), s))
            return st;
        return replay_launch<op_kinematic_wave, float, TW, TH>(net, args_from(sc), io, base, l, s, false);
    });
}

}  // extern "C"
