// C ABI of kinematic_wave (per call and on flow networks), float.
#include "api.hpp"

using namespace lue_amd;

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
    if (!network || !outflow || !current_outflow || !inflow)
        return fail(status_invalid_argument, "null argument");
    Network const& net = network->net;
    Scalars<float> sc{};
    sc.alpha_s = alpha_scalar;
    sc.beta_s = beta_scalar;
    sc.dx_s = channel_length_scalar;
    sc.dt = time_step_duration;
    std::vector<CatchmentIO<float>> io;
    for (int c = 0; c < static_cast<int>(net.cats_host.size()); ++c)
    {
        if (!current_outflow[c] || !inflow[c] || !outflow[c])
            return fail(status_invalid_argument, "null raster");
        io.push_back(CatchmentIO<float>{current_outflow[c], inflow[c], entry(alpha, c), entry(beta, c),
                                      entry(channel_length, c), outflow[c], nullptr});
    }
    return route_network<op_kinematic_wave, float>(net, sc, io, static_cast<hipStream_t>(stream));
}

}  // extern "C"
