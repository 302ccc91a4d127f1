// C ABI of kinematic_wave (per call and on flow networks), double.
#include "api.hpp"

using namespace lue_amd;

extern "C" {

int lue_amd_kinematic_wave_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* current_outflow,
                               const double* inflow, const double* alpha, double alpha_scalar, const double* beta,
                               double beta_scalar, double time_step_duration, const double* channel_length,
                               double channel_length_scalar, double* outflow, void* stream)
{
    return kinematic_wave_impl<double>(ldd, rows, cols, current_outflow, inflow, alpha, alpha_scalar, beta,
                                     beta_scalar, time_step_duration, channel_length, channel_length_scalar,
                                     outflow, stream);
}

int lue_amd_network_kinematic_wave_f64(lue_amd_network network, const double* const* current_outflow,
                                       const double* const* inflow, const double* const* alpha, double alpha_scalar,
                                       const double* const* beta, double beta_scalar, double time_step_duration,
                                       const double* const* channel_length, double channel_length_scalar,
                                       double* const* outflow, void* stream)
{
    if (!network || !outflow || !current_outflow || !inflow)
        return fail(status_invalid_argument, "null argument");
    Network const& net = network->net;
    Scalars<double> sc{};
    sc.alpha_s = alpha_scalar;
    sc.beta_s = beta_scalar;
    sc.dx_s = channel_length_scalar;
    sc.dt = time_step_duration;
    std::vector<CatchmentIO<double>> io;
    for (int c = 0; c < static_cast<int>(net.cats_host.size()); ++c)
    {
        if (!current_outflow[c] || !inflow[c] || !outflow[c])
            return fail(status_invalid_argument, "null raster");
        io.push_back(CatchmentIO<double>{current_outflow[c], inflow[c], entry(alpha, c), entry(beta, c),
                                      entry(channel_length, c), outflow[c], nullptr});
    }
    return route_network<op_kinematic_wave, double>(net, sc, io, static_cast<hipStream_t>(stream));
}

}  // extern "C"
