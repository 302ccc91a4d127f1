// C ABI of kinematic_wave (per call and on flow networks), double.
#include "api.hpp"

using namespace lue_amd;

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
    return network_kinematic_wave<double>(network, current_outflow, inflow, alpha, alpha_scalar, beta, beta_scalar,
                                      time_step_duration, channel_length, channel_length_scalar, outflow,
                                      static_cast<hipStream_t>(stream), nullptr);
}

int lue_amd_plan_kinematic_wave_f64(lue_amd_network network, const double* const* current_outflow,
                                    const double* const* inflow, const double* const* alpha, double alpha_scalar,
                                    const double* const* beta, double beta_scalar, double time_step_duration,
                                    const double* const* channel_length, double channel_length_scalar,
                                    double* const* outflow, lue_amd_plan* plan)
{
    return make_plan(plan, [&](Plan& p) {
        return network_kinematic_wave<double>(network, current_outflow, inflow, alpha, alpha_scalar, beta, beta_scalar,
                                          time_step_duration, channel_length, channel_length_scalar, outflow, nullptr,
                                          &p);
    });
}

}  // extern "C"
