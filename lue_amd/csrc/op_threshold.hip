// C ABI of accu_threshold (per call and on flow networks).
#include "api.hpp"

using namespace lue_amd;

extern "C" {

int lue_amd_accu_threshold_f32(const uint8_t* ldd, int64_t rows, int64_t cols, const float* inflow, float inflow_scalar,
                         const float* threshold, float threshold_scalar, float* outflow, float* remainder, void* stream)
{
    return split_impl<op_threshold, float>(ldd, rows, cols, inflow, inflow_scalar, threshold, threshold_scalar, outflow,
                                     remainder, stream);
}

int lue_amd_accu_threshold_f64(const uint8_t* ldd, int64_t rows, int64_t cols, const double* inflow,
                         double inflow_scalar, const double* threshold, double threshold_scalar, double* outflow,
                         double* remainder, void* stream)
{
    return split_impl<op_threshold, double>(ldd, rows, cols, inflow, inflow_scalar, threshold, threshold_scalar, outflow,
                                      remainder, stream);
}

int lue_amd_network_accu_threshold_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                                 const float* const* threshold, float threshold_scalar, float* const* outflow,
                                 float* const* remainder, void* stream)
{
    if (!network)
        return fail(status_invalid_argument, "null argument");
    return network_split<op_threshold, float>(network->net, inflow, inflow_scalar, threshold, threshold_scalar, outflow,
                                        remainder, static_cast<hipStream_t>(stream));
}

int lue_amd_network_accu_threshold_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                                 const double* const* threshold, double threshold_scalar, double* const* outflow,
                                 double* const* remainder, void* stream)
{
    if (!network)
        return fail(status_invalid_argument, "null argument");
    return network_split<op_threshold, double>(network->net, inflow, inflow_scalar, threshold, threshold_scalar, outflow,
                                         remainder, static_cast<hipStream_t>(stream));
}

int lue_amd_plan_accu_threshold_f32(lue_amd_network network, const float* const* inflow, float inflow_scalar,
                             const float* const* threshold, float threshold_scalar, float* const* outflow, float* const* remainder,
                             lue_amd_plan* plan)
{
    if (!network)
        return fail(status_invalid_argument, "null argument");
    return make_plan(plan, [&](Plan& p) {
        return network_split<op_threshold, float>(network->net, inflow, inflow_scalar, threshold, threshold_scalar, outflow,
                                          remainder, nullptr, &p);
    });
}

int lue_amd_plan_accu_threshold_f64(lue_amd_network network, const double* const* inflow, double inflow_scalar,
                             const double* const* threshold, double threshold_scalar, double* const* outflow, double* const* remainder,
                             lue_amd_plan* plan)
{
    if (!network)
        return fail(status_invalid_argument, "null argument");
    return make_plan(plan, [&](Plan& p) {
        return network_split<op_threshold, double>(network->net, inflow, inflow_scalar, threshold, threshold_scalar, outflow,
                                          remainder, nullptr, &p);
    });
}

}  // extern "C"
