// Ensemble fit (the calc_saxs tool's search, calc_saxs/calc_saxs.hpp:420-624) behind the
// C ABI: the driver of include/resaxs/calc_saxs.hpp instantiated for float bodies.
#include <string>
#include <vector>

#include "resaxs/calc_saxs.hpp"

extern "C" int rsx_fit_ensemble(const float* bodies, const int* offsets, int n_models, const float* q,
                                const float* i_exp, const float* sigma, int n_q, const float* scale_range,
                                const float* water_range, const float* expansion_range, int device, double* out,
                                float* out_iq)
{
    using namespace resaxs;
    if (!bodies || !offsets || n_models <= 0 || !q || !i_exp || !sigma || n_q <= 0 || !scale_range || !water_range ||
        !expansion_range || !out)
        return RSX_INVALID_ARG;
    try {
        std::vector<std::vector<cl_float4>> models((size_t)n_models);
        for (int m = 0; m < n_models; ++m) {
            if (offsets[m + 1] < offsets[m]) return RSX_INVALID_ARG;
            for (int i = offsets[m]; i < offsets[m + 1]; ++i) {
                cl_float4 b;
                for (int d = 0; d < 4; ++d) b.s[d] = bodies[4 * (size_t)i + d];
                models[m].push_back(b);
            }
        }
        saxs_profile<float> profile(std::vector<float>(q, q + n_q), std::vector<float>(i_exp, i_exp + n_q),
                                    std::vector<float>(sigma, sigma + n_q));
        calc_params<float> params{profile_param<float>(scale_range[0], scale_range[1]),
                                  profile_param<float>(water_range[0], water_range[1]),
                                  profile_param<float>(expansion_range[0], expansion_range[1]), std::move(profile)};
        calc_saxs<float> calc(std::move(models), 0.0f, 0.0f, 0, std::move(params));
        calc_hip_saxs<float> eval(algorithm::saxs_enum::saxs_gpu_pt_wf, "0-" + std::to_string(device), 64);
        fitted_params<float> best = calc.fit_ensemble(eval);
        out[0] = best.exp_factor_;
        out[1] = best.water_weight_;
        out[2] = best.scale_;
        out[3] = best.chi_;
        out[4] = (double)eval.calc_count;
        if (out_iq)
            for (int k = 0; k < n_q; ++k) out_iq[k] = best.intensity_[k];
        return RSX_OK;
    } catch (const error& e) {
        return e.err() != 0 ? e.err() : RSX_INVALID_ARG;
    } catch (...) {
        return RSX_HIP_ERROR;
    }
}
