// host_debye.hpp -- the one-call Debye I(q) of the resaxslib surface
// (resaxslib/include/host_debye.hpp:26-40, host_debye.cpp:54-116), same
// signatures and CalcSaxsCurveFn typedefs.
//
// The reference evaluates it on the CPU with a T accumulator; here each call runs
// the sum on HIP device 0 through a short-lived engine (FP64 running sums whatever
// T is) and rounds I(q) to FLT_T.  The q = 0 point is (sum F)^2 exactly; a pair at
// distance 0 contributes F_i F_j (sinc(0) = 1, the reference GPU rule; the
// reference CPU loop yields NaN there).
#ifndef RESAXS_AMD_HOST_DEBYE_HPP
#define RESAXS_AMD_HOST_DEBYE_HPP

#include <vector>

#include "resaxs.hpp"
#include "utils.hpp"

namespace resaxs {

typedef void (*CalcSaxsCurveFn)(const float* v_q, int n_q, const cl_float4* v_bodies, int n_bodies,
                                const float* t_factors, int n_factors, float* out_v_Iq);
typedef void (*CalcSaxsCurveFnDP)(const double* v_q, int n_q, const cl_double4* v_bodies, int n_bodies,
                                  const double* t_factors, int n_factors, double* out_v_Iq);

template <typename T, typename FLT_T, typename FLT4_T>
class host_debye {
public:
    typedef void (*CalcSaxsCurveFn)(const FLT_T* v_q, int n_q, const FLT4_T* v_bodies, int n_bodies,
                                    const FLT_T* t_factors, int n_factors, FLT_T* out_v_Iq);

    // t_factors [n_q][n_factors] row-major; out_v_Iq [n_q]
    static void calc_curve(const FLT_T* v_q, int n_q, const FLT4_T* v_bodies, int n_bodies, const FLT_T* t_factors,
                           int n_factors, FLT_T* out_v_Iq)
    {
        verify(n_q > 0 && n_bodies > 0 && n_factors > 0 && v_q && v_bodies && t_factors && out_v_Iq,
               error::SAXS_INVALID_ARG, "RESAXS::error -- host_debye: empty input");
        rsx_engine* h = nullptr;
        detail::check(rsx_create(0, &h), nullptr, "rsx_create");
        struct guard {
            rsx_engine* h;
            ~guard() { rsx_destroy(h); }
        } g{h};
        std::vector<float> q(v_q, v_q + n_q);
        std::vector<float> table(t_factors, t_factors + (size_t)n_q * n_factors);
        std::vector<float> b;
        detail::to_float4(std::vector<FLT4_T>(v_bodies, v_bodies + n_bodies), b);
        std::vector<double> iq(n_q);
        detail::check(rsx_set_q(h, q.data(), n_q), h, "set_q");
        detail::check(rsx_set_factors(h, table.data(), n_factors, -1), h, "set_factors");
        detail::check(rsx_initialize(h, n_bodies), h, "initialize");
        detail::check(rsx_calc_curve_f64(h, b.data(), n_bodies, iq.data()), h, "calc_curve");
        for (int k = 0; k < n_q; ++k) out_v_Iq[k] = (FLT_T)iq[k];
    }

    static void calc_curve(const std::vector<FLT_T>& v_q, const std::vector<FLT4_T>& v_bodies,
                           const std::vector<FLT_T>& t_factors, std::vector<FLT_T>& out_v_Iq)
    {
        calc_curve(v_q.data(), (int)v_q.size(), v_bodies.data(), (int)v_bodies.size(), t_factors.data(),
                   (int)(t_factors.size() / v_q.size()), out_v_Iq.data());
    }
};

}  // namespace resaxs

#endif
