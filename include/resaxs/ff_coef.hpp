// ff_coef.hpp -- default atomic form factors of the resaxslib surface
// (resaxslib/include/ff_coef.hpp:104-211), computed by the library
// (rsx_atomic_form_factors / rsx_atomic_radii, bit-identical to the reference's
// float instantiation).  The double instantiation widens the FP32 table, the
// precision the device sums read it in.
#ifndef RESAXS_AMD_FF_COEF_HPP
#define RESAXS_AMD_FF_COEF_HPP

#include <cmath>
#include <vector>

#include "constants.hpp"
#include "resaxs.hpp"
#include "utils.hpp"

namespace resaxs {

template <typename FLT_T>
class atomic_form_factors {
public:
    // table [n_q][ATOM_TYPE_COUNT] of f_vacuum(q) - f_excluded_volume(q), expansion factor 1
    static void generate(const std::vector<FLT_T>& v_q, std::vector<FLT_T>& t_factors)
    {
        generate(v_q, FLT_T(1.62), FLT_T(1), t_factors);
    }

    // excluded volume scaled by c^3 exp(-pi (4 pi / 3)^(2/3) k^2 r_m^2 (c^2 - 1)), r_m = avg_radius
    static void generate(const std::vector<FLT_T>& v_q, FLT_T avg_radius, FLT_T c, std::vector<FLT_T>& t_factors)
    {
        std::vector<float> q(v_q.begin(), v_q.end());
        std::vector<float> t(q.size() * atoms::ATOM_TYPE_COUNT);
        if (!q.empty())
            detail::check(rsx_atomic_form_factors(q.data(), (int)q.size(), (float)avg_radius, (float)c, t.data()),
                          nullptr, "atomic_form_factors");
        t_factors.assign(t.begin(), t.end());
    }

    // radius of each type from its excluded volume at q = 0
    static void calc_radii(std::vector<FLT_T>& radii)
    {
        float r[atoms::ATOM_TYPE_COUNT];
        detail::check(rsx_atomic_radii(r), nullptr, "calc_radii");
        radii.assign(r, r + atoms::ATOM_TYPE_COUNT);
    }

    // mean radius over the bodies' types (running sum in FLT_T)
    static FLT_T calc_avg_radius(const std::vector<real4<FLT_T>>& bodies, const std::vector<FLT_T>& radii)
    {
        FLT_T sum = 0;
        for (const auto& b : bodies) sum += radii[get_factor_index(b)];
        return sum / bodies.size();
    }

    static FLT_T excl_vol_expansion_factor(FLT_T c, FLT_T k_sq, FLT_T rm)
    {
        const FLT_T c2 = c * c;
        const FLT_T k43 = (FLT_T)std::pow((double)(4 * (FLT_T)pi / 3), (double)(2.0f / 3.0f));
        return (FLT_T)((double)(c2 * c) * std::exp((double)(-(FLT_T)pi * k43 * k_sq * rm * rm * (c2 - 1))));
    }
};

}  // namespace resaxs

#endif
