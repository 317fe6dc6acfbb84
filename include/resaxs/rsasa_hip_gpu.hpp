// rsasa_hip_gpu.hpp -- i_rsasa served by the MI355X engine's SASA kernels.
//
// rsasa_hip<FLT, OUT, BASE, DENSITY> replaces
//   rsasa_cl_gpu  (BASE = hip_base / cl_base, DENSITY = false): 32-dot dodecahedron +
//                 icosahedron spheres, rsasa_cl_gpu_alg.cpp:38-197;
//   rsasa_cpu_alg (BASE = algorithm, DENSITY = true): density-based ring spheres,
//                 rsasa_cpu_alg.cpp:33-193 (run on device 0).
// Per-type sphere sizes differ for ring spheres; the templates are padded to one
// length with dots at the atom centre, which their own atom always buries, so the
// exposed counts are those of the unpadded spheres.  Radii are the 16 atomic radii
// (ff_coef.hpp:161-184); rsasa[i] = exposed dots / dots of the atom's sphere.
#ifndef RESAXS_AMD_RSASA_HIP_GPU_HPP
#define RESAXS_AMD_RSASA_HIP_GPU_HPP

#include <algorithm>
#include <vector>

#include "constants.hpp"
#include "ff_coef.hpp"
#include "rsasa_algorithm.hpp"
#include "utils.hpp"

namespace resaxs {
namespace algorithm {

template <typename FLT_T, typename OUT_FLT_T, typename BASE, bool DENSITY>
class rsasa_hip : public i_rsasa<FLT_T, OUT_FLT_T, BASE> {
    typedef typename real_type<FLT_T>::real4 real4;
    typedef typename rsasa<FLT_T, OUT_FLT_T>::dot_spheres_params params_type;
    typedef typename rsasa<FLT_T, OUT_FLT_T>::dot_spheres_cl_params cl_params_type;

    rsx_engine* h_ = nullptr;
    std::vector<float> radii_;
    std::vector<float> dots_;  // [n_types][D][4]
    std::vector<int> n_dots_;  // real dots per type
    int D_ = 0;
    float probe_ = 0.0f;
    std::vector<float> buf_;
    std::vector<unsigned> counts_;

    void release()
    {
        if (h_) rsx_destroy(h_);
        h_ = nullptr;
    }

    // device of a hip_base object: from the dot_spheres_cl_params device list
    static int pick_device(hip_base<FLT_T>* self, const rsasa_params* p)
    {
        self->initialize(static_cast<const cl_params_type*>(p)->devices_);
        return self->device();
    }
    // plain algorithm base (the reference's CPU algorithm): device 0
    static int pick_device(algorithm*, const rsasa_params*) { return 0; }

    void run(const std::vector<real4>& atoms, std::vector<OUT_FLT_T>& out)
    {
        verify(out.size() >= atoms.size(), error::SAXS_INVALID_ARG);
        detail::to_float4(atoms, buf_);
        const int n = (int)atoms.size();
        counts_.resize(n > 0 ? n : 1);
        if (n > 0)
            detail::check(rsx_calc_rsasa(h_, buf_.data(), n, radii_.data(), (int)radii_.size(), probe_, dots_.data(),
                                         D_, counts_.data(), nullptr),
                          h_, "calc_rsasa");
        for (int i = 0; i < n; ++i)
            out[i] = static_cast<OUT_FLT_T>(counts_[i]) / static_cast<OUT_FLT_T>(n_dots_[get_factor_index(atoms[i])]);
        this->state = this->alg_computing;
    }

public:
    rsasa_hip() = default;
    rsasa_hip(const rsasa_hip&) = delete;
    rsasa_hip& operator=(const rsasa_hip&) = delete;
    ~rsasa_hip() override { release(); }

    void initialize(const rsasa_params* base_params) override
    {
        verify(base_params != nullptr, error::SAXS_INVALID_ARG);
        const params_type& p = *static_cast<const params_type*>(base_params);
        const int device = pick_device(this, base_params);
        std::vector<FLT_T> radii;
        atomic_form_factors<FLT_T>::calc_radii(radii);
        radii_.assign(radii.begin(), radii.end());
        probe_ = (float)p.probe_radius_;
        std::vector<std::vector<real4>> spheres(radii.size());
        for (size_t t = 0; t < radii.size(); ++t)
            spheres[t] = DENSITY ? p.create_dots(radii[t]) : p.create_dodeca_icosa_dots(radii[t]);
        D_ = 1;
        for (const auto& s : spheres) D_ = std::max(D_, (int)s.size());
        dots_.assign(radii.size() * (size_t)D_ * 4, 0.0f);
        n_dots_.assign(radii.size(), 0);
        for (size_t t = 0; t < radii.size(); ++t) {
            n_dots_[t] = (int)spheres[t].size();
            for (int d = 0; d < D_; ++d) {
                float* o = &dots_[((size_t)t * D_ + d) * 4];
                if (d < n_dots_[t])
                    for (int c = 0; c < 4; ++c) o[c] = (float)spheres[t][d].s[c];
                else
                    o[3] = radii_[t];  // centre dot: always buried by its own atom
            }
        }
        release();
        detail::check(rsx_create(device, &h_), nullptr, "rsx_create");
        this->state = this->alg_args_set;
    }

    void calc_rsasa(const std::vector<real4>& atoms, std::vector<OUT_FLT_T>& out) override
    {
        verify(h_ != nullptr && this->args_set(), DENSITY ? error::SAXS_ALG_NOT_INITIALIZED : error::SAXS_ARGS_NOT_SET);
        run(atoms, out);
    }

    void recalc_rsasa(const std::vector<real4>& atoms, std::vector<OUT_FLT_T>& out) override
    {
        verify(this->computing(), DENSITY ? error::SAXS_ARGS_NOT_SET : error::SAXS_NO_INITIAL_CALC);
        run(atoms, out);
    }

    // sphere sizes per type (real dots, before padding)
    const std::vector<int>& dots_per_type() const { return n_dots_; }
};

template <typename FLT_T, typename OUT_FLT_T>
struct rsasa_creator<FLT_T, OUT_FLT_T, hip_base<FLT_T>, rsasa_dot_spheres> {
    static i_rsasa<FLT_T, OUT_FLT_T, hip_base<FLT_T>>* create()
    {
        return new rsasa_hip<FLT_T, OUT_FLT_T, hip_base<FLT_T>, false>();
    }
};

template <typename FLT_T, typename OUT_FLT_T>
struct rsasa_creator<FLT_T, OUT_FLT_T, algorithm, rsasa_dot_spheres> {
    static i_rsasa<FLT_T, OUT_FLT_T, algorithm>* create() { return new rsasa_hip<FLT_T, OUT_FLT_T, algorithm, true>(); }
};

}  // namespace algorithm
}  // namespace resaxs

#endif
