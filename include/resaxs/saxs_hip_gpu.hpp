// saxs_hip_gpu.hpp -- i_saxs served by one MI355X engine (C ABI, rsx_capi.h).
//
// Replaces saxs_cl_gpu<FLT> (resaxslib/src/saxs_alg_cl_gpu_pt_wf.hpp:38-320,
// saxs_alg_cl_gpu_pt_wf.cpp:37-620) with its observable behaviour:
//   - access_params().initialize(q, bodies, devices) generates the default atomic
//     factor table at the bodies' mean radius and selects OH2 as the water row
//     (saxs_alg_cl_gpu_pt_wf.cpp:37-49); set_expansion_factor regenerates it;
//   - hydration uses the reference GPU set-up: the 16 atomic radii, a 1.8 A probe
//     and the 32-dot dodecahedron + icosahedron spheres (rsasa_cl_gpu_alg.cpp:59-84,
//     saxs_alg_cl_gpu_pt_wf.cpp:84-98); access_rsasa() holds the fractions after
//     each calculation;
//   - calc_curve re-reads the factor table every time; recalc_curve pushes changed
//     parameters first, and any change forces a full recompute
//     (saxs_alg_cl_gpu_pt_wf.cpp:468-504);
//   - errors are resaxs::error with the reference's codes and call-order checks.
// I(q) comes back in FLT_T; the engine keeps the running sum in FP64.
#ifndef RESAXS_AMD_SAXS_HIP_GPU_HPP
#define RESAXS_AMD_SAXS_HIP_GPU_HPP

#include <vector>

#include "ff_coef.hpp"
#include "saxs_algorithm.hpp"
#include "utils.hpp"

namespace resaxs {
namespace algorithm {

template <typename FLT_T>
class saxs_hip_gpu : public i_saxs<FLT_T, hip_base<FLT_T>> {
public:
    typedef typename real_type<FLT_T>::real4 real4;

protected:
    struct params_t;

    struct ff_t : form_factor_params<FLT_T> {
        params_t* owner = nullptr;
        const std::vector<FLT_T>* qq = nullptr;
        std::vector<FLT_T> radii;
        FLT_T avg_radius = FLT_T(1.62);

        void on_dirty_set() override { owner->mark_dirty(); }

        void set_expansion_factor(FLT_T c) override
        {
            if (this->expansion_factor_ != c) {
                form_factor_params<FLT_T>::set_expansion_factor(c);
                if (qq) atomic_form_factors<FLT_T>::generate(*qq, avg_radius, c, this->factors_);
            }
        }

        void setup(const std::vector<FLT_T>& q, const std::vector<real4>& bodies)
        {
            if (radii.empty()) atomic_form_factors<FLT_T>::calc_radii(radii);
            avg_radius = bodies.empty() ? FLT_T(1.62) : atomic_form_factors<FLT_T>::calc_avg_radius(bodies, radii);
            atomic_form_factors<FLT_T>::generate(q, avg_radius, this->expansion_factor_, this->factors_);
            qq = &q;
            this->set_water_ff_index(atoms::OH2);
        }
    };

    struct water_t : implicit_water_params<FLT_T> {
        params_t* owner = nullptr;
        void on_dirty_set() override { owner->mark_dirty(); }
    };

    struct ocl_t : ocl_exec_params {
        params_t* owner = nullptr;
        void on_dirty_set() override { owner->mark_dirty(); }
    };

    struct params_t : saxs_params<FLT_T> {
        ff_t ff;
        water_t water;
        ocl_t ocl;
        pivot_params pivot;

        params_t()
        {
            ff.owner = this;
            water.owner = this;
            ocl.owner = this;
        }
        params_t(const params_t&) = delete;
        params_t& operator=(const params_t&) = delete;

        void initialize(const std::vector<FLT_T>& qq, const std::vector<real4>& bodies,
                        const std::vector<dev_id>& devices, unsigned int workgroup_size = SAXS_NV_FERMI_WF_SIZE) override
        {
            saxs_params<FLT_T>::initialize(qq, bodies, devices, workgroup_size);
            ff.setup(this->qq_, bodies);
        }

        form_factor_params<FLT_T>& get_ff_params() override { return ff; }
        implicit_water_params<FLT_T>& get_implicit_water_params() override { return water; }
        ocl_exec_params& get_ocl_exec_params() override { return ocl; }
        pivot_params& get_pivot_params() override { return pivot; }

        void mark_dirty() { this->set_dirty(); }
        void clear_all()
        {
            this->clear_dirty();
            ff.clear_dirty();
            water.clear_dirty();
            ocl.clear_dirty();
        }
    };

    params_t params_;
    rsx_engine* h_ = nullptr;
    std::vector<float> buf_;
    int nq_ = 0;

    void release()
    {
        if (h_) rsx_destroy(h_);
        h_ = nullptr;
    }

    void push_factors()
    {
        const auto& f = params_.ff.factors_;
        verify(nq_ > 0 && !f.empty() && f.size() % nq_ == 0, error::SAXS_INVALID_ARG,
               "RESAXS::error -- factor table is not [n_q][n_types]");
        std::vector<float> t(f.begin(), f.end());
        detail::check(rsx_set_factors(h_, t.data(), (int)(t.size() / nq_), params_.ff.water_ff_index_), h_,
                      "set_factors");
    }

    void push_water()
    {
        detail::check(rsx_set_water_weight(h_, (float)params_.water.water_weight_), h_, "set_water_weight");
        detail::check(rsx_set_no_rsasa_recalc(h_, params_.water.no_rsasa_recalc_ ? 1 : 0), h_, "no_rsasa_recalc");
    }

    void pull_rsasa()
    {
        auto& r = params_.water.access_rsasa();
        if (params_.water.water_weight_ == 0) return;
        std::vector<float> s(params_.n_bodies_);
        detail::check(rsx_get_rsasa(h_, s.data(), (int)s.size()), h_, "get_rsasa");
        r.assign(s.begin(), s.end());
    }

    const float* body_ptr(const std::vector<real4>& bodies)
    {
        detail::to_float4(bodies, buf_);
        return buf_.data();
    }

public:
    saxs_hip_gpu() = default;
    saxs_hip_gpu(const saxs_hip_gpu&) = delete;
    saxs_hip_gpu& operator=(const saxs_hip_gpu&) = delete;
    ~saxs_hip_gpu() override { release(); }

    saxs_params<FLT_T>& access_params() override { return params_; }

    // device + engine set-up and argument binding (saxs_alg_cl_gpu_pt_wf.cpp:176-263)
    void initialize() override
    {
        hip_base<FLT_T>::initialize(params_.ocl.devices_);
        release();
        detail::check(rsx_create(this->device_, &h_), nullptr, "rsx_create");
        this->state = this->alg_program_loaded;
        nq_ = (int)params_.qq_.size();
        verify(nq_ > 0, error::SAXS_INVALID_ARG, "RESAXS::error -- empty q vector");
        std::vector<float> q(params_.qq_.begin(), params_.qq_.end());
        detail::check(rsx_set_q(h_, q.data(), nq_), h_, "set_q");
        float radii[atoms::ATOM_TYPE_COUNT];
        detail::check(rsx_atomic_radii(radii), h_, "atomic_radii");
        detail::check(rsx_set_sasa_params(h_, radii, atoms::ATOM_TYPE_COUNT, 1.8f, RSX_DOTS_DODECA_ICOSA, 0), h_,
                      "set_sasa_params");
        push_factors();
        push_water();
        detail::check(rsx_initialize(h_, (int)params_.n_bodies_), h_, "initialize");
        this->state = this->alg_args_set;
    }

    void calc_curve(const std::vector<real4>& bodies, std::vector<FLT_T>& out) override
    {
        verify(this->args_set(), error::SAXS_ARGS_NOT_SET);
        verify(bodies.size() == params_.n_bodies_, error::SAXS_INVALID_ARG,
               "RESAXS::error -- different number of bodies expected in calc_curve()");
        push_factors();
        push_water();
        const float* b = body_ptr(bodies);
        out.resize(nq_);
        if (sizeof(FLT_T) == sizeof(float))
            detail::check(rsx_calc_curve(h_, b, (int)bodies.size(), reinterpret_cast<float*>(out.data())), h_,
                          "calc_curve");
        else
            detail::check(rsx_calc_curve_f64(h_, b, (int)bodies.size(), reinterpret_cast<double*>(out.data())), h_,
                          "calc_curve");
        params_.clear_all();
        pull_rsasa();
        this->state = this->alg_computing;
    }

    void recalc_curve(const std::vector<real4>& bodies, unsigned int upd_start, unsigned int upd_length,
                      std::vector<FLT_T>& out) override
    {
        verify(this->computing(), error::SAXS_NO_INITIAL_CALC,
               "RESAXS::error -- recalc_curve() called before an initial call to calc_curve()");
        verify(bodies.size() >= (size_t)upd_start + upd_length, error::SAXS_INVALID_ARG,
               "RESAXS::error -- bodies vector too small in recalc_curve()");
        verify((size_t)upd_start + upd_length <= params_.n_bodies_, error::SAXS_INVALID_ARG,
               "RESAXS::error -- update region exceeds the number of bodies in recalc_curve()");
        if (params_.is_dirty()) {
            if (params_.ff.is_dirty()) push_factors();
            push_water();
            params_.clear_all();
        } else {
            detail::check(rsx_set_no_rsasa_recalc(h_, params_.water.no_rsasa_recalc_ ? 1 : 0), h_, "no_rsasa_recalc");
        }
        const float* b = body_ptr(bodies);
        out.resize(nq_);
        if (sizeof(FLT_T) == sizeof(float))
            detail::check(rsx_recalc_curve(h_, b, (int)bodies.size(), (int)upd_start, (int)upd_length,
                                           reinterpret_cast<float*>(out.data())),
                          h_, "recalc_curve");
        else
            detail::check(rsx_recalc_curve_f64(h_, b, (int)bodies.size(), (int)upd_start, (int)upd_length,
                                               reinterpret_cast<double*>(out.data())),
                          h_, "recalc_curve");
        pull_rsasa();
    }

    // transaction over the engine state (saxs_alg_cl_gpu_pt_wf.cpp:592-620)
    void commit() override
    {
        verify(this->computing(), error::SAXS_NO_INITIAL_CALC,
               "RESAXS::error -- commit() called before an initial call to calc_curve()");
        detail::check(rsx_commit(h_), h_, "commit");
    }

    void revert() override
    {
        verify(this->computing(), error::SAXS_NO_INITIAL_CALC,
               "RESAXS::error -- revert() called before an initial call to calc_curve()");
        detail::check(rsx_revert(h_), h_, "revert");
    }

    // ---- differentiable fit (no reference counterpart; the scale and chi of
    // saxs_profile.hpp:137-166) ----
    // chi^2 = sum_k (I_exp,k - c I_k)^2 / sigma_k^2 of the bodies' I(q) at the least-squares
    // scale c, and grad = d chi^2 / dx [n][3] (x, y, z per body) with c held fixed (it is
    // optimal).  The bodies become the engine's structure, as with calc_curve.  Returns chi^2.
    double chi2_grad(const std::vector<real4>& bodies, const std::vector<FLT_T>& i_exp,
                     const std::vector<FLT_T>& sigma, double& scale, std::vector<FLT_T>& grad)
    {
        verify(this->args_set(), error::SAXS_ARGS_NOT_SET);
        verify(bodies.size() == params_.n_bodies_ && i_exp.size() == (size_t)nq_ && sigma.size() == (size_t)nq_,
               error::SAXS_INVALID_ARG, "RESAXS::error -- chi2_grad: sizes do not match the engine");
        push_factors();
        push_water();
        std::vector<float> ie(i_exp.begin(), i_exp.end()), sg(sigma.begin(), sigma.end()), g(bodies.size() * 3);
        double chi2 = 0.0;
        detail::check(rsx_chi2_grad(h_, body_ptr(bodies), (int)bodies.size(), ie.data(), sg.data(), &chi2, &scale,
                                    g.data(), nullptr),
                      h_, "chi2_grad");
        params_.clear_all();
        this->state = this->alg_computing;
        grad.assign(g.begin(), g.end());
        return chi2;
    }

    // vector-Jacobian product grad_i = sum_k w_k dI_k / dx_i, grad [n][3] (the backward of I(q))
    void vjp(const std::vector<real4>& bodies, const std::vector<double>& w, std::vector<FLT_T>& grad)
    {
        verify(this->args_set(), error::SAXS_ARGS_NOT_SET);
        verify(bodies.size() == params_.n_bodies_ && w.size() == (size_t)nq_, error::SAXS_INVALID_ARG,
               "RESAXS::error -- vjp: sizes do not match the engine");
        push_factors();
        push_water();
        std::vector<float> g(bodies.size() * 3);
        detail::check(rsx_debye_vjp(h_, body_ptr(bodies), (int)bodies.size(), w.data(), g.data()), h_, "vjp");
        params_.clear_all();
        this->state = this->alg_computing;
        grad.assign(g.begin(), g.end());
    }

    // the engine handle, for callers that mix this object with the C ABI
    rsx_engine* engine() const { return h_; }
};

template <typename FLT_T>
struct saxs_creator<FLT_T, hip_base<FLT_T>, saxs_enum::saxs_gpu_pt_wf> {
    static i_saxs<FLT_T, hip_base<FLT_T>>* create() { return new saxs_hip_gpu<FLT_T>(); }
};

template <typename FLT_T>
struct saxs_creator<FLT_T, hip_base<FLT_T>, saxs_enum::saxs_gpu_v2> {
    static i_saxs<FLT_T, hip_base<FLT_T>>* create() { return new saxs_hip_gpu<FLT_T>(); }
};

// Distance-histogram engine (no counterpart in the reference; constructed directly,
// `new saxs_hip_hist<float>()`, not through the factory):
// I(q) from pair-distance histograms at bin_width (rsx_calc_curve_hist; approximate,
// the error falls with the square of the bin width).  Set-up as saxs_hip_gpu; every
// calc_curve / recalc_curve recomputes the histogram of the given bodies (no incremental
// state), commit and revert keep nothing to restore, and access_rsasa() is not updated.
template <typename FLT_T>
class saxs_hip_hist : public saxs_hip_gpu<FLT_T> {
    using base = saxs_hip_gpu<FLT_T>;
    using typename base::real4;

public:
    FLT_T bin_width = FLT_T(0.01);

    void calc_curve(const std::vector<real4>& bodies, std::vector<FLT_T>& out) override
    {
        verify(this->args_set(), error::SAXS_ARGS_NOT_SET);
        verify(bodies.size() == this->params_.n_bodies_, error::SAXS_INVALID_ARG,
               "RESAXS::error -- different number of bodies expected in calc_curve()");
        this->push_factors();
        this->push_water();
        this->params_.clear_all();
        run(bodies, out);
        this->state = this->alg_computing;
    }

    void recalc_curve(const std::vector<real4>& bodies, unsigned int upd_start, unsigned int upd_length,
                      std::vector<FLT_T>& out) override
    {
        verify(this->computing(), error::SAXS_NO_INITIAL_CALC,
               "RESAXS::error -- recalc_curve() called before an initial call to calc_curve()");
        verify(bodies.size() >= (size_t)upd_start + upd_length && bodies.size() == this->params_.n_bodies_,
               error::SAXS_INVALID_ARG, "RESAXS::error -- bodies vector too small in recalc_curve()");
        if (this->params_.is_dirty()) {
            if (this->params_.ff.is_dirty()) this->push_factors();
            this->push_water();
            this->params_.clear_all();
        }
        run(bodies, out);
    }

    void commit() override
    {
        verify(this->computing(), error::SAXS_NO_INITIAL_CALC,
               "RESAXS::error -- commit() called before an initial call to calc_curve()");
    }

    void revert() override
    {
        verify(this->computing(), error::SAXS_NO_INITIAL_CALC,
               "RESAXS::error -- revert() called before an initial call to calc_curve()");
    }

private:
    void run(const std::vector<real4>& bodies, std::vector<FLT_T>& out)
    {
        std::vector<double> iq((size_t)this->nq_);
        detail::check(rsx_calc_curve_hist(this->h_, this->body_ptr(bodies), (int)bodies.size(), (float)bin_width,
                                          iq.data()),
                      this->h_, "calc_curve_hist");
        out.assign(iq.begin(), iq.end());
    }
};

// The reference's saxs_hist_cl_gpu (alg_saxs_hist_cl_gpu.cpp, saxs_hist_gpu.cl:22-333)
// runs the pt_wf block kernels unchanged, so its slot gets the exact engine.
template <typename FLT_T>
struct saxs_creator<FLT_T, hip_base<FLT_T>, saxs_enum::saxs_hist_cl_gpu> {
    static i_saxs<FLT_T, hip_base<FLT_T>>* create() { return new saxs_hip_gpu<FLT_T>(); }
};

}  // namespace algorithm
}  // namespace resaxs

#endif
