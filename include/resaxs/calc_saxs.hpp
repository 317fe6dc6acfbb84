// calc_saxs.hpp -- the calc_saxs tool's ensemble fit (calc_saxs/calc_saxs.hpp:262-624,
// calc_saxs/calc_saxs.cpp:41-91) on the batched engine.
//
// The fit looks for the excluded-volume expansion factor c and the hydration-layer
// weight w at which the uniform mean of the models' I(q), scaled by its least-squares
// factor, is closest (chi of saxs_profile) to a reference profile.
//
// Evaluation.  calc_hip_saxs lays every model out once as a frame of one engine's batch
// (rsx_batch_set_frames): positions, exposed-dot counts and class-sorted columns stay in
// HBM.  A point (c, w) then costs one regenerated atomic factor table, one class-factor
// upload and one replay of the captured batch launch, which yields the ensemble mean
// directly (the tool's evaluator instead runs one engine per model call, re-uploading the
// model and recounting its SASA at every point, calc_saxs.hpp:223-243).
//
// Search (behaviour kept from calc_saxs.hpp:531-607 so that a fit lands on the same
// point).  Each parameter that is fitted spans its range in 5 steps in the first round;
// a fixed one is a single point.  A round visits the grid c-line by c-line and leaves a
// line as soon as chi rises along w.  The next round's window is +-one step around the
// best point so far, clipped to the range, again in 5 steps, or a single point once the
// step is at most 1e-4 (c) or 1e-3 (w).  The search stops when both steps are that small,
// or when a round moved the best chi by less than 1e-5 relative and 1e-4 absolute.
#ifndef RESAXS_AMD_CALC_SAXS_HPP
#define RESAXS_AMD_CALC_SAXS_HPP

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "../rsx_capi.h"
#include "ff_coef.hpp"
#include "parse_pdb.hpp"
#include "resaxs.hpp"
#include "saxs_algorithm.hpp"
#include "saxs_profile.hpp"
#include "utils.hpp"

namespace resaxs {

// A fitted parameter's range [lo, hi], or a fixed value (lo == hi: not fitted).  A reversed
// range is taken in order.  (Interface of calc_saxs.hpp:74-104.)
template <typename FLT_T>
struct profile_param {
public:
    profile_param() : lo_(std::numeric_limits<FLT_T>::min()), hi_(std::numeric_limits<FLT_T>::max()), fit_(true) {}
    profile_param(FLT_T value) : lo_(value), hi_(value), fit_(false) {}
    profile_param(FLT_T a, FLT_T b) : lo_(std::min(a, b)), hi_(std::max(a, b)), fit_(a != b) {}

    bool fit() const { return fit_; }
    FLT_T lo() const { return lo_; }
    FLT_T hi() const { return hi_; }
    void fix(FLT_T value) { *this = profile_param(value); }

private:
    FLT_T lo_, hi_;
    bool fit_;
};

// What the fit varies and what it fits to (interface of calc_saxs.hpp:107-126).  The
// defaults are what the tool's brace-initialised (value, false) members select, namely the
// (lo, hi) constructor: scale and expansion factor fitted over [0, 1], water weight fixed
// at 0 (calc_saxs.hpp:121-123).  The tool itself always passes all three.
template <typename FLT_T>
struct calc_params {
    calc_params() = default;
    calc_params(const profile_param<FLT_T>& scale, const profile_param<FLT_T>& water_weight,
                const profile_param<FLT_T>& exp_factor, saxs_profile<FLT_T>&& ref_profile)
        : scale_(scale), water_weight_(water_weight), exp_factor_(exp_factor), ref_profile_(std::move(ref_profile))
    {
    }
    calc_params(const profile_param<FLT_T>& scale, const profile_param<FLT_T>& water_weight,
                const profile_param<FLT_T>& exp_factor, const saxs_profile<FLT_T>& ref_profile)
        : scale_(scale), water_weight_(water_weight), exp_factor_(exp_factor), ref_profile_(ref_profile)
    {
    }

    profile_param<FLT_T> scale_{FLT_T(0), FLT_T(1)};
    profile_param<FLT_T> water_weight_{FLT_T(0)};
    profile_param<FLT_T> exp_factor_{FLT_T(0), FLT_T(1)};
    saxs_profile<FLT_T> ref_profile_;
};

// One evaluated point of the search and its scaled intensity; points order by chi
// (interface of calc_saxs.hpp:129-189: same_fit compares chi, == the parameters).
template <typename FLT_T>
struct fitted_params {
    FLT_T scale_ = 1;
    FLT_T water_weight_ = 0;
    FLT_T exp_factor_ = 1;
    FLT_T chi_ = std::numeric_limits<FLT_T>::max();
    std::vector<FLT_T> intensity_;

    bool operator==(const fitted_params& o) const
    {
        return exp_factor_ == o.exp_factor_ && water_weight_ == o.water_weight_ && scale_ == o.scale_;
    }
    bool same_fit(const fitted_params& o) const { return chi_ == o.chi_; }
    bool operator<(const fitted_params& o) const { return chi_ < o.chi_; }
    bool operator>(const fitted_params& o) const { return o < *this; }
    bool operator<=(const fitted_params& o) const { return !(o < *this); }
    bool operator>=(const fitted_params& o) const { return !(*this < o); }
};

template <typename T>
inline std::ostream& operator<<(std::ostream& os, const fitted_params<T>& p)
{
    return os << "# scale: " << p.scale_ << ", water weight: " << p.water_weight_
              << ", expansion factor: " << p.exp_factor_ << ", Chi: " << p.chi_ << std::endl;
}

// Ensemble evaluator on one batched engine (the tool's calc_cl_saxs, calc_saxs.hpp:192-260).
// The engine is set up as the tool's evaluator sets up its engine from the first model it
// sees: q of the reference profile, the default atomic factors at that model's mean
// atomic radius, water row OH2, SASA with the 32 dodeca-icosahedron dots and a 1.8 A probe
// (saxs_alg_cl_gpu_pt_wf.cpp:38-48, 86-97).  The ensemble is laid out on first use and
// again whenever a different set of models is passed.
template <typename FLT_T>
class calc_hip_saxs {
    using real4 = typename real_type<FLT_T>::real4;

public:
    calc_params<FLT_T> params_;   // the point evaluated next: water_weight_.lo(), exp_factor_.lo()
    unsigned int calc_count = 0;  // model evaluations (models x ensemble launches)

    calc_hip_saxs(algorithm::saxs_enum /*alg_pick*/, const std::string& dev_spec, unsigned int /*wf_size*/)
    {
        std::vector<dev_id> ids;
        parse_dev_spec(dev_spec, ids);
        algorithm::hip_base<FLT_T> dev;
        dev.initialize(ids);
        device_ = dev.device();
    }
    calc_hip_saxs(const calc_hip_saxs&) = delete;
    calc_hip_saxs& operator=(const calc_hip_saxs&) = delete;
    ~calc_hip_saxs()
    {
        if (h_) rsx_destroy(h_);
    }

    // uniform mean of the models' I(q) at the current parameters: one batched launch
    void ensemble_mean(const std::vector<std::vector<real4>>& models, std::vector<FLT_T>& mean)
    {
        bind(models);
        launch();
        calc_count += (unsigned int)models.size();
        mean.assign(mean_.begin(), mean_.end());
    }

    // I(q) of one model at the current parameters; a model outside the laid-out ensemble
    // becomes an ensemble of its own
    void calc(const std::vector<real4>& bodies, std::vector<FLT_T>& v_Iq)
    {
        int f = frame_of(bodies);
        if (f < 0) {
            bind(std::vector<std::vector<real4>>{bodies});
            f = 0;
        }
        launch();
        ++calc_count;
        v_Iq.assign(frames_.begin() + (size_t)f * nq_, frames_.begin() + (size_t)(f + 1) * nq_);
    }

    size_t n_engines() const { return h_ ? 1 : 0; }
    rsx_engine* engine() const { return h_; }

private:
    int device_ = 0;
    rsx_engine* h_ = nullptr;
    int nq_ = 0;
    std::vector<std::vector<real4>> models_;  // the laid-out ensemble (a copy)
    std::vector<const real4*> keys_;          // the callers' storage of each model
    std::vector<double> frames_, mean_;

    void check(int rc, const char* what) const { detail::check(rc, h_, what); }

    int frame_of(const std::vector<real4>& bodies) const
    {
        for (size_t f = 0; f < models_.size(); ++f)
            if (keys_[f] == bodies.data() && same(models_[f], bodies)) return (int)f;
        return -1;
    }

    static bool same(const std::vector<real4>& a, const std::vector<real4>& b)
    {
        return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(real4)) == 0;
    }

    void bind(const std::vector<std::vector<real4>>& models)
    {
        verify(params_.ref_profile_.initialized(), error::SAXS_ARGS_NOT_SET,
               "RESAXS::error -- the evaluator's reference profile is not initialised");
        verify(!models.empty(), error::SAXS_INVALID_ARG, "RESAXS::error -- no models to evaluate");
        bool fresh = h_ && models.size() == models_.size();
        for (size_t f = 0; fresh && f < models.size(); ++f)
            fresh = keys_[f] == models[f].data() && same(models_[f], models[f]);
        if (fresh) return;
        if (!h_) {
            check(rsx_create(device_, &h_), "rsx_create");
            const std::vector<FLT_T>& q = params_.ref_profile_.v_q_;
            std::vector<float> qf(q.begin(), q.end());
            nq_ = (int)qf.size();
            check(rsx_set_q(h_, qf.data(), nq_), "set_q");
            float radii[atoms::ATOM_TYPE_COUNT];
            check(rsx_atomic_radii(radii), "atomic_radii");
            check(rsx_set_sasa_params(h_, radii, atoms::ATOM_TYPE_COUNT, 1.8f, RSX_DOTS_DODECA_ICOSA, 0),
                  "set_sasa_params");
        }
        // factor table at the first model's mean atomic radius
        std::vector<FLT_T> radii;
        atomic_form_factors<FLT_T>::calc_radii(radii);
        check(rsx_set_expansion_factor(h_, (float)params_.exp_factor_.lo()), "set_expansion_factor");
        check(rsx_set_atomic_factors(h_, (float)atomic_form_factors<FLT_T>::calc_avg_radius(models[0], radii)),
              "set_atomic_factors");
        check(rsx_set_water_weight(h_, (float)params_.water_weight_.lo()), "set_water_weight");
        std::vector<float> flat;
        std::vector<int> offsets(1, 0);
        for (const auto& m : models) {
            for (const auto& b : m)
                for (int d = 0; d < 4; ++d) flat.push_back((float)b.s[d]);
            offsets.push_back(offsets.back() + (int)m.size());
        }
        check(rsx_batch_set_frames(h_, flat.data(), offsets.data(), (int)models.size()), "batch_set_frames");
        models_ = models;
        keys_.clear();
        for (const auto& m : models) keys_.push_back(m.data());
    }

    void launch()
    {
        check(rsx_set_expansion_factor(h_, (float)params_.exp_factor_.lo()), "set_expansion_factor");
        check(rsx_set_water_weight(h_, (float)params_.water_weight_.lo()), "set_water_weight");
        frames_.resize(models_.size() * (size_t)nq_);
        mean_.resize(nq_);
        check(rsx_batch_calc(h_, frames_.data(), mean_.data()), "batch_calc");
    }
};

// the tool's evaluator name
template <typename FLT_T>
using calc_cl_saxs = calc_hip_saxs<FLT_T>;

namespace detail {

// One axis of the grid search: the window [lo, hi] in n steps of `step` (n = 0: one point),
// inside the parameter's range [range_lo, range_hi].
template <typename T>
struct grid_axis {
    T range_lo, range_hi, lo, hi, step = 0;
    unsigned n;

    explicit grid_axis(const profile_param<T>& p)
        : range_lo(p.lo()), range_hi(p.hi()), lo(p.lo()), hi(p.hi()), n(p.fit() ? 5u : 0u)
    {
    }
    void begin_round() { step = (hi - lo) / std::max(n, 1u); }
    T at(unsigned i) const { return lo + i * step; }
    // the next round's window: one step either side of the best value; a single point once
    // the step is at most eps (the step then counts as 0)
    void narrow(T best, T eps)
    {
        const bool coarse = step > eps;
        n = coarse ? 5u : 0u;
        step = coarse ? step : T(0);
        lo = std::max(best - step, range_lo);
        hi = std::min(best + step, range_hi);
    }
};

}  // namespace detail

// The tool's driver (calc_saxs.hpp:262-304) over models given in memory or read from PDB
// files; q from the reference profile when it has one, else q_n points from q_min to q_max
// (calc_saxs.cpp:49-61).
template <typename FLT_T>
struct calc_saxs {
    using real4 = typename real_type<FLT_T>::real4;

    enum verbose_levels { QUIET = 0, NORMAL, DETAILS, DEBUG };

    calc_saxs(std::vector<std::vector<real4>> models, FLT_T q_min, FLT_T q_max, unsigned int q_n,
              calc_params<FLT_T> params, verbose_levels verbose_lvl = QUIET)
        : v_models_(std::move(models)), params_(std::move(params)), verbose_lvl_(verbose_lvl)
    {
        if (params_.ref_profile_) {
            v_q_ = params_.ref_profile_.v_q_;
        } else {
            v_q_.resize(q_n);
            const FLT_T dq = q_n > 1 ? (q_max - q_min) / (q_n - 1) : 0;
            for (unsigned int k = 0; k < q_n; ++k) v_q_[k] = q_min + k * dq;
        }
        intensity_.resize(v_q_.size());
        if (!v_models_.empty()) v_bodies_ = v_models_.front();
    }

    // atomic PDB models (calc_saxs.cpp:48-91, 279-303): every MODEL of every file is one
    // member of the ensemble.  The coarse two-body model (atomic = false) depends on a
    // third-party PDB library of the reference and is not provided.
    calc_saxs(const std::vector<std::string>& bodies_filenames, const std::string& /*exe_base_path*/, bool atomic,
              FLT_T q_min, FLT_T q_max, unsigned int q_n, calc_params<FLT_T> params, verbose_levels verbose_lvl)
        : calc_saxs(std::vector<std::vector<real4>>(), q_min, q_max, q_n, std::move(params), verbose_lvl)
    {
        verify(atomic, error::SAXS_INVALID_ARG, "RESAXS::error -- only atomic PDB models are supported");
        for (const auto& name : bodies_filenames) {
            parse_pdb<FLT_T> pdb(name);
            pdb.set_verbose_level(typename parse_pdb<FLT_T>::verbose_levels(verbose_lvl_));
            auto models = pdb.all_models_as_complex_atoms();
            for (auto& m : models) v_models_.push_back(std::move(m));
        }
        verify(!v_models_.empty(), error::SAXS_INVALID_ARG, "RESAXS::error -- no models in the PDB files");
        v_bodies_ = v_models_.front();
    }

    void set_verbose_level(verbose_levels lvl = NORMAL) { verbose_lvl_ = lvl; }

    // The ensemble at (c, w): mean I(q), its least-squares scale (or the fixed one) and chi.
    template <typename CALC_T>
    fitted_params<FLT_T> fit_ensemble(CALC_T& eval, FLT_T exp_factor, FLT_T water_weight)
    {
        eval.params_.scale_ = params_.scale_;
        eval.params_.exp_factor_.fix(exp_factor);
        eval.params_.water_weight_.fix(water_weight);
        fitted_params<FLT_T> pt;
        pt.exp_factor_ = exp_factor;
        pt.water_weight_ = water_weight;
        pt.intensity_ = avg_ensemble(eval);
        const saxs_profile<FLT_T>& ref = params_.ref_profile_;
        pt.scale_ = params_.scale_.fit() ? ref.optimize_scale_for(pt.intensity_) : params_.scale_.lo();
        for (auto& v : pt.intensity_) v *= pt.scale_;
        pt.chi_ = ref.chi_square_chi(pt.intensity_);
        return pt;
    }

    // The grid search over (c, w) described at the top of this file.
    template <typename CALC_T>
    fitted_params<FLT_T> fit_ensemble(CALC_T& eval)
    {
        if (verbose_lvl_ >= NORMAL)
            std::cout << "Calculating ensemble average for " << v_models_.size() << " conformations." << std::endl;
        eval.params_.ref_profile_.initialize(v_q_);
        const FLT_T c_eps = FLT_T(0.0001), w_eps = FLT_T(0.001);
        detail::grid_axis<FLT_T> c_ax(params_.exp_factor_), w_ax(params_.water_weight_);
        fitted_params<FLT_T> best;
        for (;;) {
            const FLT_T chi_before = best.chi_;
            c_ax.begin_round();
            w_ax.begin_round();
            for (unsigned i = 0; i <= c_ax.n; ++i) {
                FLT_T chi_prev = std::numeric_limits<FLT_T>::max();
                for (unsigned j = 0; j <= w_ax.n; ++j) {
                    fitted_params<FLT_T> pt = fit_ensemble(eval, c_ax.at(i), w_ax.at(j));
                    if (pt.chi_ > chi_prev) break;  // past the minimum along this line
                    chi_prev = pt.chi_;
                    if (pt < best) best = std::move(pt);
                }
            }
            if (verbose_lvl_ >= DETAILS && best.chi_ < chi_before) std::cout << "found min: " << best << std::endl;
            c_ax.narrow(best.exp_factor_, c_eps);
            w_ax.narrow(best.water_weight_, w_eps);
            const FLT_T moved = std::fabs(chi_before - best.chi_);
            const bool settled = moved / chi_before < FLT_T(0.00001) && moved < FLT_T(0.0001);
            if (settled || !(c_ax.step > c_eps || w_ax.step > w_eps)) break;
        }
        if (verbose_lvl_ >= DETAILS) std::cout << eval.calc_count << " calculations." << std::endl;
        return best;
    }

    // uniform mean of the models' I(q) at the evaluator's parameters (one batched launch)
    template <typename CALC_T>
    std::vector<FLT_T> avg_ensemble(CALC_T& eval) const
    {
        std::vector<FLT_T> mean;
        eval.ensemble_mean(v_models_, mean);
        return mean;
    }

    std::vector<std::vector<real4>> v_models_;
    std::vector<real4> v_bodies_;
    std::vector<FLT_T> v_q_;
    std::vector<FLT_T> intensity_;
    calc_params<FLT_T> params_;
    verbose_levels verbose_lvl_;
};

}  // namespace resaxs

#endif
