// calc_saxs.hpp -- ensemble fitting driver of the calc_saxs tool
// (calc_saxs/calc_saxs.hpp:35-624, calc_saxs/calc_saxs.cpp:41-90), on the HIP engine.
//
// The fit searches the excluded-volume expansion factor c and the hydration-layer
// weight w for the ensemble average of per-model I(q), with the least-squares scale
// and chi of saxs_profile at every point (fit_ensemble(eval, c, w),
// calc_saxs.hpp:420-439).  The search is the tool's: a 6 x 6 grid over the ranges,
// lines abandoned once chi rises, then rounds of 6 x 6 grids of +-delta around the
// best point until both deltas fall below 1e-4 (c) and 1e-3 (w) or the best chi
// moves by less than 1e-5 relative (calc_saxs.hpp:531-607).
//
// The evaluator keeps one engine per model, each holding its model's positions,
// SASA counts and class layout in HBM, so a grid point costs one factor-table
// upload and one forward pass per model (the tool's evaluator re-uploads every model
// and recomputes its SASA at every point, calc_saxs.hpp:196-260).
#ifndef RESAXS_AMD_CALC_SAXS_HPP
#define RESAXS_AMD_CALC_SAXS_HPP

#include <algorithm>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <iostream>
#include <limits>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "parse_pdb.hpp"
#include "resaxs.hpp"
#include "saxs_algorithm.hpp"
#include "saxs_hip_gpu.hpp"
#include "saxs_profile.hpp"
#include "utils.hpp"

namespace resaxs {

// element-wise vector helpers the driver and its callers use (calc_saxs.hpp:38-67)
template <typename T, typename U, typename = typename std::enable_if<std::is_arithmetic<T>::value>::type,
          typename = typename std::enable_if<std::is_arithmetic<U>::value>::type>
inline std::vector<T>& operator*=(std::vector<T>& v, U c)
{
    for (auto& e : v) e *= c;
    return v;
}

template <typename T, typename U, typename = typename std::enable_if<std::is_arithmetic<T>::value>::type,
          typename = typename std::enable_if<std::is_arithmetic<U>::value>::type>
inline std::vector<T>& operator/=(std::vector<T>& v, U c)
{
    for (auto& e : v) e /= c;
    return v;
}

template <typename T, typename = typename std::enable_if<std::is_arithmetic<T>::value>::type>
inline std::vector<T>& operator+=(std::vector<T>& v1, const std::vector<T>& v2)
{
    for (size_t i = 0; i < v1.size(); ++i) v1[i] += v2[i];
    return v1;
}

template <typename T>
inline bool contains(std::initializer_list<T> il, const T& value)
{
    return std::find(std::begin(il), std::end(il), value) != std::end(il);
}

template <typename T, typename = typename std::enable_if<std::is_arithmetic<T>::value>::type>
constexpr T clamp_to_zero(T value, T lo)
{
    return value > lo ? value : 0;
}

// a fitted or fixed parameter: [lo, hi] with fit() when the range is not a point
// (calc_saxs.hpp:74-104; the (lo, hi) constructor swaps a reversed range)
template <typename FLT_T>
struct profile_param {
public:
    profile_param() : fit_(true), lo_(std::numeric_limits<FLT_T>::min()), hi_(std::numeric_limits<FLT_T>::max()) {}
    profile_param(FLT_T lo, FLT_T hi) : fit_(true), lo_(lo), hi_(hi)
    {
        if (lo_ == hi_) fit_ = false;
        if (lo_ > hi_) std::swap(lo_, hi_);
    }
    profile_param(FLT_T value) : fit_(false), lo_(value), hi_(value) {}

    bool fit() const { return fit_; }
    FLT_T lo() const { return lo_; }
    FLT_T hi() const { return hi_; }
    void fix(FLT_T value)
    {
        lo_ = hi_ = value;
        fit_ = false;
    }

private:
    bool fit_;
    FLT_T lo_;
    FLT_T hi_;
};

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
        : calc_params(scale, water_weight, exp_factor, saxs_profile<FLT_T>(ref_profile))
    {
    }

    profile_param<FLT_T> scale_ = {1, false};         // (the reference's defaults, kept as written)
    profile_param<FLT_T> water_weight_ = {0, false};
    profile_param<FLT_T> exp_factor_ = {1, false};
    saxs_profile<FLT_T> ref_profile_;
};

template <typename FLT_T>
struct fitted_params {
    FLT_T scale_ = 1;
    FLT_T water_weight_ = 0;
    FLT_T exp_factor_ = 1;
    FLT_T chi_ = std::numeric_limits<FLT_T>::max();
    std::vector<FLT_T> intensity_;

    bool operator==(const fitted_params& o) const
    {
        return scale_ == o.scale_ && water_weight_ == o.water_weight_ && exp_factor_ == o.exp_factor_;
    }
    bool same_fit(const fitted_params& o) const { return chi_ == o.chi_; }
    bool operator<(const fitted_params& o) const { return chi_ < o.chi_; }
    bool operator<=(const fitted_params& o) const { return chi_ <= o.chi_; }
    bool operator>(const fitted_params& o) const { return !operator<=(o); }
    bool operator>=(const fitted_params& o) const { return !operator<(o); }
};

template <typename T>
inline std::ostream& operator<<(std::ostream& os, const fitted_params<T>& p)
{
    os << "# scale: " << p.scale_ << ", water weight: " << p.water_weight_ << ", expansion factor: " << p.exp_factor_
       << ", Chi: " << p.chi_ << std::endl;
    return os;
}

// Evaluator: I(q) of one model at the current (c, w).  One engine per distinct model
// (keyed by the bodies' storage, checked against a copy), each set up as the tool's
// evaluator sets up its single engine from the first model it sees: default atomic
// factors at that model's mean radius, water row OH2, GPU SASA hydration
// (calc_saxs.hpp:196-260; the tool recalculates every later model on that engine, so
// all models share the first one's factor table and atom count).
template <typename FLT_T>
class calc_hip_saxs {
    using real4 = typename real_type<FLT_T>::real4;
    using engine_t = algorithm::saxs_hip_gpu<FLT_T>;

public:
    calc_params<FLT_T> params_;   // current parameters (water weight and expansion factor are read per call)
    unsigned int calc_count = 0;  // number of model evaluations

    calc_hip_saxs(algorithm::saxs_enum /*alg_pick*/, const std::string& dev_spec, unsigned int wf_size)
        : wf_size_(wf_size)
    {
        parse_dev_spec(dev_spec, devices_);
    }

    // I(q) of the model at the current parameters
    void calc(const std::vector<real4>& bodies, std::vector<FLT_T>& v_Iq)
    {
        ++calc_count;
        model_t& m = model_for(bodies);
        auto& sp = m.engine->access_params();
        sp.get_implicit_water_params().set_water_weight(params_.water_weight_.lo());
        sp.get_ff_params().set_expansion_factor(params_.exp_factor_.lo());
        if (m.engine->computing())
            m.engine->recalc_curve(bodies, 0, 0, v_Iq);
        else
            m.engine->calc_curve(bodies, v_Iq);
    }

    size_t n_engines() const { return models_.size(); }

private:
    struct model_t {
        const real4* key = nullptr;
        std::vector<real4> copy;
        std::unique_ptr<engine_t> engine;
    };

    std::vector<dev_id> devices_;
    unsigned int wf_size_;
    std::vector<model_t> models_;
    std::vector<real4> first_;  // the model the factor table's mean radius comes from

    model_t& model_for(const std::vector<real4>& bodies)
    {
        verify(params_.ref_profile_.initialized(), error::SAXS_ARGS_NOT_SET,
               "RESAXS::error -- the evaluator's reference profile is not initialised");
        for (auto& m : models_)
            if (m.key == bodies.data() && m.copy.size() == bodies.size() &&
                std::memcmp(m.copy.data(), bodies.data(), bodies.size() * sizeof(real4)) == 0)
                return m;
        // a new model (or storage reused with other positions): its own engine
        for (auto it = models_.begin(); it != models_.end(); ++it)
            if (it->key == bodies.data()) {
                models_.erase(it);
                break;
            }
        if (first_.empty()) first_ = bodies;
        verify(bodies.size() == first_.size(), error::SAXS_INVALID_ARG,
               "RESAXS::error -- different number of bodies expected in calc_curve()");
        model_t m;
        m.key = bodies.data();
        m.copy = bodies;
        m.engine.reset(new engine_t());
        auto& sp = m.engine->access_params();
        sp.initialize(params_.ref_profile_.v_q_, first_, devices_, wf_size_);
        sp.get_implicit_water_params().set_water_weight(params_.water_weight_.lo());
        sp.get_ff_params().set_expansion_factor(params_.exp_factor_.lo());
        m.engine->initialize();
        models_.push_back(std::move(m));
        return models_.back();
    }
};

// the tool's evaluator name
template <typename FLT_T>
using calc_cl_saxs = calc_hip_saxs<FLT_T>;

// The fitting part of the tool's driver (calc_saxs.hpp:262-624) over models given in
// memory; q from the reference profile when it has one, else q_n points from q_min
// to q_max (calc_saxs.cpp:49-61).
template <typename FLT_T>
struct calc_saxs {
    using real4 = typename real_type<FLT_T>::real4;

    enum verbose_levels { QUIET = 0, NORMAL, DETAILS, DEBUG };

    calc_saxs(std::vector<std::vector<real4>> models, FLT_T q_min, FLT_T q_max, unsigned int q_n,
              calc_params<FLT_T> params, verbose_levels verbose_lvl = QUIET)
        : v_models_(std::move(models)), params_(std::move(params)), verbose_lvl_(verbose_lvl)
    {
        if (params_.ref_profile_)
            v_q_ = params_.ref_profile_.v_q_;
        else {
            v_q_.resize(q_n);
            const FLT_T step = (q_n > 0) ? (q_max - q_min) / (q_n - 1) : 0;
            for (unsigned int i = 0; i < q_n; ++i) v_q_[i] = q_min + i * step;
        }
        intensity_.resize(v_q_.size());
        if (!v_models_.empty()) v_bodies_ = v_models_[0];
    }

    // the tool's constructor for atomic models (calc_saxs.cpp:48-91, 279-303): every MODEL of
    // every PDB file is one ensemble member; bodies are the first model.  The coarse
    // two-body model (atomic = false) reads its tables and residues with the reference's
    // third-party PDB library and is not provided here.
    calc_saxs(const std::vector<std::string>& bodies_filenames, const std::string& /*exe_base_path*/, bool atomic,
              FLT_T q_min, FLT_T q_max, unsigned int q_n, calc_params<FLT_T> params, verbose_levels verbose_lvl)
        : calc_saxs(std::vector<std::vector<real4>>(), q_min, q_max, q_n, std::move(params), verbose_lvl)
    {
        verify(atomic, error::SAXS_INVALID_ARG, "RESAXS::error -- only atomic PDB models are supported");
        for (const auto& f : bodies_filenames) {
            parse_pdb<FLT_T> parser(f);
            parser.set_verbose_level(typename parse_pdb<FLT_T>::verbose_levels(verbose_lvl_));
            for (auto& m : parser.all_models_as_complex_atoms()) v_models_.push_back(std::move(m));
        }
        verify(!v_models_.empty(), error::SAXS_INVALID_ARG, "RESAXS::error -- no models in the PDB files");
        v_bodies_ = v_models_[0];
    }

    void set_verbose_level(verbose_levels lvl = NORMAL) { verbose_lvl_ = lvl; }

    // the ensemble at (c, w): average I(q), least-squares scale (or the fixed one), chi
    template <typename CALC_T>
    fitted_params<FLT_T> fit_ensemble(CALC_T& eval, FLT_T exp_factor, FLT_T water_weight)
    {
        eval.params_.scale_ = params_.scale_;
        eval.params_.exp_factor_.fix(exp_factor);
        eval.params_.water_weight_.fix(water_weight);

        fitted_params<FLT_T> fit;
        fit.exp_factor_ = exp_factor;
        fit.water_weight_ = water_weight;
        fit.intensity_ = avg_ensemble(eval);
        fit.scale_ = params_.scale_.fit() ? params_.ref_profile_.optimize_scale_for(fit.intensity_) : params_.scale_.lo();
        fit.intensity_ *= fit.scale_;
        fit.chi_ = params_.ref_profile_.chi_square_chi(fit.intensity_);
        return fit;
    }

    // grid search of the (assumed convex) (c, w) space
    template <typename CALC_T>
    fitted_params<FLT_T> fit_ensemble(CALC_T& eval)
    {
        if (verbose_lvl_ >= NORMAL)
            std::cout << "Calculating ensemble average for " << v_models_.size() << " conformations." << std::endl;
        eval.params_.ref_profile_.initialize(v_q_);

        const FLT_T ef_epsilon = FLT_T(0.0001);
        const FLT_T ww_epsilon = FLT_T(0.001);
        const FLT_T chi_threshold = FLT_T(0.00001);
        const unsigned int default_grid_slices = 5, first_round_grid_slices = 5, fixed_value_slices = 0;

        unsigned int ef_slices = params_.exp_factor_.fit() ? first_round_grid_slices : fixed_value_slices;
        FLT_T ef_min = params_.exp_factor_.lo(), ef_max = params_.exp_factor_.hi(), ef_delta;
        unsigned int ww_slices = params_.water_weight_.fit() ? first_round_grid_slices : fixed_value_slices;
        FLT_T ww_min = params_.water_weight_.lo(), ww_max = params_.water_weight_.hi(), ww_delta;

        fitted_params<FLT_T> best;
        FLT_T prev_round_chi = best.chi_;
        bool chi_threshold_reached = false;
        do {
            prev_round_chi = best.chi_;
            ef_delta = (ef_max - ef_min) / std::max(ef_slices, 1u);
            ww_delta = (ww_max - ww_min) / std::max(ww_slices, 1u);
            for (unsigned int i = 0; i <= ef_slices; ++i) {
                FLT_T prev_pt_chi = std::numeric_limits<FLT_T>::max();
                for (unsigned int j = 0; j <= ww_slices; ++j) {
                    fitted_params<FLT_T> p = fit_ensemble(eval, ef_min + i * ef_delta, ww_min + j * ww_delta);
                    if (prev_pt_chi < p.chi_) break;  // moving away from the minimum along this line
                    prev_pt_chi = p.chi_;
                    if (p < best) best = std::move(p);
                }
            }
            if (verbose_lvl_ >= DETAILS && best.chi_ < prev_round_chi) std::cout << "found min: " << best << std::endl;
            ef_slices = ef_delta > ef_epsilon ? default_grid_slices : fixed_value_slices;
            ww_slices = ww_delta > ww_epsilon ? default_grid_slices : fixed_value_slices;
            ef_delta = clamp_to_zero(ef_delta, ef_epsilon);
            ww_delta = clamp_to_zero(ww_delta, ww_epsilon);
            ef_min = std::max(best.exp_factor_ - ef_delta, params_.exp_factor_.lo());
            ef_max = std::min(best.exp_factor_ + ef_delta, params_.exp_factor_.hi());
            ww_min = std::max(best.water_weight_ - ww_delta, params_.water_weight_.lo());
            ww_max = std::min(best.water_weight_ + ww_delta, params_.water_weight_.hi());
            const FLT_T chi_delta = std::fabs(prev_round_chi - best.chi_);
            chi_threshold_reached = chi_delta / prev_round_chi < chi_threshold && chi_delta < FLT_T(0.0001);
        } while ((ef_delta > ef_epsilon || ww_delta > ww_epsilon) && !chi_threshold_reached);

        if (verbose_lvl_ >= DETAILS) std::cout << eval.calc_count << " calculations." << std::endl;
        return best;
    }

    // uniform-weight average of the models' I(q) at the evaluator's parameters
    template <typename CALC_T>
    std::vector<FLT_T> avg_ensemble(CALC_T& eval) const
    {
        std::vector<FLT_T> avg(v_q_.size()), iq(v_q_.size());
        for (const auto& bodies : v_models_) {
            eval.calc(bodies, iq);
            avg += iq;
        }
        avg /= v_models_.size();
        return avg;
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
