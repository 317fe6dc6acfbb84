// saxs_algorithm.hpp -- the Debye I(q) interface of the resaxslib surface
// (resaxslib/include/saxs_algorithm.hpp:29-244): parameter blocks, i_saxs, the
// algorithm enum, saxs_creator / saxs<FLT>::create and create_saxs_ocl.  The
// algorithms are served by saxs_hip_gpu (saxs_hip_gpu.hpp), one MI355X engine per
// object.
#ifndef RESAXS_AMD_SAXS_ALGORITHM_HPP
#define RESAXS_AMD_SAXS_ALGORITHM_HPP

#include <cmath>
#include <vector>

#include "re_algorithm.hpp"
#include "re_cl_algorithm_base.hpp"

namespace resaxs {
namespace algorithm {

// device list and workgroup size (saxs_algorithm.hpp:33-53); the engine chooses its
// own launch shapes, so the workgroup size is recorded only
class ocl_exec_params : public params_base {
public:
    std::vector<dev_id> devices_;
    unsigned int workgroup_size_ = 64;

    virtual void set_devices(const std::vector<dev_id>& devices)
    {
        devices_ = devices;
        set_dirty();
    }

    void set_workgroup_size(unsigned int size)
    {
        if (workgroup_size_ != size) {
            workgroup_size_ = size;
            set_dirty();
        }
    }
};

// hydration shell: F_i = f_type_i + s_i * w * f_water, s_i the exposed fraction
// (saxs_algorithm.hpp:55-77)
template <typename FLT_T>
class implicit_water_params : public params_base {
protected:
    std::vector<FLT_T> rsasa_;

public:
    FLT_T water_weight_ = 0;
    bool no_rsasa_recalc_ = false;

    void set_water_weight(FLT_T weight)
    {
        if (water_weight_ != weight) {
            water_weight_ = weight;
            set_dirty();
        }
    }

    std::vector<FLT_T>& access_rsasa() { return rsasa_; }
};

// factor table [n_q][n_types] row-major, water row index, expansion factor
// (saxs_algorithm.hpp:79-109)
template <typename FLT_T>
class form_factor_params : public params_base {
public:
    std::vector<FLT_T> factors_;
    int water_ff_index_ = -1;
    FLT_T expansion_factor_ = 1;

    void set_water_ff_index(int index)
    {
        if (water_ff_index_ != index) {
            water_ff_index_ = index;
            set_dirty();
        }
    }

    std::vector<FLT_T>& access_factors() { return factors_; }

    virtual void set_expansion_factor(FLT_T c)
    {
        if (expansion_factor_ != c) {
            expansion_factor_ = c;
            set_dirty();
        }
    }
};

// kept for source compatibility (saxs_algorithm.hpp:111-116); the engine's ΔI path
// does not need pivots
class pivot_params : public params_base {
public:
    int pivot_start_ = -1;
    int pivot_end_ = -1;
};

template <typename FLT_T>
struct saxs_params : public params_base {
    std::vector<FLT_T> qq_;
    unsigned int n_bodies_ = 0;

    virtual void initialize(const std::vector<FLT_T>& qq, const std::vector<real4<FLT_T>>& bodies,
                            const std::vector<dev_id>& devices, unsigned int workgroup_size = SAXS_NV_FERMI_WF_SIZE)
    {
        qq_ = qq;
        n_bodies_ = (unsigned int)bodies.size();
        get_ocl_exec_params().set_devices(devices);
        get_ocl_exec_params().set_workgroup_size(workgroup_size);
        set_dirty();
    }

    virtual form_factor_params<FLT_T>& get_ff_params() = 0;
    virtual implicit_water_params<FLT_T>& get_implicit_water_params() = 0;
    virtual ocl_exec_params& get_ocl_exec_params() = 0;
    virtual pivot_params& get_pivot_params() = 0;
};

// The SAXS algorithm interface (saxs_algorithm.hpp:143-183).  Calls are synchronous;
// one object per host thread.
template <typename FLT_T, typename BASE = algorithm>
struct i_saxs : public BASE {
    typedef typename real_type<FLT_T>::real4 real4;

    virtual saxs_params<FLT_T>& access_params() = 0;
    virtual void initialize() = 0;
    // full I(q) of `bodies` (size must equal the initialised body count)
    virtual void calc_curve(const std::vector<real4>& bodies, std::vector<FLT_T>& out_Iqq) = 0;
    // I(q) after bodies [upd_start, upd_start + upd_length) moved
    virtual void recalc_curve(const std::vector<real4>& bodies, unsigned int upd_start, unsigned int upd_length,
                              std::vector<FLT_T>& out_Iqq) = 0;
    virtual void commit() {}
    virtual void revert() {}
};

enum class saxs_enum {
    saxs_gpu_pt_wf,   // served by the MI355X engine
    saxs_gpu_v2,      // the same engine (the reference's older GPU decomposition of the same sum)
    saxs_hist_cl_gpu, // the same engine (the reference's class of that name runs the pt_wf kernels)
};

template <typename FLT_T, typename BASE, saxs_enum alg>
struct saxs_creator {
    static i_saxs<FLT_T, BASE>* create();
};

template <typename FLT_T>
class saxs {
public:
    template <typename BASE, saxs_enum alg>
    static i_saxs<FLT_T, BASE>* create()
    {
        return saxs_creator<FLT_T, BASE, alg>::create();
    }

    template <typename BASE>
    static i_saxs<FLT_T, BASE>* create(saxs_enum alg)
    {
        switch (alg) {
        case saxs_enum::saxs_gpu_pt_wf:
            return create<BASE, saxs_enum::saxs_gpu_pt_wf>();
        case saxs_enum::saxs_gpu_v2:
            return create<BASE, saxs_enum::saxs_gpu_v2>();
        case saxs_enum::saxs_hist_cl_gpu:
            return create<BASE, saxs_enum::saxs_hist_cl_gpu>();
        default:
            return nullptr;
        }
    }
};

template <typename FLT_T = float>
i_saxs<FLT_T, cl_base<FLT_T>>* create_saxs_ocl()
{
    return saxs<FLT_T>::template create<cl_base<FLT_T>, saxs_enum::saxs_gpu_pt_wf>();
}

}  // namespace algorithm
}  // namespace resaxs

#include "saxs_hip_gpu.hpp"

#endif
