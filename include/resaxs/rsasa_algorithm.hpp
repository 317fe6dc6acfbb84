// rsasa_algorithm.hpp -- the SASA interface of the resaxslib surface
// (resaxslib/include/rsasa_algorithm.hpp:29-146): i_rsasa, the dot-sphere
// parameters with their generators, rsasa_creator and rsasa<FLT, OUT>::create.
// Both bases are served by the MI355X engine (rsasa_hip_gpu.hpp):
//   create<cl_base<FLT>>  -> the reference GPU algorithm's set-up: 32-dot
//                            dodecahedron + icosahedron spheres (rsasa_cl_gpu_alg.cpp:59-84);
//   create<algorithm>     -> the reference CPU algorithm's set-up: density-based
//                            ring spheres from create_dots (rsasa_cpu_alg.cpp:38-59),
//                            computed on the GPU (there is no CPU path).
#ifndef RESAXS_AMD_RSASA_ALGORITHM_HPP
#define RESAXS_AMD_RSASA_ALGORITHM_HPP

#include <cmath>
#include <vector>

#include "re_algorithm.hpp"
#include "re_cl_algorithm_base.hpp"

namespace resaxs {
namespace algorithm {

struct rsasa_params {
};

template <typename FLT_T, typename OUT_FLT_T = FLT_T, typename BASE = algorithm>
class i_rsasa : public BASE {
    typedef typename real_type<FLT_T>::real4 real4;

public:
    virtual void initialize(const rsasa_params*) = 0;
    // exposed fraction per atom; rsasa.size() >= atoms.size() is required
    virtual void calc_rsasa(const std::vector<real4>& atoms, std::vector<OUT_FLT_T>& rsasa) = 0;
    virtual void recalc_rsasa(const std::vector<real4>& atoms, std::vector<OUT_FLT_T>& rsasa) = 0;
};

enum rsasa_enum {
    rsasa_dot_spheres,
};

template <typename FLT_T, typename OUT_FLT_T, typename BASE, unsigned int alg>
struct rsasa_creator {
    static i_rsasa<FLT_T, OUT_FLT_T, BASE>* create();
};

template <typename FLT_T, typename OUT_FLT_T = FLT_T>
class rsasa {
    typedef typename real_type<FLT_T>::real4 real4;

    static std::vector<real4> unpack(const std::vector<float>& f, int n)
    {
        std::vector<real4> out(n);
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 4; ++c) out[i].s[c] = (FLT_T)f[4 * i + c];
        return out;
    }

    static std::vector<real4> ring_dots(FLT_T radius, float n_sect, FLT_T probe)
    {
        int n = 0;
        detail::check(rsx_sphere_dots((float)radius, n_sect, (float)probe, nullptr, 0, &n), nullptr, "sphere_dots");
        std::vector<float> f(4 * (size_t)(n > 0 ? n : 1));
        detail::check(rsx_sphere_dots((float)radius, n_sect, (float)probe, f.data(), n, &n), nullptr, "sphere_dots");
        return unpack(f, n);
    }

public:
    struct dot_spheres_params : public rsasa_params {
        FLT_T sqrt_dot_density_;
        FLT_T probe_radius_;

        dot_spheres_params() : sqrt_dot_density_(0), probe_radius_(0) {}
        dot_spheres_params(FLT_T dot_density, FLT_T probe_radius)
            : sqrt_dot_density_(std::sqrt(dot_density)), probe_radius_(probe_radius)
        {
        }
        virtual ~dot_spheres_params() {}

        bool operator==(const dot_spheres_params& o) const
        {
            return sqrt_dot_density_ == o.sqrt_dot_density_ && probe_radius_ == o.probe_radius_;
        }
        bool operator!=(const dot_spheres_params& o) const { return !(*this == o); }

        // rings of dots at the density sqrt_dot_density_^2 per A^2 (rsasa_algorithm.cpp:82-87)
        const std::vector<real4> create_dots(FLT_T radius) const
        {
            const float n_sect = 2 * (float)pi * (float)radius * (float)sqrt_dot_density_;
            return ring_dots(radius, n_sect, probe_radius_);
        }
        // rsasa_algorithm.cpp:89-93
        const std::vector<real4> create_32_dots(FLT_T radius) const { return ring_dots(radius, 10.5f, probe_radius_); }
        // 64-dot Fibonacci sphere (rsasa_algorithm.cpp:95-116)
        const std::vector<real4> create_fibonacci_dots(FLT_T radius) const
        {
            std::vector<float> f(4 * 64);
            detail::check(rsx_fibonacci_dots(32, (float)radius, (float)probe_radius_, f.data()), nullptr,
                          "fibonacci_dots");
            return unpack(f, 64);
        }
        // 20 dodecahedron vertices (rsasa_algorithm.cpp:118-155): the first 20 dodeca-icosa dots
        const std::vector<real4> create_dodecahedron_dots(FLT_T radius) const
        {
            std::vector<real4> d = create_dodeca_icosa_dots(radius);
            d.resize(20);
            return d;
        }
        // 20 dodecahedron + 12 icosahedron vertices (rsasa_algorithm.cpp:157-208)
        const std::vector<real4> create_dodeca_icosa_dots(FLT_T radius) const
        {
            std::vector<float> f(4 * 32);
            detail::check(rsx_dodeca_icosa_dots((float)radius, (float)probe_radius_, f.data()), nullptr,
                          "dodeca_icosa_dots");
            return unpack(f, 32);
        }
    };

    struct dot_spheres_cl_params : public dot_spheres_params {
        std::vector<dev_id> devices_;
        unsigned int n_atoms_ = 0;

        dot_spheres_cl_params() : dot_spheres_params() {}
        dot_spheres_cl_params(FLT_T dot_density, FLT_T probe_radius, const std::vector<dev_id>& devices,
                              unsigned int n_atoms)
            : dot_spheres_params(dot_density, probe_radius), devices_(devices), n_atoms_(n_atoms)
        {
        }
    };

    template <typename BASE, unsigned int alg>
    static i_rsasa<FLT_T, OUT_FLT_T, BASE>* create()
    {
        return rsasa_creator<FLT_T, OUT_FLT_T, BASE, alg>::create();
    }

    template <typename BASE>
    static i_rsasa<FLT_T, OUT_FLT_T, BASE>* create(rsasa_enum alg)
    {
        switch (alg) {
        case rsasa_dot_spheres:
            return create<BASE, rsasa_dot_spheres>();
        default:
            return nullptr;
        }
    }
};

}  // namespace algorithm
}  // namespace resaxs

#include "rsasa_hip_gpu.hpp"

#endif
