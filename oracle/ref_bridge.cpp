// ref_bridge.cpp -- C entry points into the reference's own CPU code.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile (target `ref`) compiles this file
// together with the reference's CPU sources where they lie (host_debye.cpp,
// rsasa_algorithm.cpp, rsasa_cpu_alg.cpp and the files they link against) into
// oracle/_ref/libresaxs_ref.so.  The tests load it to pin the oracle against
// the reference itself; bench.py times ref_host_debye as the CPU baseline.
// Nothing in the product library links or calls it.
//
// Bodies are float4 records {x, y, z, type bits}, as cl_float4.
// utils.hpp first: ff_coef.hpp's templates name get_factor_index unqualified
#include "utils.hpp"
#include "ff_coef.hpp"
#include "host_debye.hpp"
#include "parse_pdb.hpp"
#include "rsasa_algorithm.hpp"

#include <memory>
#include <vector>

namespace {

std::vector<cl_float4> to_bodies(const float* b, int n)
{
    std::vector<cl_float4> v(n);
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < 4; ++c) v[i].s[c] = b[4 * i + c];
    return v;
}

int copy_dots(const std::vector<cl_float4>& d, float* out, int cap)
{
    const int n = (int)d.size();
    for (int i = 0; i < n && i < cap; ++i)
        for (int c = 0; c < 4; ++c) out[4 * i + c] = d[i].s[c];
    return n;
}

using dot_params = resaxs::algorithm::rsasa<float, float>::dot_spheres_params;

}  // namespace

extern "C" {

// host_debye<double, float, cl_float4>::calc_curve (host_debye.cpp:54-109)
int ref_host_debye(const float* q, int nq, const float* bodies, int n, const float* table, int n_factors, float* out)
{
    const std::vector<cl_float4> b = to_bodies(bodies, n);
    resaxs::host_debye<double, float, cl_float4>::calc_curve(q, nq, b.data(), n, table, n_factors, out);
    return 0;
}

// atomic_form_factors<float>::generate (ff_coef.hpp:118-159): table [nq][16]
int ref_atomic_ff(const float* q, int nq, float avg_radius, float c, float* out)
{
    std::vector<float> vq(q, q + nq), t;
    resaxs::atomic_form_factors<float>::generate(vq, avg_radius, c, t);
    for (size_t i = 0; i < t.size(); ++i) out[i] = t[i];
    return (int)t.size();
}

// atomic_form_factors<float>::calc_radii (ff_coef.hpp:161-184)
int ref_atomic_radii(float* out16)
{
    std::vector<float> r;
    resaxs::atomic_form_factors<float>::calc_radii(r);
    for (size_t i = 0; i < r.size() && i < 16; ++i) out16[i] = r[i];
    return (int)r.size();
}

// atomic_form_factors<float>::calc_avg_radius (ff_coef.hpp:186-194)
float ref_avg_radius(const float* bodies, int n)
{
    std::vector<float> r;
    resaxs::atomic_form_factors<float>::calc_radii(r);
    return resaxs::atomic_form_factors<float>::calc_avg_radius(to_bodies(bodies, n), r);
}

// rsasa_algorithm.cpp:96-116 (64 dots)
int ref_fibonacci_dots(float radius, float probe, float* out, int cap)
{
    return copy_dots(dot_params(0.5f, probe).create_fibonacci_dots(radius), out, cap);
}

// rsasa_algorithm.cpp:157-208 (32 dots)
int ref_dodeca_icosa_dots(float radius, float probe, float* out, int cap)
{
    return copy_dots(dot_params(0.5f, probe).create_dodeca_icosa_dots(radius), out, cap);
}

// rsasa_algorithm.cpp:58-86 (density-based sphere of dots, used by the CPU SASA)
int ref_density_dots(float density, float radius, float probe, float* out, int cap)
{
    return copy_dots(dot_params(density, probe).create_dots(radius), out, cap);
}

// rsasa_cpu_alg::calc_rsasa (rsasa_cpu_alg.cpp:112-185): exposed fraction per atom
int ref_rsasa_cpu(const float* bodies, int n, float density, float probe, double* out_frac)
{
    using namespace resaxs::algorithm;
    std::unique_ptr<i_rsasa<float, double, algorithm>> alg(
        rsasa<float, double>::create<algorithm>(rsasa_dot_spheres));
    if (!alg) return -1;
    rsasa<float, double>::dot_spheres_params p(density, probe);
    alg->initialize(&p);
    std::vector<double> frac(n);
    alg->calc_rsasa(to_bodies(bodies, n), frac);
    for (int i = 0; i < n; ++i) out_frac[i] = frac[i];
    return 0;
}

// parse_pdb<float>::first_model_as_complex_atoms / all_models_as_complex_atoms
// (parse_pdb.cpp:35-104, records parse_pdb.cpp:302-364), quiet.  Bodies of model m are
// out[offsets[m] .. offsets[m+1]); returns 0, or the resaxs::error code.
int ref_pdb_read(const char* path, int all_models, float* out, long long cap, long long* n_atoms, int* offsets,
                 int off_cap, int* n_models)
{
    try {
        resaxs::parse_pdb<float> p(path);
        p.set_verbose_level(resaxs::parse_pdb<float>::QUIET);
        std::vector<std::vector<cl_float4>> models;
        if (all_models)
            models = p.all_models_as_complex_atoms();
        else
            models.push_back(p.first_model_as_complex_atoms());
        long long k = 0;
        int m = 0;
        for (const auto& mod : models) {
            if (m < off_cap) offsets[m] = (int)k;
            for (const auto& b : mod) {
                if (k < cap)
                    for (int c = 0; c < 4; ++c) out[4 * k + c] = b.s[c];
                ++k;
            }
            ++m;
        }
        if (m < off_cap) offsets[m] = (int)k;
        *n_atoms = k;
        *n_models = m;
        return 0;
    } catch (const resaxs::error& e) {
        return e.err();
    }
}

// map_pdb_atom_to_ff_type (utils.cpp:283-314): element NULL -> the label overload (element
// from the label's first letter); else the element symbol through pdb_to_ff_simple_atom_map,
// an unknown symbol typed N as parse_pdb.cpp:329-337 does
int ref_pdb_atom_type(const char* label, const char* res, const char* element)
{
    if (!element) return resaxs::map_pdb_atom_to_ff_type(std::string(label), std::string(res));
    const auto it = resaxs::pdb_to_ff_simple_atom_map.find(element);
    const resaxs::atoms::atom_type el = it != resaxs::pdb_to_ff_simple_atom_map.end() ? it->second : resaxs::atoms::N;
    return resaxs::map_pdb_atom_to_ff_type(el, std::string(label), std::string(res));
}

}  // extern "C"
