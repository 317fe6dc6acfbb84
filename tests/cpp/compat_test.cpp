// compat_test.cpp -- drives the resaxslib-compatible C++ surface (include/resaxs/)
// the way the reference's own harnesses do (test_resaxs/SAXSTest.cpp:276-330,
// test_common/alg_test_data.cpp:90-130): factory, params.initialize, initialize,
// calc_curve, recalc_curve with commit / revert, i_rsasa and host_debye.
//
//   compat_test cpu          checks that need no device (parse_dev_spec, errors,
//                            body bits, generators, saxs_profile); prints "ok ..."
//   compat_test gpu DIR      reads DIR/q.bin (float32 [n_q]), DIR/bodies.bin
//                            (float32 [n][4], atomic types 0..15), DIR/moves.bin
//                            (int32 count, then per move int32 start, len and
//                            float32 [len][4] bodies), runs every path on device 0 and
//                            writes its results to DIR/*.bin for the Python side to
//                            check against the oracle.
// Exit status 0 only when every in-program check passed.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "resaxs/ff_coef.hpp"
#include "resaxs/host_debye.hpp"
#include "resaxs/rsasa_algorithm.hpp"
#include "resaxs/saxs_algorithm.hpp"
#include "resaxs/saxs_profile.hpp"
#include "resaxs/utils.hpp"

using namespace resaxs;

static int failures = 0;

#define CHECK(cond, what)                                                          \
    do {                                                                           \
        if (cond) {                                                                \
            std::printf("ok %s\n", what);                                          \
        } else {                                                                   \
            std::printf("FAIL %s (%s:%d)\n", what, __FILE__, __LINE__);            \
            ++failures;                                                            \
        }                                                                          \
        std::fflush(stdout);                                                       \
    } while (0)

static int error_code_of(const std::function<void()>& f)
{
    try {
        f();
    } catch (const error& e) {
        return e.err();
    } catch (int e) {
        return 1000 + e;
    }
    return 0;
}

template <typename T>
static std::vector<T> read_bin(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    std::vector<char> raw((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::vector<T> v(raw.size() / sizeof(T));
    if (!v.empty()) std::memcpy(v.data(), raw.data(), v.size() * sizeof(T));
    return v;
}

template <typename T>
static void write_bin(const std::string& path, const std::vector<T>& v)
{
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

static int run_cpu()
{
    // parse_dev_spec grammar (utils.cpp:101-142)
    std::vector<dev_id> ids;
    parse_dev_spec("gpu", ids);
    parse_dev_spec("cpu", ids);
    parse_dev_spec("0-1", ids);
    parse_dev_spec("1-2(cpu)", ids);
    CHECK(ids.size() == 4, "parse_dev_spec count");
    CHECK(ids[0] == (dev_id{-1, -1, GPU}), "parse gpu");
    CHECK(ids[1] == (dev_id{-1, -1, CPU}), "parse cpu");
    CHECK(ids[2] == (dev_id{0, 1, GPU}), "parse 0-1");
    CHECK(ids[3] == (dev_id{1, 2, CPU}), "parse 1-2(cpu)");
    CHECK(error_code_of([] { std::vector<dev_id> v; parse_dev_spec("gpu0", v); }) == 1000 - 1, "malformed spec throws -1");
    CHECK(error_code_of([] { std::vector<dev_id> v; parse_dev_spec("-3", v); }) == 1000 - 1, "empty platform throws -1");
    CHECK(ids[2].to_string() == "0-1", "dev_id to_string");

    // errors and verify (resaxs.hpp:187-221, resaxs.cpp:33-46)
    CHECK(error::SAXS_INVALID_DEVICE == -1 && error::SAXS_NO_INITIAL_CALC == -6, "error codes");
    try {
        verify(false, error::SAXS_INVALID_ARG);
    } catch (const error& e) {
        CHECK(e.err() == -3 && std::string(e.what()) == "RESAXS::error", "verify default message");
    }

    // body bits (utils.hpp:61-88)
    cl_float4 bf;
    set_factor_index(bf, 14u);
    cl_double4 bd;
    set_factor_index(bd, 9u);
    CHECK(get_factor_index(bf) == 14u && get_factor_index(bd) == 9u, "factor index bits");
    CHECK(aligned_num(65u, 64) == 128u && aligned_num(64u, 64) == 64u, "aligned_num");
    CHECK(trim("  a b  ") == "a b" && split("1,2,,3", ',').size() == 4, "trim / split");

    // default factors and radii: the library's table (bit-exact to ff_coef.hpp's float path)
    std::vector<float> q{0.0f, 0.1f, 0.3f}, t, r;
    atomic_form_factors<float>::generate(q, t);
    atomic_form_factors<float>::calc_radii(r);
    CHECK(t.size() == 3 * 16 && r.size() == 16, "form factor table shape");
    CHECK(std::fabs(t[atoms::OH2] - 3.51f) < 0.01f, "f_OH2(0) ~ 3.5");
    std::vector<cl_float4> two(2);
    set_factor_index(two[0], atoms::C);
    set_factor_index(two[1], atoms::O);
    CHECK(atomic_form_factors<float>::calc_avg_radius(two, r) == (r[atoms::C] + r[atoms::O]) / 2, "avg radius");

    // dot spheres (rsasa_algorithm.cpp:58-208)
    algorithm::rsasa<float>::dot_spheres_params dp(0.5f, 1.8f);
    CHECK(dp.create_fibonacci_dots(1.5f).size() == 64, "fibonacci 64 dots");
    CHECK(dp.create_dodeca_icosa_dots(1.5f).size() == 32, "dodeca-icosa 32 dots");
    CHECK(dp.create_dodecahedron_dots(1.5f).size() == 20, "dodecahedron 20 dots");
    const auto dd = dp.create_dots(1.7f);
    bool on_sphere = !dd.empty();
    for (const auto& d : dd) on_sphere &= std::fabs(std::sqrt(d.s[0] * d.s[0] + d.s[1] * d.s[1] + d.s[2] * d.s[2]) - 3.5f) < 1e-4f;
    CHECK(on_sphere, "density dots on the probe sphere");

    // saxs_profile: file format, scale and chi (saxs_profile.hpp:95-166)
    const char* path = "compat_profile.tmp";
    {
        std::ofstream f(path);
        f << "# q I sigma\n0.1 10 1\n  0.2 8 2\nbad line\n0.3 6 0.5\n";
    }
    auto prof = saxs_profile<double>::read_from_file(path);
    std::remove(path);
    CHECK(prof.size() == 3 && prof.intensity_[1] == 8.0 && prof.v_error_[2] == 0.5, "read_from_file");
    const std::vector<double> other{5.0, 4.0, 3.0};
    const double c = prof.optimize_scale_for(other);
    const double num = 10 * 5 / 1.0 + 8 * 4 / 4.0 + 6 * 3 / 0.25, den = 25 / 1.0 + 16 / 4.0 + 9 / 0.25;
    CHECK(std::fabs(c - num / den) < 1e-12, "optimize_scale_for closed form");
    const double chi = prof.chi_square_chi(other);
    CHECK(std::fabs(chi - std::sqrt((25.0 / 1 + 16.0 / 4 + 9.0 / 0.25) / 3)) < 1e-12, "chi_square_chi closed form");
    CHECK(error_code_of([] { saxs_profile<float>::read_from_file("/nonexistent/x.dat"); }) == SAXS_PROFILE_FILE_NOT_FOUND,
          "missing profile file");
    return failures ? 1 : 0;
}

static int run_gpu(const std::string& dir)
{
    using namespace resaxs::algorithm;
    const std::vector<float> q = read_bin<float>(dir + "/q.bin");
    const std::vector<float> flat = read_bin<float>(dir + "/bodies.bin");
    std::vector<cl_float4> bodies(flat.size() / 4);
    std::memcpy(bodies.data(), flat.data(), flat.size() * sizeof(float));
    const std::vector<int> mv_raw = read_bin<int>(dir + "/moves.bin");
    std::printf("case: compat gpu, %zu bodies x %zu q, %d moves\n", bodies.size(), q.size(), mv_raw.empty() ? 0 : mv_raw[0]);
    std::fflush(stdout);

    std::vector<dev_id> devices;
    parse_dev_spec("gpu", devices);

    // ---- error paths before any device work
    {
        std::unique_ptr<i_saxs<float, cl_base<float>>> s(saxs<float>::create<cl_base<float>>(saxs_enum::saxs_gpu_pt_wf));
        std::vector<float> out;
        CHECK(error_code_of([&] { s->calc_curve(bodies, out); }) == error::SAXS_ARGS_NOT_SET, "calc before initialize");
        s->access_params().initialize(q, bodies, std::vector<dev_id>{dev_id{-1, -1, CPU}});
        CHECK(error_code_of([&] { s->initialize(); }) == error::SAXS_INVALID_DEVICE, "CPU device rejected");
    }

    // ---- plain I(q), harness flow (SAXSTest.cpp:276-330)
    std::unique_ptr<i_saxs<float, cl_base<float>>> alg(saxs<float>::create<cl_base<float>>(saxs_enum::saxs_gpu_pt_wf));
    auto& params = alg->access_params();
    params.initialize(q, bodies, devices, SAXS_AMD_WF_SIZE);
    params.get_implicit_water_params().set_water_weight(0);
    alg->initialize();
    write_bin(dir + "/factors.bin", params.get_ff_params().access_factors());
    write_bin(dir + "/water_index.bin", std::vector<int>{params.get_ff_params().water_ff_index_});
    std::vector<float> iq;
    CHECK(error_code_of([&] { alg->recalc_curve(bodies, 0, 1, iq); }) == error::SAXS_NO_INITIAL_CALC,
          "recalc before calc");
    CHECK(error_code_of([&] { alg->commit(); }) == error::SAXS_NO_INITIAL_CALC, "commit before calc");
    std::vector<cl_float4> short_bodies(bodies.begin(), bodies.end() - 1);
    CHECK(error_code_of([&] { alg->calc_curve(short_bodies, iq); }) == error::SAXS_INVALID_ARG, "wrong body count");
    alg->calc_curve(bodies, iq);
    CHECK(iq.size() == q.size(), "I(q) size");
    write_bin(dir + "/iq_plain.bin", iq);
    CHECK(error_code_of([&] { alg->recalc_curve(bodies, (unsigned)bodies.size(), 1, iq); }) == error::SAXS_INVALID_ARG,
          "update region past the end");

    // ---- moves: recalc, commit every move except the last, which is reverted
    std::vector<cl_float4> cur = bodies;
    size_t p = 1;
    const int n_moves = mv_raw.empty() ? 0 : mv_raw[0];
    std::vector<float> committed;
    for (int m = 0; m < n_moves; ++m) {
        const int start = mv_raw[p], len = mv_raw[p + 1];
        p += 2;
        std::vector<cl_float4> saved(cur.begin() + start, cur.begin() + start + len);
        std::memcpy(&cur[start], &mv_raw[p], (size_t)len * sizeof(cl_float4));
        p += (size_t)len * 4;
        alg->recalc_curve(cur, (unsigned)start, (unsigned)len, iq);
        if (m + 1 < n_moves) {
            alg->commit();
            committed = iq;
        } else {
            write_bin(dir + "/iq_last_move.bin", iq);
            alg->revert();
            std::copy(saved.begin(), saved.end(), cur.begin() + start);
            alg->recalc_curve(cur, 0, 0, iq);
            write_bin(dir + "/iq_reverted.bin", iq);
        }
    }
    write_bin(dir + "/iq_committed.bin", committed);
    write_bin(dir + "/bodies_committed.bin", std::vector<float>(&cur[0].s[0], &cur[0].s[0] + 4 * cur.size()));

    // ---- hydration: the water weight is a dirty parameter -> full recompute on the next recalc
    params.get_implicit_water_params().set_water_weight(1.5f);
    alg->recalc_curve(cur, 0, 0, iq);
    write_bin(dir + "/iq_hydrated.bin", iq);
    write_bin(dir + "/rsasa_hydrated.bin", params.get_implicit_water_params().access_rsasa());
    // expansion factor regenerates the table (saxs_alg_cl_gpu_pt_wf.cpp:51-63)
    params.get_ff_params().set_expansion_factor(1.05f);
    alg->recalc_curve(cur, 0, 0, iq);
    write_bin(dir + "/iq_expanded.bin", iq);
    write_bin(dir + "/factors_expanded.bin", params.get_ff_params().access_factors());

    // ---- double instantiation and the convenience factory
    {
        std::unique_ptr<i_saxs<double, cl_base<double>>> d(create_saxs_ocl<double>());
        std::vector<double> qd(q.begin(), q.end());
        std::vector<cl_double4> bd(bodies.size());
        for (size_t i = 0; i < bodies.size(); ++i) {
            for (int c = 0; c < 3; ++c) bd[i].s[c] = bodies[i].s[c];
            set_factor_index(bd[i], get_factor_index(bodies[i]));
        }
        d->access_params().initialize(qd, bd, devices);
        d->initialize();
        std::vector<double> id;
        d->calc_curve(bd, id);
        write_bin(dir + "/iq_double.bin", id);
    }

    // ---- the differentiable-fit extension of the engine class: chi^2 and its gradient, the VJP;
    // the gradient against central differences of chi^2 (h = 0.05 A) on the body with the
    // largest gradient
    {
        saxs_hip_gpu<float> g;
        g.access_params().initialize(q, bodies, devices);
        g.initialize();
        std::vector<float> i0;
        g.calc_curve(bodies, i0);
        std::vector<float> i_exp(i0.size()), sigma(i0.size());
        for (size_t k = 0; k < i0.size(); ++k) {
            i_exp[k] = i0[k] * (1.0f + 0.02f * std::sin(0.7f * (float)k));
            sigma[k] = 0.01f * i0[k];
        }
        std::vector<cl_float4> moved = bodies;
        moved[0].s[0] += 0.3f;
        double scale = 0.0;
        std::vector<float> grad;
        const double chi2 = g.chi2_grad(moved, i_exp, sigma, scale, grad);
        CHECK(grad.size() == 3 * bodies.size() && chi2 > 0.0 && scale > 0.0, "chi2_grad sizes and values");
        size_t top = 0;
        for (size_t i = 0; i < bodies.size(); ++i)
            if (std::fabs(grad[3 * i]) + std::fabs(grad[3 * i + 1]) + std::fabs(grad[3 * i + 2]) >
                std::fabs(grad[3 * top]) + std::fabs(grad[3 * top + 1]) + std::fabs(grad[3 * top + 2]))
                top = i;
        for (int c = 0; c < 3; ++c) {
            const float h = 0.05f;
            std::vector<cl_float4> p = moved, m = moved;
            p[top].s[c] += h;
            m[top].s[c] -= h;
            double sp = 0.0, sm = 0.0;
            std::vector<float> gp, gm;
            const double cp = g.chi2_grad(p, i_exp, sigma, sp, gp), cm = g.chi2_grad(m, i_exp, sigma, sm, gm);
            const double fd = (cp - cm) / (2.0 * h);
            const double gmax = std::fabs(grad[3 * top]) + std::fabs(grad[3 * top + 1]) + std::fabs(grad[3 * top + 2]);
            std::printf("  body %zu axis %d: gradient %.6g, central difference %.6g\n", top, c, grad[3 * top + c], fd);
            CHECK(std::fabs(fd - grad[3 * top + c]) <= 1e-2 * gmax, "chi2 gradient vs central difference");
        }
        std::vector<double> w(q.size());
        for (size_t k = 0; k < w.size(); ++k) w[k] = (double)k / (double)w.size() - 0.5;
        std::vector<float> gv;
        g.vjp(moved, w, gv);
        std::vector<float> gc(3 * bodies.size());
        std::vector<float> flat(4 * bodies.size());
        for (size_t i = 0; i < bodies.size(); ++i)
            for (int c = 0; c < 4; ++c) flat[4 * i + c] = moved[i].s[c];
        detail::check(rsx_debye_vjp(g.engine(), flat.data(), (int)bodies.size(), w.data(), gc.data()), g.engine(),
                      "vjp");
        CHECK(gv == gc, "vjp equals the C ABI's");
    }

    // ---- the saxs_hist_cl_gpu slot: the exact sum, as the reference's class of that name
    {
        std::unique_ptr<i_saxs<float, cl_base<float>>> sg(
            saxs<float>::create<cl_base<float>>(saxs_enum::saxs_hist_cl_gpu));
        CHECK(sg != nullptr, "saxs_hist_cl_gpu engine created");
        sg->access_params().initialize(q, bodies, devices);
        sg->initialize();
        std::vector<float> is;
        sg->calc_curve(bodies, is);
        write_bin(dir + "/iq_hist_slot.bin", is);
    }

    // ---- the distance-histogram engine, constructed directly (default 0.01 A bins, plain)
    {
        std::unique_ptr<i_saxs<float, cl_base<float>>> hg(new saxs_hip_hist<float>());
        hg->access_params().initialize(q, bodies, devices);
        hg->initialize();
        std::vector<float> ih;
        hg->calc_curve(bodies, ih);
        write_bin(dir + "/iq_hist.bin", ih);
        std::vector<float> ih2;
        hg->recalc_curve(bodies, 0, 0, ih2);
        CHECK(ih2 == ih, "histogram recalc of the same bodies repeats");
    }

    // ---- host_debye (host_debye.hpp:30-40) with the default factor table
    {
        std::vector<float> t = params.get_ff_params().access_factors();
        std::vector<float> qh(q), out(q.size());
        std::vector<float> t0;
        atomic_form_factors<float>::generate(qh, t0);
        host_debye<double, float, cl_float4>::calc_curve(qh, bodies, t0, out);
        write_bin(dir + "/iq_host_debye.bin", out);
        CalcSaxsCurveFn fn = &host_debye<double, float, cl_float4>::calc_curve;
        std::vector<float> out2(q.size());
        fn(qh.data(), (int)qh.size(), bodies.data(), (int)bodies.size(), t0.data(), 16, out2.data());
        CHECK(out == out2, "CalcSaxsCurveFn pointer");
    }

    // ---- i_rsasa, GPU set-up (dodeca-icosa) and CPU set-up (density rings)
    {
        std::unique_ptr<i_rsasa<float, float, cl_base<float>>> r(
            rsasa<float, float>::create<cl_base<float>>(rsasa_dot_spheres));
        std::vector<float> out(bodies.size() - 1);
        rsasa<float, float>::dot_spheres_cl_params rp(0.5f, 1.8f, devices, (unsigned)bodies.size());
        CHECK(error_code_of([&] { r->recalc_rsasa(bodies, out); }) == error::SAXS_NO_INITIAL_CALC, "rsasa recalc first");
        r->initialize(&rp);
        CHECK(error_code_of([&] { r->calc_rsasa(bodies, out); }) == error::SAXS_INVALID_ARG, "rsasa output too small");
        out.resize(bodies.size());
        r->calc_rsasa(bodies, out);
        write_bin(dir + "/rsasa_gpu.bin", out);
        r->recalc_rsasa(bodies, out);

        std::unique_ptr<i_rsasa<float, double, algorithm::algorithm>> rc(
            rsasa<float, double>::create<algorithm::algorithm>(rsasa_dot_spheres));
        rsasa<float, double>::dot_spheres_params cp(2.0f, 1.8f);
        rc->initialize(&cp);
        std::vector<double> outc(bodies.size());
        rc->calc_rsasa(bodies, outc);
        write_bin(dir + "/rsasa_cpu_setup.bin", outc);
    }
    std::printf("done, %d failure(s)\n", failures);
    return failures ? 1 : 0;
}

int main(int argc, char** argv)
{
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    try {
        if (mode == "cpu") return run_cpu();
        if (mode == "gpu" && argc > 2) return run_gpu(argv[2]);
    } catch (const error& e) {
        std::printf("FAIL resaxs::error %d: %s\n", e.err(), e.what());
        return 2;
    }
    std::printf("usage: compat_test cpu | gpu DIR\n");
    return 2;
}
