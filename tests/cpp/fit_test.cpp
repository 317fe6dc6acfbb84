// fit_test.cpp -- drives the ensemble fitting driver (include/resaxs/calc_saxs.hpp) the
// way the calc_saxs tool does (calc_saxs/main.cpp:264-296): the reference profile read
// from a file, scale fitted, water weight over [-2, 4], expansion factor over
// [0.95, 1.05], the calc_cl_saxs evaluator on the GPU.
//
//   fit_test pdb FILE    (no device) builds the driver from a PDB file as the tool does
//                        and prints "models M atoms N q Q" (q: 0.01..0.5, 50 points)
//   fit_test DIR [c w]   reads DIR/models.bin (int32 n_models, n_atoms, then float32
//                        [n_models][n_atoms][4], atomic types in the w bits) and
//                        DIR/profile.dat (q I sigma), runs the fit on device 0 and
//                        writes DIR/fit.bin (float64 expansion factor, water weight,
//                        scale, chi, evaluations, engines) and DIR/fit_iq.bin (float32
//                        [n_q], the scaled best intensity).  With c and w it also
//                        evaluates that point alone: DIR/point.bin (float64 scale,
//                        chi) and DIR/point_iq.bin.
// Exit status 0 only when every in-program check passed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "resaxs/calc_saxs.hpp"

using namespace resaxs;

static int failures = 0;

#define CHECK(cond, what)                                               \
    do {                                                                \
        if (cond) {                                                     \
            std::printf("ok %s\n", what);                               \
        } else {                                                        \
            std::printf("FAIL %s (%s:%d)\n", what, __FILE__, __LINE__); \
            ++failures;                                                 \
        }                                                               \
        std::fflush(stdout);                                            \
    } while (0)

template <typename T>
static void write_bin(const std::string& path, const std::vector<T>& v)
{
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::printf("usage: fit_test DIR [c w]\n");
        return 2;
    }
    const std::string dir = argv[1];
    if (dir == "pdb" && argc >= 3) {
        try {
            calc_params<float> params;
            calc_saxs<float> calc(std::vector<std::string>{argv[2]}, "", true, 0.01f, 0.5f, 50, params,
                                  calc_saxs<float>::QUIET);
            std::printf("models %zu atoms %zu q %zu\n", calc.v_models_.size(), calc.v_bodies_.size(),
                        calc.v_q_.size());
            CHECK(calc.v_q_.size() == 50 && std::fabs(calc.v_q_[49] - 0.5f) < 1e-6f, "q grid");
        } catch (const error& e) {
            std::printf("error %d\n", e.err());
            return 3;
        }
        std::printf("%d failures\n", failures);
        return failures ? 1 : 0;
    }
    try {
        std::ifstream in(dir + "/models.bin", std::ios::binary);
        int hdr[2] = {0, 0};
        in.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
        const int n_models = hdr[0], n_atoms = hdr[1];
        CHECK(in && n_models > 0 && n_atoms > 0, "models header");
        std::vector<float> raw((size_t)n_models * n_atoms * 4);
        in.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)(raw.size() * sizeof(float)));
        CHECK(bool(in), "models read");
        std::vector<std::vector<cl_float4>> models(n_models, std::vector<cl_float4>(n_atoms));
        for (int m = 0; m < n_models; ++m)
            for (int i = 0; i < n_atoms; ++i) {
                const float* b = &raw[((size_t)m * n_atoms + i) * 4];
                cl_float4& d = models[m][i];
                d.s[0] = b[0];
                d.s[1] = b[1];
                d.s[2] = b[2];
                d.s[3] = b[3];
            }

        // the tool's parameters when fitting: scale over (min, max), default ranges for w and c
        calc_params<float> params{profile_param<float>(), profile_param<float>(-2.0f, 4.0f),
                                  profile_param<float>(0.95f, 1.05f),
                                  saxs_profile<float>::read_from_file(dir + "/profile.dat")};
        CHECK(params.scale_.fit() && params.water_weight_.fit() && params.exp_factor_.fit(), "all three fitted");
        const size_t nq = params.ref_profile_.size();
        CHECK(nq > 0, "profile read");
        calc_saxs<float> calc(models, 0.0f, 0.0f, 0, params, calc_saxs<float>::QUIET);
        CHECK(calc.v_q_.size() == nq, "q from the profile");

        calc_cl_saxs<float> eval(algorithm::saxs_enum::saxs_gpu_pt_wf, "gpu", 64);
        fitted_params<float> best = calc.fit_ensemble(eval);
        CHECK(best.intensity_.size() == nq, "best intensity size");
        CHECK(eval.calc_count > 0 && eval.calc_count % (unsigned)n_models == 0, "whole ensembles evaluated");
        CHECK(eval.n_engines() == 1, "one batched engine holds the ensemble");
        std::printf("fit: c %.6f w %.5f scale %.6g chi %.6g after %u model evaluations\n", best.exp_factor_,
                    best.water_weight_, best.scale_, best.chi_, eval.calc_count);
        std::vector<double> res = {best.exp_factor_, best.water_weight_, best.scale_, best.chi_,
                                   (double)eval.calc_count, (double)eval.n_engines()};
        write_bin(dir + "/fit.bin", res);
        write_bin(dir + "/fit_iq.bin", best.intensity_);

        // re-evaluating the best point reproduces it (the engine keeps the laid-out ensemble)
        const unsigned before = eval.calc_count;
        fitted_params<float> again = calc.fit_ensemble(eval, best.exp_factor_, best.water_weight_);
        CHECK(eval.calc_count == before + (unsigned)n_models, "one evaluation per model");
        CHECK(again.chi_ == best.chi_ && again.scale_ == best.scale_, "best point reproduces");
        // the tool's per-model call gives the models' own I(q); their mean is the ensemble mean
        std::vector<double> acc(nq, 0.0);
        std::vector<float> iq;
        for (const auto& m : calc.v_models_) {
            eval.calc(m, iq);
            for (size_t k = 0; k < nq; ++k) acc[k] += iq[k];
        }
        double worst = 0.0;
        for (size_t k = 0; k < nq; ++k) {
            const double mean = acc[k] / n_models, ens = again.intensity_[k] / again.scale_;
            worst = std::max(worst, std::fabs(mean - ens) / std::fabs(ens));
        }
        CHECK(worst < 1e-6, "per-model I(q) average to the ensemble mean");
        CHECK(eval.n_engines() == 1, "per-model calls reuse the laid-out ensemble");

        if (argc >= 4) {
            const float c = (float)std::atof(argv[2]), w = (float)std::atof(argv[3]);
            fitted_params<float> pt = calc.fit_ensemble(eval, c, w);
            write_bin(dir + "/point.bin", std::vector<double>{pt.scale_, pt.chi_});
            write_bin(dir + "/point_iq.bin", pt.intensity_);
            CHECK(pt.chi_ >= best.chi_ * 0.999f || pt.chi_ < 1e-3f, "the search found a minimum at least as good");
        }
    } catch (const error& e) {
        std::printf("FAIL resaxs::error %d %s\n", e.err(), e.what());
        ++failures;
    }
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
