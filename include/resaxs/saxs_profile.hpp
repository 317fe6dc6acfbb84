// saxs_profile.hpp -- experimental profile, least-squares scale and chi of the
// resaxslib surface (resaxslib/include/saxs_profile.hpp:26-189).
//
// File format (an interface, kept): whitespace-separated columns q I [sigma], lines
// starting with '#' are comments, malformed lines are skipped; missing sigmas are
// synthesised as 0.15 I (q + 0.001) (|Poisson(10)/10 - 1| + 1) from a
// default-seeded std::default_random_engine.  Sums run in FLT_T in index order.
#ifndef RESAXS_AMD_SAXS_PROFILE_HPP
#define RESAXS_AMD_SAXS_PROFILE_HPP

#include <cassert>
#include <cmath>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "resaxs.hpp"
#include "utils.hpp"

#define SAXS_PROFILE_FILE_NOT_FOUND -1101

namespace resaxs {

template <typename FLT_T>
class saxs_profile {
public:
    saxs_profile() = default;
    saxs_profile(saxs_profile&&) = default;
    saxs_profile(const saxs_profile&) = default;
    saxs_profile& operator=(const saxs_profile&) = default;
    saxs_profile& operator=(saxs_profile&&) = default;

    saxs_profile(std::vector<FLT_T>&& v_q, std::vector<FLT_T>&& v_Iq, std::vector<FLT_T>&& v_error)
        : v_q_(std::move(v_q)), intensity_(std::move(v_Iq)), v_error_(std::move(v_error))
    {
        assert(intensity_.size() == size());
        if (v_error_.size() < size()) generate_errors();
        w_.resize(size());
        for (size_t i = 0; i < size(); ++i) w_[i] = 1 / (v_error_[i] * v_error_[i]);
    }

    void initialize(std::vector<FLT_T>&& v_q)
    {
        v_q_ = std::move(v_q);
        intensity_.resize(v_q_.size());
    }

    void initialize(const std::vector<FLT_T>& v_q) { initialize(std::vector<FLT_T>(v_q)); }

    size_t size() const { return v_q_.size(); }
    bool initialized() const { return size() > 0; }
    explicit operator bool() const { return initialized(); }

    std::vector<FLT_T> v_q_;
    std::vector<FLT_T> intensity_;
    std::vector<FLT_T> v_error_;

    static saxs_profile read_from_file(const std::string& filename)
    {
        std::ifstream in(filename.c_str());
        verify(bool(in), SAXS_PROFILE_FILE_NOT_FOUND, filename + " file not found");
        std::vector<FLT_T> q, iq, err;
        std::string line;
        while (std::getline(in, line)) {
            line = trim(line);
            if (line.empty() || line[0] == '#') continue;
            std::istringstream ls(line);
            FLT_T a, b, c;
            ls >> a >> b;
            if (!ls) continue;
            q.push_back(a);
            iq.push_back(b);
            ls >> c;
            if (ls) err.push_back(c);
        }
        return saxs_profile(std::move(q), std::move(iq), std::move(err));
    }

    // c minimising sum ((I - c I_other) / sigma)^2: sum(I I_other w) / sum(I_other^2 w)
    FLT_T optimize_scale_for(const std::vector<FLT_T>& other) const
    {
        assert(other.size() == size());
        FLT_T num = 0, den = 0;
        for (size_t i = 0; i < size(); ++i) {
            num += intensity_[i] * other[i] * w_[i];
            den += other[i] * other[i] * w_[i];
        }
        return num / den;
    }

    // chi = sqrt(sum ((I - I_other) / sigma)^2 / n)  (chi, not chi^2)
    FLT_T chi_square_chi(const std::vector<FLT_T>& other) const
    {
        assert(other.size() == size());
        FLT_T s = 0;
        for (size_t i = 0; i < size(); ++i) {
            const FLT_T d = intensity_[i] - other[i];
            s += d * d * w_[i];
        }
        return std::sqrt(s / size());
    }

private:
    std::vector<FLT_T> w_;  // 1 / sigma^2

    void generate_errors()
    {
        v_error_.resize(size());
        std::default_random_engine gen;
        std::poisson_distribution<int> poisson(10.0);
        for (size_t i = 0; i < size(); ++i) {
            const double f = std::abs(poisson(gen) / 10.0 - 1.0) + 1.0;
            v_error_[i] = FLT_T(0.15 * intensity_[i] * (v_q_[i] + 0.001) * f);
        }
    }
};

}  // namespace resaxs

#endif
