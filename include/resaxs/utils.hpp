// utils.hpp -- helpers of the resaxslib surface (resaxslib/include/utils.hpp:29-131):
// device specs, the type index in a body's w bits, alignment and string helpers.
#ifndef RESAXS_AMD_UTILS_HPP
#define RESAXS_AMD_UTILS_HPP

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "constants.hpp"
#include "resaxs.hpp"

namespace resaxs {

// Lists the HIP devices the engine can run on (the reference lists OpenCL platforms).
inline void report_opencl_caps()
{
    int n = 0;
    rsx_device_count(&n);
    std::printf("HIP platform 0: %d device(s)\n", n);
    for (int d = 0; d < n; ++d) {
        char name[256];
        if (rsx_device_name(d, name, sizeof name) == RSX_OK) std::printf("  device 0-%d: %s\n", d, name);
    }
}

// Parses one device spec: "gpu", "cpu", "<platform>-<device>" or "<platform>-<device>(cpu)"
// (utils.cpp:101-142 grammar).  A malformed spec throws int -1, as the reference does.
inline void parse_dev_spec(const std::string& spec, std::vector<dev_id>& ids)
{
    dev_id d{-1, -1, GPU};
    if (spec == "gpu") {
        d.type = GPU;
    } else if (spec == "cpu") {
        d.type = CPU;
    } else {
        const std::string::size_type dash = spec.find('-');
        if (dash == std::string::npos) throw -1;
        const std::string plat = spec.substr(0, dash);
        std::string dev;
        const std::string::size_type paren = spec.find('(', dash + 1);
        if (paren == std::string::npos) {
            dev = spec.substr(dash + 1);
        } else {
            dev = spec.substr(dash + 1, paren - dash - 1);
            if (spec.compare(paren + 1, 3, "cpu") == 0) d.type = CPU;
        }
        if (plat.empty() || dev.empty()) throw -1;
        char* end;
        d.platform = std::strtol(plat.c_str(), &end, 0);
        d.device = std::strtol(dev.c_str(), &end, 0);
    }
    ids.push_back(d);
}

template <class T>
inline T sinc(T x)
{
    return std::sin(x) / x;
}

// the atom-type index kept in the bit pattern of w (uint32 for float bodies, uint64 for double)
inline unsigned int get_factor_index(const cl_float4& b)
{
    unsigned int i;
    std::memcpy(&i, &b.s[3], sizeof i);
    return i;
}

inline unsigned int get_factor_index(const cl_double4& b)
{
    unsigned long long i;
    std::memcpy(&i, &b.s[3], sizeof i);
    return (unsigned int)i;
}

inline void set_factor_index(cl_float4& b, unsigned int index)
{
    std::memcpy(&b.s[3], &index, sizeof index);
}

inline void set_factor_index(cl_double4& b, unsigned int index)
{
    const unsigned long long i = index;
    std::memcpy(&b.s[3], &i, sizeof i);
}

// PDB atom typing through the library's tables (utils.cpp:292-314): the element from the
// label's first letter (unknown: N), or given as a type
inline atoms::atom_type map_pdb_atom_to_ff_type(const std::string& atom_label, const std::string& res_type)
{
    return (atoms::atom_type)rsx_pdb_atom_type(atom_label.c_str(), res_type.c_str(), nullptr, -1);
}

inline atoms::atom_type map_pdb_atom_to_ff_type(atoms::atom_type element, const std::string& atom_label,
                                                const std::string& res_type)
{
    return (atoms::atom_type)rsx_pdb_atom_type(atom_label.c_str(), res_type.c_str(), nullptr, (int)element);
}

// num rounded up to a multiple of stride
template <typename T>
inline T aligned_num(T num, unsigned int stride)
{
    return ((num - 1) / stride + 1) * stride;
}

inline std::string trim(const std::string& s, const std::string& delim)
{
    const std::string::size_type b = s.find_first_not_of(delim);
    if (b == std::string::npos) return std::string();
    const std::string::size_type e = s.find_last_not_of(delim);
    return s.substr(b, e - b + 1);
}

// spaces only, as utils.cpp:319-325
inline std::string trim(const std::string& s)
{
    return trim(s, " ");
}

inline std::vector<std::string> split(const std::string& s, char delim)
{
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, delim)) out.push_back(item);
    return out;
}

}  // namespace resaxs

#endif
