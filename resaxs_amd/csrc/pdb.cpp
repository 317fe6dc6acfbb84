// PDB input for real proteins: ATOM / HETATM records -> bodies with atomic form-factor
// types (united-atom "complex" types where hydrogens are folded into their heavy atom).
//
// Follows resaxslib/src/parse_pdb.cpp:35-364 (record handling, skipped waters,
// hydrogens and alternate locations, models) and resaxslib/src/utils.cpp:201-314 (the
// element and residue-specific typing tables).  Host code only.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "rsx_capi.h"

namespace {

// atomic form-factor types (the order of the factor table's columns)
enum : int { H, C, N, O, P, S, Zn, CH, CH2, CH3, NH, NH2, NH3, OH, OH2, SH };

std::string trim_spaces(const std::string& s)
{
    const size_t b = s.find_first_not_of(' ');
    if (b == std::string::npos) return std::string();
    const size_t e = s.find_last_not_of(' ');
    return s.substr(b, e - b + 1);
}

// The reference's trim of a record field (utils.cpp:319-325): spaces removed at both ends,
// except that a field of spaces only keeps all but its last character.  Typing depends on
// it: a blank element column reads as " ", an unknown element (typed N), while an element
// column the line does not reach reads as "" and the label's first letter is used.
std::string trim_field(const std::string& s)
{
    const size_t b = s.find_first_not_of(' ');
    if (b == std::string::npos) return s.empty() ? s : s.substr(0, s.size() - 1);
    const size_t e = s.find_last_not_of(' ');
    return s.substr(b, e - b + 1);
}

// element symbol -> type; anything else is typed N by the caller (utils.cpp:201-210)
bool element_type(const std::string& el, int& t)
{
    static const std::map<std::string, int> m = {{"H", H}, {"C", C}, {"N", N}, {"O", O},
                                                 {"P", P}, {"S", S}, {"Zn", Zn}};
    const auto it = m.find(el);
    if (it == m.end()) return false;
    t = it->second;
    return true;
}

// atom label -> default united-atom type, then (label, residue) exceptions
// (utils.cpp:212-290): chemistry tables, kept entry for entry
int united_atom_type(int element, const std::string& label, const std::string& res)
{
    static const std::map<std::string, int> by_label = {{"CH2", CH2}, {"CH3", CH3}, {"CA", CH}, {"CB", CH2},
                                                        {"CG", CH2},  {"CG2", CH3}, {"CD", CH2}, {"N", NH}};
    static const std::map<std::pair<std::string, std::string>, int> by_residue = {
        {{"CA", "GLY"}, CH2},  {{"CB", "ALA"}, CH3},  {{"CB", "ILE"}, CH},   {{"CB", "THR"}, CH},
        {{"CB", "VAL"}, CH},   {{"CG", "ASN"}, C},    {{"CG", "ASP"}, C},    {{"CG", "HIS"}, C},
        {{"CG", "LEU"}, CH},   {{"CG", "PHE"}, C},    {{"CG", "TRP"}, C},    {{"CG", "TYR"}, C},
        {{"CG1", "ILE"}, CH2}, {{"CG1", "VAL"}, CH3}, {{"CD", "GLN"}, C},    {{"CD", "GLU"}, C},
        {{"CD1", "ILE"}, CH3}, {{"CD1", "LEU"}, CH3}, {{"CD1", "PHE"}, CH},  {{"CD1", "TRP"}, CH},
        {{"CD1", "TYR"}, CH},  {{"CD2", "HIS"}, CH},  {{"CD2", "LEU"}, CH3}, {{"CD2", "PHE"}, CH},
        {{"CD2", "TYR"}, CH},  {{"CE", "LYS"}, CH2},  {{"CE", "MET"}, CH3},  {{"CE1", "HIS"}, CH},
        {{"CE1", "PHE"}, CH},  {{"CE1", "TYR"}, CH},  {{"CE2", "PHE"}, CH},  {{"CE2", "TYR"}, CH},
        {{"CE3", "TRP"}, CH},  {{"CZ", "PHE"}, CH},   {{"CZ2", "TRP"}, CH},  {{"CZ3", "TRP"}, CH},
        {{"N", "PRO"}, N},     {{"ND1", "HIS"}, NH},  {{"ND2", "ASN"}, NH2}, {{"NH1", "ARG"}, NH2},
        {{"NH2", "ARG"}, NH2}, {{"NE", "ARG"}, NH},   {{"NE1", "TRP"}, NH},  {{"NE2", "GLN"}, NH2},
        {{"NZ", "LYS"}, NH3},  {{"OG", "SER"}, OH},   {{"OG1", "THR"}, OH},  {{"OH", "TYR"}, OH},
        {{"SG", "CYS"}, SH}};
    int t = element;
    const auto a = by_label.find(label);
    if (a != by_label.end()) t = a->second;
    const auto b = by_residue.find({label, res});
    if (b != by_residue.end()) t = b->second;
    return t;
}

std::string field(const std::string& line, size_t at, size_t len)
{
    return at < line.size() ? line.substr(at, len) : std::string();
}

float number(const std::string& s)
{
    const std::string t = trim_spaces(s);
    return t.empty() ? 0.0f : std::strtof(t.c_str(), nullptr);
}

struct Atom {
    float x, y, z;
    int type;
};

// one ATOM / HETATM record (fixed columns of the PDB format, as the reference's
// sscanf format reads them); false when the record is skipped
bool parse_atom(const std::string& line, int verbose, Atom& out)
{
    const std::string label = trim_field(field(line, 12, 4));
    const char alt = line.size() > 16 ? line[16] : ' ';
    const std::string res = trim_field(field(line, 17, 3));
    std::string el = line.size() >= 78 ? trim_field(line.substr(76, 2)) : std::string();
    if (el.empty()) el = label.substr(0, 1);
    if (res == "HOH") return false;               // waters
    if (alt != ' ' && alt != 'A') return false;   // extra alternate locations
    int t;
    if (!element_type(el, t)) {
        if (verbose >= 1)
            std::fprintf(stderr, "Missing form factor for element %s in residue %s - using N instead\n", el.c_str(),
                         res.c_str());
        t = N;
    }
    if (t == H) return false;  // hydrogens live in the united-atom types
    out.type = united_atom_type(t, label, res);
    out.x = number(field(line, 30, 8));
    out.y = number(field(line, 38, 8));
    out.z = number(field(line, 46, 8));
    return true;
}

}  // namespace

extern "C" {

int rsx_pdb_atom_type(const char* atom_label, const char* res_name, const char* element, int element_index)
{
    if (!atom_label || !res_name || element_index > SH) return RSX_INVALID_ARG;
    const std::string label(atom_label), res(res_name);
    int t = element_index;
    if (t < 0) {
        const std::string el = element ? std::string(element) : label.substr(0, 1);
        if (!element_type(el, t)) {
            if (!element)
                std::fprintf(stderr, "Missing form factor for element %s in residue %s - using N instead\n",
                             el.c_str(), res.c_str());
            t = N;
        }
    }
    return united_atom_type(t, label, res);
}

int rsx_pdb_read(const char* path, int all_models, int verbose, float* bodies, long long capacity,
                 long long* n_atoms, int* offsets, int offsets_capacity, int* n_models)
{
    if (!path || !n_atoms || !n_models) return RSX_INVALID_ARG;
    std::ifstream file(path);
    if (!file.is_open()) return RSX_PDB_FILE_NOT_FOUND;
    std::vector<Atom> atoms;
    std::vector<int> starts;
    bool in_model = false;
    std::string line;
    while (std::getline(file, line)) {
        // (the all-models read drops a last line that has no newline, parse_pdb.cpp:72-73)
        if (all_models && file.eof()) break;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        // records are read into a 100-character buffer: a longer line is cut to 99
        // characters, and ends the single-model read (parse_pdb.cpp:43-48, 73-83)
        if (line.size() > 99) {
            if (!all_models) break;
            line.resize(99);
        }
        const std::string rec = trim_spaces(field(line, 0, 6));
        if (rec == "END") break;
        if (rec == "ENDMDL") {
            if (!all_models) break;
            in_model = false;
        } else if (rec == "MODEL") {
            if (all_models) {
                starts.push_back((int)atoms.size());
                in_model = true;
            }
        } else if (rec == "ATOM" || rec == "HETATM") {
            if (all_models && !in_model) {
                starts.push_back((int)atoms.size());
                in_model = true;
            }
            Atom a;
            if (parse_atom(line, verbose, a)) atoms.push_back(a);
        }
    }
    if (!all_models) starts.assign(1, 0);
    *n_atoms = (long long)atoms.size();
    *n_models = (int)starts.size();
    if (bodies) {
        if (capacity < (long long)atoms.size()) return RSX_INVALID_ARG;
        for (size_t i = 0; i < atoms.size(); ++i) {
            bodies[4 * i] = atoms[i].x;
            bodies[4 * i + 1] = atoms[i].y;
            bodies[4 * i + 2] = atoms[i].z;
            uint32_t t = (uint32_t)atoms[i].type;
            std::memcpy(&bodies[4 * i + 3], &t, 4);
        }
    }
    if (offsets) {
        if (offsets_capacity < (int)starts.size() + 1) return RSX_INVALID_ARG;
        for (size_t m = 0; m < starts.size(); ++m) offsets[m] = starts[m];
        offsets[starts.size()] = (int)atoms.size();
    }
    return RSX_OK;
}

}  // extern "C"
