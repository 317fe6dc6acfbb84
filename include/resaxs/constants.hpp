// constants.hpp -- physical constants and atom types of the resaxslib surface
// (resaxslib/include/constants.hpp:26-81).  The enum order and the tables are an
// interface: factor tables are indexed by these type ids, so they are kept
// identical to the reference's.
#ifndef RESAXS_AMD_CONSTANTS_HPP
#define RESAXS_AMD_CONSTANTS_HPP

namespace resaxs {

const double pi = 3.14159265358979323846;

// solvent (HOH) electron density, e / A^3
const float default_solvent_rho = 0.334f;

const float excl_vol_to_r3_coef = static_cast<float>(3.0 / (4.0 * pi * default_solvent_rho));

namespace atoms {

enum atom_type {
    H,
    C,
    N,
    O,
    P,
    S,
    Zn,
    SIMPLE_ATOM_TYPE_LAST = Zn,
    SIMPLE_ATOM_TYPE_COUNT,
    CH = SIMPLE_ATOM_TYPE_COUNT,
    CH2,
    CH3,
    NH,
    NH2,
    NH3,
    OH,
    OH2,
    SH,
    COMPLEX_ATOM_TYPE_LAST = SH,
    ATOM_TYPE_COUNT
};

// excluded volumes (A^3) of H, C, N, O, P, S, Zn
const float excluded_volume[] = {5.15f, 16.44f, 2.49f, 9.13f, 5.73f, 19.86f, 9.85f};

// complex types: base simple type and hydrogen count
const unsigned int complex_atom_stucture[][2] = {{C, 1}, {C, 2}, {C, 3}, {N, 1}, {N, 2}, {N, 3}, {O, 1}, {O, 2}, {S, 1}};

}  // namespace atoms
}  // namespace resaxs

#endif
