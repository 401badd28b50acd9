// Generated by nt2_amd/csrc/gen_bs_coeffs.py. Do not edit.
// erfcx on [0, 12]: max relative error 4.29e-11; exp on |r| <= ln2/2: 1.55e-15
#pragma once
namespace nt2 { namespace hip { namespace bsacc {
constexpr double kErfcxZmax = 12.0;
constexpr double kErfcxC = 2.5;
constexpr double kErfcxA = 2.4166666666666665;
constexpr double kErfcxB = -1.4166666666666667;
// erfcx(z) ~ sum_k kErfcx[k] t^k, highest degree first
constexpr double kErfcx[13] = {
    6.187374375672795e-07,
    -3.5383843025712804e-07,
    -7.690960485617884e-06,
    9.15129521535385e-06,
    8.135763182068328e-05,
    -0.00011936340120005095,
    -0.00111201090610538,
    0.00030742244224651633,
    0.019723403418650705,
    0.08698156962705317,
    0.22168333138662857,
    0.3893944633783154,
    0.28305810121215746,
};
// exp(r) ~ sum_k kExp[k] r^k, highest degree first
constexpr double kExp[11] = {
    2.7345196233046845e-07,
    2.764160852741326e-06,
    2.4802455191062774e-05,
    0.00019841166292849566,
    0.0013888887755108541,
    0.008333333388601226,
    0.04166666667286697,
    0.16666666666556415,
    0.4999999999998759,
    1.0000000000000038,
    1.0000000000000004,
};
}}}  // namespace nt2::hip::bsacc
