// bf_lattice.h — compile-time velocity sets shared by host and device code.
//
// Direction order and weights follow BoltzmannFluids/Source/LBMConstants/
// VelocitySet.cpp:53-141 (D2Q9, D3Q15, D3Q19, D3Q27) and :180-203 (D2Q5,
// D3Q7).  The order pairs opposite directions as (odd, odd+1), which the
// in-place update relies on (collide.comp:219-222).  The thermal sets are the
// leading 5 (2-D) / 7 (3-D) directions of the momentum sets, so a thermal
// direction k points the same way as momentum direction k.
//
// rest_fix / rest_fix_t: 1 minus the exact sum of the FP32 weights (a power of
// two, exact in FP32).  The floats nearest to the rational weights sum to a
// little more than 1, and the equilibrium's mass part sum_i w_i (rho - 1)
// would gain (rho - 1)(sum - 1) per cell and tick (7e-5 of the mass after
// 20000 ticks of a rho = 2 liquid).  The rest population's mass part is
// computed as fma(w_0, rho - 1, rest_fix (rho - 1)), one rounding of
// (w_0 + rest_fix)(rho - 1), which leaves the moving weights, and with them
// the lattice's isotropy, untouched.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define BF_HD __host__ __device__
#else
#define BF_HD
#endif

namespace bf {

enum VelocitySetId : int { D2Q5 = 0, D3Q7 = 1, D2Q9 = 2, D3Q15 = 3, D3Q19 = 4, D3Q27 = 5 };

template <int VS> struct Lattice;

template <> struct Lattice<D2Q9> {
    static constexpr int Q = 9, QT = 5, DIM = 2;
    static constexpr int8_t c[9][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
                                       {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}};
    static constexpr float w[9] = {float(4.0 / 9), float(1.0 / 9), float(1.0 / 9), float(1.0 / 9),
                                   float(1.0 / 9), float(1.0 / 36), float(1.0 / 36), float(1.0 / 36),
                                   float(1.0 / 36)};
    static constexpr float rest_fix = -0x1p-27f, rest_fix_t = -0x1p-25f;
    static constexpr float wt[5] = {float(1.0 / 3), float(1.0 / 6), float(1.0 / 6), float(1.0 / 6),
                                    float(1.0 / 6)};
};

template <> struct Lattice<D3Q15> {
    static constexpr int Q = 15, QT = 7, DIM = 3;
    static constexpr int8_t c[15][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                                        {0, 0, 1},   {0, 0, -1},  {1, 1, 1},  {-1, -1, -1}, {1, 1, -1},
                                        {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}};
    static constexpr float w[15] = {float(2.0 / 9),  float(1.0 / 9),  float(1.0 / 9),  float(1.0 / 9),
                                    float(1.0 / 9),  float(1.0 / 9),  float(1.0 / 9),  float(1.0 / 72),
                                    float(1.0 / 72), float(1.0 / 72), float(1.0 / 72), float(1.0 / 72),
                                    float(1.0 / 72), float(1.0 / 72), float(1.0 / 72)};
    static constexpr float rest_fix = -0x1p-27f, rest_fix_t = 0.0f;
    static constexpr float wt[7] = {float(1.0 / 4), float(1.0 / 8), float(1.0 / 8), float(1.0 / 8),
                                    float(1.0 / 8), float(1.0 / 8), float(1.0 / 8)};
};

template <> struct Lattice<D3Q19> {
    static constexpr int Q = 19, QT = 7, DIM = 3;
    static constexpr int8_t c[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                                        {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},
                                        {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
                                        {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
    static constexpr float w[19] = {float(1.0 / 3),  float(1.0 / 18), float(1.0 / 18), float(1.0 / 18),
                                    float(1.0 / 18), float(1.0 / 18), float(1.0 / 18), float(1.0 / 36),
                                    float(1.0 / 36), float(1.0 / 36), float(1.0 / 36), float(1.0 / 36),
                                    float(1.0 / 36), float(1.0 / 36), float(1.0 / 36), float(1.0 / 36),
                                    float(1.0 / 36), float(1.0 / 36), float(1.0 / 36)};
    static constexpr float rest_fix = -0x1p-26f, rest_fix_t = 0.0f;
    static constexpr float wt[7] = {float(1.0 / 4), float(1.0 / 8), float(1.0 / 8), float(1.0 / 8),
                                    float(1.0 / 8), float(1.0 / 8), float(1.0 / 8)};
};

template <> struct Lattice<D3Q27> {
    static constexpr int Q = 27, QT = 7, DIM = 3;
    static constexpr int8_t c[27][3] = {
        {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0}, {0, 0, 1},   {0, 0, -1},
        {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},  {-1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {1, -1, 0},
        {-1, 1, 0}, {1, 0, -1},  {-1, 0, 1}, {0, 1, -1},  {0, -1, 1}, {1, 1, 1},   {-1, -1, -1},
        {1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}};
    static constexpr float w[27] = {
        float(8.0 / 27),  float(2.0 / 27),  float(2.0 / 27),  float(2.0 / 27),  float(2.0 / 27),
        float(2.0 / 27),  float(2.0 / 27),  float(1.0 / 54),  float(1.0 / 54),  float(1.0 / 54),
        float(1.0 / 54),  float(1.0 / 54),  float(1.0 / 54),  float(1.0 / 54),  float(1.0 / 54),
        float(1.0 / 54),  float(1.0 / 54),  float(1.0 / 54),  float(1.0 / 54),  float(1.0 / 216),
        float(1.0 / 216), float(1.0 / 216), float(1.0 / 216), float(1.0 / 216), float(1.0 / 216),
        float(1.0 / 216), float(1.0 / 216)};
    static constexpr float rest_fix = -0x1p-27f, rest_fix_t = 0.0f;
    static constexpr float wt[7] = {float(1.0 / 4), float(1.0 / 8), float(1.0 / 8), float(1.0 / 8),
                                    float(1.0 / 8), float(1.0 / 8), float(1.0 / 8)};
};

// Opposite direction: odd i <-> i+1, 0 <-> 0 (collide.comp:219-222).
BF_HD constexpr int opposite(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

inline int velocity_count(int vs) {
    switch (vs) {
    case D2Q9: return 9;
    case D3Q15: return 15;
    case D3Q19: return 19;
    case D3Q27: return 27;
    }
    return 0;
}

}  // namespace bf
