// sky_host.cpp — host-side Hosek-Wilkie coefficient fit.
//
// Restates HosekWilkieSkyModel::update's CPU part
// (src/engine/gfx/hosek_wilkie_sky_model.cpp:41-75, 658-701): quintic Bezier
// splines in elevation^(1/3), lerped over turbidity and ground albedo, with
// H and I swapped in the dataset, and Z normalized so that the luminance of
// the sun direction equals normalized_sun_y. The dataset
// (helios_amd/data/hosek_rgb.bin, 3600 float64) is linked into the library.
#include <algorithm>
#include <cmath>
#include <cstddef>

#include "internal.hpp"

#if !defined(__HIP_DEVICE_COMPILE__)
__asm__(".section .rodata\n.balign 16\n.global hh_hosek_rgb_data\nhh_hosek_rgb_data:\n.incbin \"" HH_HOSEK_BIN "\"\n.previous\n");
#endif
extern "C" const double hh_hosek_rgb_data[3600];

namespace hh {

static const double* rgb(int ch) { return hh_hosek_rgb_data + ch * 1080; }
static const double* rad(int ch) { return hh_hosek_rgb_data + 3 * 1080 + ch * 120; }

static double evaluate_spline(const double* spline, size_t stride, double value) {
    return 1 * std::pow(1 - value, 5) * spline[0 * stride] + 5 * std::pow(1 - value, 4) * std::pow(value, 1) * spline[1 * stride] +
           10 * std::pow(1 - value, 3) * std::pow(value, 2) * spline[2 * stride] +
           10 * std::pow(1 - value, 2) * std::pow(value, 3) * spline[3 * stride] +
           5 * std::pow(1 - value, 1) * std::pow(value, 4) * spline[4 * stride] + 1 * std::pow(value, 5) * spline[5 * stride];
}

static double evaluate(const double* dataset, size_t stride, float turbidity, float albedo, float sun_theta) {
    // The reference rounds the elevation term to float (std::max<float>) and
    // takes the float overload of pow.
    float e = std::max<float>(0.f, (float)(1.f - sun_theta / (3.14159265358979323846 / 2.f)));
    double elevation_k = std::pow(e, 1.f / 3.0f);
    int t0 = std::min(std::max((int)turbidity, 1), 10);
    int t1 = std::min(t0 + 1, 10);
    float tk = std::min(std::max(turbidity - t0, 0.f), 1.f);
    const double* a0 = dataset;
    const double* a1 = dataset + stride * 6 * 10;
    double a0t0 = evaluate_spline(a0 + stride * 6 * (t0 - 1), stride, elevation_k);
    double a1t0 = evaluate_spline(a1 + stride * 6 * (t0 - 1), stride, elevation_k);
    double a0t1 = evaluate_spline(a0 + stride * 6 * (t1 - 1), stride, elevation_k);
    double a1t1 = evaluate_spline(a1 + stride * 6 * (t1 - 1), stride, elevation_k);
    return a0t0 * (1 - albedo) * (1 - tk) + a1t0 * albedo * (1 - tk) + a0t1 * (1 - albedo) * tk + a1t1 * albedo * tk;
}

static void hosek_host(float cos_theta, float gamma, float cos_gamma, const float* c, float out[3]) {
    for (int i = 0; i < 3; i++) {
        float A = c[0 + i], B = c[3 + i], C = c[6 + i], D = c[9 + i], E = c[12 + i], F = c[15 + i], G = c[18 + i],
              H = c[21 + i], I = c[24 + i];
        float chi = (1.f + cos_gamma * cos_gamma) / std::pow(1.f + H * H - 2.f * cos_gamma * H, 1.5f);
        out[i] = (1.f + A * std::exp(B / (cos_theta + 0.01f))) *
                 (C + D * std::exp(E * gamma) + F * (cos_gamma * cos_gamma) + G * chi + I * (float)std::sqrt(std::max(0.f, cos_theta)));
    }
}

void sky_coefficients(float turbidity, float albedo, float normalized_sun_y, const float sun_dir[3], float out[30]) {
    const float sun_theta = std::acos(std::min(std::max(sun_dir[1], 0.f), 1.f));
    for (int i = 0; i < 3; i++) {
        const double* d = rgb(i);
        for (int k = 0; k < 7; k++) out[k * 3 + i] = (float)evaluate(d + k, 9, turbidity, albedo, sun_theta);
        out[7 * 3 + i] = (float)evaluate(d + 8, 9, turbidity, albedo, sun_theta);
        out[8 * 3 + i] = (float)evaluate(d + 7, 9, turbidity, albedo, sun_theta);
        out[9 * 3 + i] = (float)evaluate(rad(i), 1, turbidity, albedo, sun_theta);
    }
    if (normalized_sun_y != 0.f) {
        float S[3];
        hosek_host(std::cos(sun_theta), 0, 1.f, out, S);
        for (int i = 0; i < 3; i++) S[i] *= out[27 + i];
        float lum = S[0] * 0.2126f + S[1] * 0.7152f + S[2] * 0.0722f;
        for (int i = 0; i < 3; i++) out[27 + i] = out[27 + i] / lum * normalized_sun_y;
    }
}

}  // namespace hh
