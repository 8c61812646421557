// as_model32.h -- per-cell arithmetic of the FP32-compensated precision.
//
// State and right-hand sides are FP32 with hardware exp/log/reciprocal. The
// astrocyte variables Ca, IP3 and h change by ~1e-6 of their value per step,
// below FP32 resolution over thousands of steps, so each is carried as an
// unevaluated hi+lo pair of floats and every increment is added with an
// error-free TwoSum and renormalised (about 48 significant bits). The AdEx
// spike current g_L*DeltaT*exp((V-V_T)/DeltaT) is evaluated in FP64 for every
// V (spike_current32: one fused argument, a shifter reduction and a degree-8
// polynomial, rounded to FP32 once): the band |V - V_T| < 3*DeltaT, where the
// spike onset is decided and FP64 is required, holds some lane of almost
// every wave, so a per-lane choice would run both paths. Poisson counts use
// the FP64 path's 64-bit draws and FP64 products (bit-identical counts);
// Gaussian noise uses the same two uniforms per draw with the Box-Muller
// transform in FP32 (normal32: hardware log2 and cosine), within FP32
// rounding of the FP64 draw. Formulas follow src/astrosim_fused.cu:236-290,
// 508-576.
#pragma once

#include <cmath>

#include "as_model.h"

namespace astrosim {

struct AstroConst32 {
    float Kd_IP3_1, Kd_IP3_2, Kd_act, Kd_inh, Km_SERCA2;
    float k_IP3R, rate_IP3R, rate_L, rate_SERCA, ratio_ER_cyt, inv_ratio_ER_cyt;
    float SIC_scale, SIC_th;
};

struct NeuronConst32 {
    float inv_C_m, g_L, E_L, V_th, Delta_T, a, inv_tau_w, V_peak, E_ex, E_in, I_e;
    float psc_init_ex, psc_init_in;
    int refractory_steps;                                   // int(t_ref/h + 0.5)
    double x_scale_d, x_shift_d;  // spike-current argument: V * x_scale_d + x_shift_d
};

struct Params32 {
    AstroConst32 ac;
    NeuronConst32 ne, ni;
    float h_step, h_half, h_sixth, ex_decay_e, in_decay_e, ex_decay_i, in_decay_i;
    float w_e_abs, w_i_abs, w_n2a, w_a2n, w_poiss_astro, w_poiss_exc, w_poiss_inh;
    float noise_astro, noise_exc, noise_inh;
};

inline Params32 make_params32(const RunParams& P) {
    Params32 q{};
    const AstroConst& c = P.ac;
    q.ac = {float(c.Kd_IP3_1), float(c.Kd_IP3_2), float(c.Kd_act), float(c.Kd_inh),
            float(c.Km_SERCA * c.Km_SERCA), float(c.k_IP3R), float(c.rate_IP3R), float(c.rate_L),
            float(c.rate_SERCA), float(c.ratio_ER_cyt), float(1.0 / c.ratio_ER_cyt), float(c.SIC_scale),
            float(c.SIC_th)};
    auto nc = [&](const NeuronConst& p) {
        NeuronConst32 n{};
        n.inv_C_m = float(1.0 / p.C_m);
        n.g_L = float(p.g_L);
        n.E_L = float(p.E_L);
        n.V_th = float(p.V_th);
        n.Delta_T = float(p.Delta_T);
        n.a = float(p.a);
        n.inv_tau_w = float(1.0 / p.tau_w);
        n.V_peak = float(p.V_peak);
        n.E_ex = float(p.E_ex);
        n.E_in = float(p.E_in);
        n.I_e = float(p.I_e);
        n.psc_init_ex = float(p.psc_init_ex);
        n.psc_init_in = float(p.psc_init_in);
        n.refractory_steps = static_cast<int>(p.t_ref / P.h_step + 0.5);
        n.x_scale_d = 1.0 / p.Delta_T;
        n.x_shift_d = std::log(p.g_L * p.Delta_T) - p.V_th / p.Delta_T;
        return n;
    };
    q.ne = nc(P.ne);
    q.ni = nc(P.ni);
    q.h_step = float(P.h_step);
    q.h_half = float(0.5 * P.h_step);
    q.h_sixth = float(P.h_step / 6.0);
    q.ex_decay_e = float(P.ex_decay_e);
    q.in_decay_e = float(P.in_decay_e);
    q.ex_decay_i = float(P.ex_decay_i);
    q.in_decay_i = float(P.in_decay_i);
    q.w_e_abs = float(P.w_e);
    q.w_i_abs = float(-P.w_i);
    q.w_n2a = float(P.w_n2a);
    q.w_a2n = float(P.w_a2n);
    q.w_poiss_astro = float(P.w_poiss_astro);
    q.w_poiss_exc = float(P.w_poiss_exc);
    q.w_poiss_inh = float(P.w_poiss_inh);
    q.noise_astro = float(P.noise_astro);
    q.noise_exc = float(P.noise_exc);
    q.noise_inh = float(P.noise_inh);
    return q;
}

#if defined(__HIPCC__)

__device__ __forceinline__ float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }

// The Box-Muller draw of rng_normal_z (src/astrosim_fused.cu:143-149)
// evaluated in FP32 for FP32 state: the same two uniforms rounded to float
// (u1 clamped at 2^-126), the hardware log2, and the hardware cosine, whose
// argument is in turns, so cos(2*pi*u2) needs no range reduction.
__device__ __forceinline__ float normal32(std::uint64_t z) {
    float u1 = static_cast<float>(rng_to_unit(rng_mix(z)));
    u1 = u1 < 0x1p-126f ? 0x1p-126f : u1;
    const float u2 = static_cast<float>(rng_to_unit(rng_mix(z + kRngCounterMul)));
    constexpr float kMinus2Ln2 = -1.38629436111989061883f;  // -2 ln 2: ln u = ln 2 * log2 u
    return __builtin_sqrtf(kMinus2Ln2 * __builtin_amdgcn_logf(u1)) * __builtin_amdgcn_cosf(u2);
}

// hi+lo += x, error-free (TwoSum) then renormalised (Fast2Sum). Every
// operation here must round on its own: this function is never contracted.
__device__ __forceinline__ void comp_add(float& hi, float& lo, float x) {
    const float s = hi + x;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (x - bp);
    const float l = lo + err;
    const float h2 = s + l;
    lo = l - (h2 - s);
    hi = h2;
}

// The AdEx spike current g_L * Delta_T * exp((V - V_T) / Delta_T) in FP64 for
// FP32 state, rounded to FP32 once. The factor g_L*Delta_T enters the
// argument as its logarithm, so x = V * x_scale + x_shift is one fused
// operation (x_scale = 1/Delta_T, x_shift = ln(g_L*Delta_T) - V_T/Delta_T);
// then k = round(x / ln2) by a rounding shifter (the low 32 bits of the sum's
// encoding), r = x - k*ln2 with |r| <= ln2/2, a degree-8 Taylor polynomial
// (truncation < 2.2e-10 relative: under 0.005 ulp of the FP32 result) and
// 2^k built in the exponent field. x is held at -700 or above, which keeps
// 2^k normal (the current is zero in FP32 long before).
__device__ __forceinline__ float spike_current32(const NeuronConst32& p, float Vc) {
    double x = __builtin_fma(static_cast<double>(Vc), p.x_scale_d, p.x_shift_d);
    x = __builtin_fmax(x, -700.0);
    const double t = __builtin_fma(x, 0x1.71547652b82fep+0, 0x1.8p52);
    const double kd = t - 0x1.8p52;
    const int k = static_cast<int>(static_cast<std::uint32_t>(__double_as_longlong(t)));
    const double r = __builtin_fma(-kd, 0x1.62e42fefa39efp-1, x);
    double q = 1.0 / 40320.0;
    q = __builtin_fma(r, q, 1.0 / 5040.0);
    q = __builtin_fma(r, q, 1.0 / 720.0);
    q = __builtin_fma(r, q, 1.0 / 120.0);
    q = __builtin_fma(r, q, 1.0 / 24.0);
    q = __builtin_fma(r, q, 1.0 / 6.0);
    q = __builtin_fma(r, q, 0.5);
    q = __builtin_fma(r, q, 1.0);
    q = __builtin_fma(r, q, 1.0);
    const double two_k = __longlong_as_double(static_cast<long long>(k + 1023) << 52);
    return static_cast<float>(q * two_k);
}

// Li-Rinzel right-hand side in FP32 (src/astrosim_fused.cu:236-253).
// The FP32 right-hand sides let the compiler fuse multiply-adds.
__device__ __forceinline__ void astro_rhs32(const AstroConst32& c, float Ca_tot, float IP3_0, float inv_tau,
                                            float Ca, float IP3, float h, float noise, float& dCa, float& dIP3,
                                            float& dh) {
#pragma clang fp contract(fast)
    const float Ca_ER = (Ca_tot - Ca) * c.inv_ratio_ER_cyt;
    const float m_inf = IP3 * rcp32(IP3 + c.Kd_IP3_1);
    const float n_inf = Ca * rcp32(Ca + c.Kd_act);
    const float m3 = m_inf * m_inf * m_inf;
    const float n3 = n_inf * n_inf * n_inf;
    const float h3 = h * h * h;
    const float J_channel = c.ratio_ER_cyt * c.rate_IP3R * m3 * n3 * h3 * (Ca_ER - Ca);
    const float Ca2 = Ca * Ca;
    const float J_pump = c.rate_SERCA * Ca2 * rcp32(c.Km_SERCA2 + Ca2);
    const float J_leak = c.ratio_ER_cyt * c.rate_L * (Ca_ER - Ca);
    const float alpha_h = c.k_IP3R * c.Kd_inh * (IP3 + c.Kd_IP3_1) * rcp32(IP3 + c.Kd_IP3_2);
    const float beta_h = c.k_IP3R * Ca;
    dCa = J_channel - J_pump + J_leak + noise;
    dIP3 = (IP3_0 - IP3) * inv_tau;
    dh = alpha_h * (1.0f - h) - beta_h * h;
}

// IP3 jump, RK4 substeps and clamps with compensated state
// (src/astrosim_fused.cu:255-278).
__device__ __forceinline__ void astro_advance32(const AstroConst32& c, const Params32& q, float Ca_tot,
                                                float IP3_0, float inv_tau, float delta_IP3, float ip3_input,
                                                float noise, int substeps, float& Ca, float& Ca_lo, float& IP3,
                                                float& IP3_lo, float& h, float& h_lo) {
#pragma clang fp contract(fast)
    if (ip3_input != 0.0f) comp_add(IP3, IP3_lo, delta_IP3 * ip3_input);
    const float hs = q.h_step, hh = q.h_half;
    for (int s = 0; s < substeps; ++s) {
        float k1C, k1I, k1h, k2C, k2I, k2h, k3C, k3I, k3h, k4C, k4I, k4h;
        astro_rhs32(c, Ca_tot, IP3_0, inv_tau, Ca, IP3, h, noise, k1C, k1I, k1h);
        astro_rhs32(c, Ca_tot, IP3_0, inv_tau, Ca + hh * k1C, IP3 + hh * k1I, h + hh * k1h, noise, k2C, k2I, k2h);
        astro_rhs32(c, Ca_tot, IP3_0, inv_tau, Ca + hh * k2C, IP3 + hh * k2I, h + hh * k2h, noise, k3C, k3I, k3h);
        astro_rhs32(c, Ca_tot, IP3_0, inv_tau, Ca + hs * k3C, IP3 + hs * k3I, h + hs * k3h, noise, k4C, k4I, k4h);
        comp_add(Ca, Ca_lo, q.h_sixth * (k1C + 2.0f * k2C + 2.0f * k3C + k4C));
        comp_add(IP3, IP3_lo, q.h_sixth * (k1I + 2.0f * k2I + 2.0f * k3I + k4I));
        comp_add(h, h_lo, q.h_sixth * (k1h + 2.0f * k2h + 2.0f * k3h + k4h));
        if (Ca < 0.0f) { Ca = 0.0f; Ca_lo = 0.0f; }
        else if (Ca > Ca_tot) { Ca = Ca_tot; Ca_lo = 0.0f; }
        if (h < 0.0f) { h = 0.0f; h_lo = 0.0f; }
        else if (h > 1.0f) { h = 1.0f; h_lo = 0.0f; }
        if (IP3 < 0.0f) { IP3 = 0.0f; IP3_lo = 0.0f; }
    }
}

__device__ __forceinline__ float sic_factor32(const AstroConst32& c, float Ca, float Ca_lo) {
    const float y = ((Ca - c.SIC_th) + Ca_lo) * 1000.0f;
    return (y <= 1.0f) ? 0.0f : c.SIC_scale * __logf(y);
}

// AdEx right-hand side in FP32 with an FP64 exponential
// (src/astrosim_fused.cu:281-290). FP64 is required within 3*Delta_T of V_T;
// it is taken for every V because a wave holds neurons on both sides of that
// band, and one FP64 path costs less than executing both the FP64 and the
// FP32 exponential per wave (measured: 15% of the FP32 step loop).
__device__ __forceinline__ void neuron_rhs32(const NeuronConst32& p, float V, float w, float g_ex, float g_in,
                                             float I_ext, float& dV, float& dw) {
#pragma clang fp contract(fast)
    const float Vc = V < p.V_peak ? V : p.V_peak;
    const float I_spike = spike_current32(p, Vc);
    dV = (-p.g_L * (Vc - p.E_L) + I_spike - g_ex * (Vc - p.E_ex) - g_in * (Vc - p.E_in) - w + p.I_e + I_ext) *
         p.inv_C_m;
    dw = (p.a * (Vc - p.E_L) - w) * p.inv_tau_w;
}

struct NeuronRegs32 {
    float V, w, g_ex, dg_ex, g_in, dg_in;
    int refractory;
};

// Substep loop with exact alpha-cascade propagator, refractory clamp, RK4,
// spike and reset (src/astrosim_fused.cu:514-553).
__device__ __forceinline__ bool neuron_substeps32(const NeuronConst32& p, const Params32& q, NeuronRegs32& r,
                                                  float I_ext, float V_reset, float b, int substeps,
                                                  float ex_decay, float in_decay) {
#pragma clang fp contract(fast)
    bool spiked = false;
    const float hs = q.h_step, hh = q.h_half;
    for (int sub = 0; sub < substeps; ++sub) {
        const float g_ex_next = ex_decay * (r.g_ex + hs * r.dg_ex);
        const float dg_ex_next = ex_decay * r.dg_ex;
        const float g_in_next = in_decay * (r.g_in + hs * r.dg_in);
        const float dg_in_next = in_decay * r.dg_in;
        if (r.refractory > 0) {
            r.V = V_reset;
            float dV, dw;
            neuron_rhs32(p, r.V, r.w, r.g_ex, r.g_in, I_ext, dV, dw);
            r.w += hs * dw;
        } else {
            const float g_ex_mid = 0.5f * (r.g_ex + g_ex_next);
            const float g_in_mid = 0.5f * (r.g_in + g_in_next);
            float k1V, k1w, k2V, k2w, k3V, k3w, k4V, k4w;
            neuron_rhs32(p, r.V, r.w, r.g_ex, r.g_in, I_ext, k1V, k1w);
            neuron_rhs32(p, r.V + hh * k1V, r.w + hh * k1w, g_ex_mid, g_in_mid, I_ext, k2V, k2w);
            neuron_rhs32(p, r.V + hh * k2V, r.w + hh * k2w, g_ex_mid, g_in_mid, I_ext, k3V, k3w);
            neuron_rhs32(p, r.V + hs * k3V, r.w + hs * k3w, g_ex_next, g_in_next, I_ext, k4V, k4w);
            r.V += q.h_sixth * (k1V + 2.0f * k2V + 2.0f * k3V + k4V);
            r.w += q.h_sixth * (k1w + 2.0f * k2w + 2.0f * k3w + k4w);
        }
        r.g_ex = g_ex_next;
        r.dg_ex = dg_ex_next;
        r.g_in = g_in_next;
        r.dg_in = dg_in_next;
        if (r.refractory > 0) {
            --r.refractory;
        } else if (r.V >= p.V_peak) {
            r.V = V_reset;
            r.w += b;
            r.refractory = p.refractory_steps;
            spiked = true;
        }
    }
    return spiked;
}

#endif  // __HIPCC__

}  // namespace astrosim
