// as_model32.h -- per-cell arithmetic of the FP32-compensated precision.
//
// State and right-hand sides are FP32 with hardware exp/log/reciprocal. The
// astrocyte variables Ca, IP3 and h change by ~1e-6 of their value per step,
// below FP32 resolution over thousands of steps, so each is carried as an
// unevaluated hi+lo pair of floats and every increment is added with an
// error-free TwoSum and renormalised (about 48 significant bits). The AdEx
// spike current g_L*DeltaT*exp((V-V_T)/DeltaT) has an exponent that near
// threshold is a small difference of large terms: it is formed and split
// into integer and fraction in FP64, and the fraction's power of two is the
// hardware exponential (spike_current32: about 1 ulp of FP32). Poisson counts
// use the FP64 path's 64-bit draws and FP64 products (bit-identical counts);
// Gaussian noise uses the same two uniforms per draw with the Box-Muller
// transform in FP32 (normal32: hardware log2 and cosine), within FP32
// rounding of the FP64 draw. Formulas follow src/astrosim_fused.cu:236-290,
// 508-576.
#pragma once

#include <cmath>

#include "as_model.h"

namespace astrosim {

// Li-Rinzel constants for FP32 state, with the products of constants the
// right-hand side uses formed on the host.
struct AstroConst32 {
    float Kd_IP3_1, Kd_IP3_2, Kd_act, Km_SERCA2, k_IP3R, rate_SERCA, inv_ratio_ER_cyt;
    float c_channel;  // ratio_ER_cyt * rate_IP3R
    float c_leak;     // ratio_ER_cyt * rate_L
    float c_alpha;    // k_IP3R * Kd_inh
    float SIC_scale, SIC_th;
};

// AdEx constants of one population for FP32 state, in the forms the
// right-hand side uses: with C_m folded in, dV = B - A*V - w/C_m + I_spike/C_m
// where A = (g_L + g_ex + g_in)/C_m and B = (g_L*E_L + g_ex*E_ex + g_in*E_in +
// I_e + I_ext)/C_m, and dw = (a/tau_w)*V - a*E_L/tau_w - w/tau_w.
struct NeuronConst32 {
    float inv_C_m, gL_c, Eex_c, Ein_c, base_c;  // 1/C_m, g_L/C_m, E_ex/C_m, E_in/C_m, (g_L*E_L + I_e)/C_m
    float aw, dw0, inv_tau_w;                   // a/tau_w, -a*E_L/tau_w, 1/tau_w
    float V_peak, V_floor;                      // V clamp of the right-hand side (see spike_current32)
    float psc_init_ex, psc_init_in;
    int refractory_steps;                       // int(t_ref/h + 0.5)
    double y_scale_d, y_shift_d;  // spike-current exponent in base 2: V * y_scale_d + y_shift_d
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
    q.ac = {float(c.Kd_IP3_1), float(c.Kd_IP3_2), float(c.Kd_act), float(c.Km_SERCA * c.Km_SERCA),
            float(c.k_IP3R), float(c.rate_SERCA), float(1.0 / c.ratio_ER_cyt),
            float(c.ratio_ER_cyt * c.rate_IP3R), float(c.ratio_ER_cyt * c.rate_L), float(c.k_IP3R * c.Kd_inh),
            float(c.SIC_scale), float(c.SIC_th)};
    auto nc = [&](const NeuronConst& p) {
        NeuronConst32 n{};
        n.inv_C_m = float(1.0 / p.C_m);
        n.gL_c = float(p.g_L / p.C_m);
        n.Eex_c = float(p.E_ex / p.C_m);
        n.Ein_c = float(p.E_in / p.C_m);
        n.base_c = float((p.g_L * p.E_L + p.I_e) / p.C_m);
        n.aw = float(p.a / p.tau_w);
        n.dw0 = float(-p.a * p.E_L / p.tau_w);
        n.inv_tau_w = float(1.0 / p.tau_w);
        n.V_peak = float(p.V_peak);
        n.psc_init_ex = float(p.psc_init_ex);
        n.psc_init_in = float(p.psc_init_in);
        n.refractory_steps = static_cast<int>(p.t_ref / P.h_step + 0.5);
        // the spike current divided by C_m as a power of two: g_L*Delta_T/C_m
        // enters the exponent as its base-2 logarithm
        n.y_scale_d = M_LOG2E / p.Delta_T;
        n.y_shift_d = std::log2(p.g_L * p.Delta_T / p.C_m) - p.V_th * M_LOG2E / p.Delta_T;
        // 1000*Delta_T below threshold (2 V at Delta_T = 2 mV): never reached
        // by the dynamics; the exponent stays above -1500
        n.V_floor = float(p.V_th - 1000.0 * p.Delta_T);
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
// (u1 clamped at 2^-126), the hardware log2, square root and cosine (each
// within about 1 ulp; the cosine's argument is in turns, so cos(2*pi*u2)
// needs no range reduction).
__device__ __forceinline__ float normal32(std::uint64_t z) {
    float u1 = static_cast<float>(rng_to_unit(rng_mix(z)));
    u1 = u1 < 0x1p-126f ? 0x1p-126f : u1;
    const float u2 = static_cast<float>(rng_to_unit(rng_mix(z + kRngCounterMul)));
    constexpr float kMinus2Ln2 = -1.38629436111989061883f;  // -2 ln 2: ln u = ln 2 * log2 u
    return __builtin_amdgcn_sqrtf(kMinus2Ln2 * __builtin_amdgcn_logf(u1)) * __builtin_amdgcn_cosf(u2);
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

// The AdEx spike current divided by the capacitance,
// g_L * Delta_T / C_m * exp((V - V_T) / Delta_T), for FP32 state, as
// 2^y with y = V * log2(e)/Delta_T + log2(g_L*Delta_T/C_m) - V_T*log2(e)/Delta_T.
// Near threshold y is a small difference of two terms of about 36, so it is
// formed in FP64 (one fused operation on the exact FP32 V) and split there,
// exactly, into an integer n and f = y - n in [-1/2, 1/2] (a rounding
// shifter: n is in the low 32 bits of y + 1.5*2^52). f rounded to FP32 is
// off by at most 2^-26, 1.1e-8 relative in 2^f; 2^f is the hardware base-2
// exponential (v_exp_f32, within 1 ulp on [-1/2, 1/2]) and the scaling by
// 2^n is exact (v_ldexp_f32; it underflows to zero far below threshold).
// V >= V_floor keeps |y| < 2^31, where the shifter is exact.
__device__ __forceinline__ float spike_current32(const NeuronConst32& p, float Vc) {
    const double y = __builtin_fma(static_cast<double>(Vc), p.y_scale_d, p.y_shift_d);
    const double t = y + 0x1.8p52;
    const double nd = t - 0x1.8p52;
    const int n = static_cast<int>(static_cast<std::uint32_t>(__double_as_longlong(t)));
    const float f = static_cast<float>(y - nd);
    return __builtin_amdgcn_ldexpf(__builtin_amdgcn_exp2f(f), n);
}

// Li-Rinzel right-hand side in FP32 (src/astrosim_fused.cu:236-253),
// rearranged to fewer operations with the same terms:
//   J_channel + J_leak = (Ca_ER - Ca) * (c_channel * (m_inf*n_inf*h)^3 + c_leak)
//   m_inf * n_inf = IP3 * Ca / ((IP3 + Kd_IP3_1) * (Ca + Kd_act))  (one reciprocal)
//   dh = alpha_h * (1 - h) - beta_h * h = alpha_h - h * (alpha_h + beta_h).
// The FP32 right-hand sides let the compiler fuse multiply-adds.
__device__ __forceinline__ void astro_rhs32(const AstroConst32& c, float Ca_tot, float IP3_0, float inv_tau,
                                            float Ca, float IP3, float h, float noise, float& dCa, float& dIP3,
                                            float& dh) {
#pragma clang fp contract(fast)
    const float Ca_ER = (Ca_tot - Ca) * c.inv_ratio_ER_cyt;
    const float ip3_k1 = IP3 + c.Kd_IP3_1;
    const float mnh = IP3 * Ca * rcp32(ip3_k1 * (Ca + c.Kd_act)) * h;
    const float mnh3 = mnh * mnh * mnh;
    const float Ca2 = Ca * Ca;
    const float J_pump = c.rate_SERCA * Ca2 * rcp32(c.Km_SERCA2 + Ca2);
    const float alpha_h = c.c_alpha * ip3_k1 * rcp32(IP3 + c.Kd_IP3_2);
    dCa = (Ca_ER - Ca) * (c.c_channel * mnh3 + c.c_leak) - J_pump + noise;
    dIP3 = (IP3_0 - IP3) * inv_tau;
    dh = alpha_h - h * (alpha_h + c.k_IP3R * Ca);
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

// AdEx right-hand side in FP32 (src/astrosim_fused.cu:281-290), at a stage
// whose conductances give the coefficients A = (g_L + g_ex + g_in)/C_m and
// B = (g_L*E_L + g_ex*E_ex + g_in*E_in + I_e + I_ext)/C_m, with the spike
// current of spike_current32. V is clamped to [V_floor, V_peak] (the
// reference clamps at V_peak; V_floor is never reached).
__device__ __forceinline__ void neuron_rhs32(const NeuronConst32& p, float V, float w, float A, float B,
                                             float& dV, float& dw) {
#pragma clang fp contract(fast)
    const float Vc = __builtin_amdgcn_fmed3f(V, p.V_floor, p.V_peak);
    dV = __builtin_fmaf(-A, Vc, __builtin_fmaf(-p.inv_C_m, w, B)) + spike_current32(p, Vc);
    dw = __builtin_fmaf(p.aw, Vc, __builtin_fmaf(-p.inv_tau_w, w, p.dw0));
}

struct NeuronRegs32 {
    float V, w, g_ex, dg_ex, g_in, dg_in;
    int refractory;
};

// Substep loop with exact alpha-cascade propagator, refractory clamp, RK4,
// spike and reset (src/astrosim_fused.cu:514-553). The RK4 stages take the
// conductances at t, at the midpoint as 0.5*(g(t) + g(t+h)), and at t+h, as
// the reference; A and B are linear in them, so the midpoint's are the mean
// of the ends'.
__device__ __forceinline__ bool neuron_substeps32(const NeuronConst32& p, const Params32& q, NeuronRegs32& r,
                                                  float I_ext, float V_reset, float b, int substeps,
                                                  float ex_decay, float in_decay) {
#pragma clang fp contract(fast)
    bool spiked = false;
    const float hs = q.h_step, hh = q.h_half;
    const float K = __builtin_fmaf(I_ext, p.inv_C_m, p.base_c);
    for (int sub = 0; sub < substeps; ++sub) {
        const float g_ex_next = ex_decay * (r.g_ex + hs * r.dg_ex);
        const float dg_ex_next = ex_decay * r.dg_ex;
        const float g_in_next = in_decay * (r.g_in + hs * r.dg_in);
        const float dg_in_next = in_decay * r.dg_in;
        if (r.refractory > 0) {
            r.V = V_reset;
            const float Vc = __builtin_amdgcn_fmed3f(r.V, p.V_floor, p.V_peak);
            r.w += hs * __builtin_fmaf(p.aw, Vc, __builtin_fmaf(-p.inv_tau_w, r.w, p.dw0));
        } else {
            const float A0 = __builtin_fmaf(r.g_ex, p.inv_C_m, __builtin_fmaf(r.g_in, p.inv_C_m, p.gL_c));
            const float B0 = __builtin_fmaf(r.g_ex, p.Eex_c, __builtin_fmaf(r.g_in, p.Ein_c, K));
            const float A1 = __builtin_fmaf(g_ex_next, p.inv_C_m, __builtin_fmaf(g_in_next, p.inv_C_m, p.gL_c));
            const float B1 = __builtin_fmaf(g_ex_next, p.Eex_c, __builtin_fmaf(g_in_next, p.Ein_c, K));
            const float Am = 0.5f * (A0 + A1), Bm = 0.5f * (B0 + B1);
            float k1V, k1w, k2V, k2w, k3V, k3w, k4V, k4w;
            neuron_rhs32(p, r.V, r.w, A0, B0, k1V, k1w);
            neuron_rhs32(p, r.V + hh * k1V, r.w + hh * k1w, Am, Bm, k2V, k2w);
            neuron_rhs32(p, r.V + hh * k2V, r.w + hh * k2w, Am, Bm, k3V, k3w);
            neuron_rhs32(p, r.V + hs * k3V, r.w + hs * k3w, A1, B1, k4V, k4w);
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
