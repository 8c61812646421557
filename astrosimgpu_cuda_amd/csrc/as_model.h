// as_model.h -- model constants, scaling law, and the per-cell FP64 step
// arithmetic (Li-Rinzel astrocyte, AdEx neuron with alpha-conductance
// cascade, Tsodyks-Markram STP) shared by every FP64 kernel.
//
// The arithmetic is written so that, compiled with floating-point
// contraction off, it evaluates in exactly the order of the reference's
// expressions; the FP64 kernels therefore reproduce the host oracle bit for
// bit. Transcendentals come from as_dmath.h.
#pragma once

#include <cstdint>

#include "as_rng.h"

namespace astrosim {

using index_t = std::uint32_t;  // cell and edge ids (reference: src/astrosim_fused.cu:107)

// Fixed-in-degree scaling law: neurons, astrocytes and synapses all grow
// linearly with N_total. Reference: src/astrosim_fused.cu:595-615.
struct ScalingConfig {
    index_t n_total = 1000;
    index_t n_exc = 0, n_inh = 0, n_astro = 0;
    double p_primary = 0.0;
    static constexpr double kSyn = 80.0;
    static constexpr double kPThird = 0.2;
    static constexpr double kEFraction = 0.8;
    static constexpr double kAstroRatio = 5.0;
    static constexpr int kPoolSize = 5;
    static constexpr index_t kMaxAstroOutDegree = 10000;
    index_t n_neurons() const { return n_exc + n_inh; }
};
ScalingConfig make_scaling(index_t n_total);

// Biology and drive constants. Reference: src/astrosim_fused.cu:742-781.
struct Biology {
    // Li-Rinzel astrocyte.
    double Ca_tot = 1.8958264782990153, IP3_0 = 0.010925583871830699;
    double tau_IP3 = 1058.7589568400845, delta_IP3 = 0.04911592657233997;
    double SIC_scale = 1.0, Kd_IP3_1 = 0.13, Kd_IP3_2 = 0.9434, Kd_act = 0.08234, Kd_inh = 1.049;
    double Km_SERCA = 0.1, SIC_th = 0.19669, k_IP3R = 0.0002, rate_IP3R = 0.006, rate_L = 0.00011;
    double rate_SERCA = 0.0009, ratio_ER_cyt = 0.185, Ca_init = 0.073, h_init = 0.793;
    double IP3_init = 0.010925583871830699;
    bool randomize_astro = true;
    double ra_lower = 0.9, ra_upper = 1.05, ra_var = 0.1;
    // AdEx, excitatory population.
    double e_Delta_T = 2.0, e_a = 4.0, e_tau_w = 450.0, e_V_reset = -50.0, e_b = 300.0, e_V_m = -50.0;
    double e_C_m = 130.0, e_g_L = 18.0, e_E_L = -58.0, e_V_th = -50.0, e_V_peak = 0.0, e_t_ref = 0.0;
    double e_E_ex = 0.0, e_E_in = -85.0, e_tau_syn_ex = 0.2, e_tau_syn_in = 2.0;
    // AdEx, inhibitory population.
    double i_Delta_T = 2.0, i_a = -0.8, i_tau_w = 264.0, i_V_reset = -60.0, i_b = 130.0, i_V_m = -60.0;
    double i_C_m = 104.0, i_g_L = 4.3, i_E_L = -65.0, i_V_th = -52.0, i_V_peak = 0.0, i_t_ref = 0.0;
    double i_E_ex = 0.0, i_E_in = -85.0, i_tau_syn_ex = 0.2, i_tau_syn_in = 2.0;
    bool randomize_neuron = true;
    double rn_lower = 0.9, rn_upper = 1.1, rn_var = 0.05;
    // Synapses: one scalar weight and delay per connection type.
    double w_e = 5.0, w_i = -5.0, w_n2a = 0.2, w_a2n = 1.0;
    double d_e = 1.0, d_i = 1.0, d_a2n = 1.0;
    bool stp_enabled = true;
    double stp_U = 0.5, stp_tau_rec = 800.0, stp_tau_fac = 0.0;
    int sic_interval = 1;
    // Background drive (the *_noise_std values are standard deviations).
    double astro_poiss_rate = 3.029715502945856, astro_poiss_weight = 1.0, astro_noise_std = 0.0003;
    double exc_poiss_rate = 2700.0, exc_poiss_weight = 1.0, exc_noise_std = 100.0;
    double inh_poiss_rate = 2500.0, inh_poiss_weight = 1.0, inh_noise_std = 100.0;
    bool independent_noise = true;
    double noise_dt = 0.0;  // <= 0: resolved to 10*dt
};

// Delay in ms to whole steps, at least one. Reference: src/astrosim_fused.cu:783-786.
inline int delay_to_steps(double delay_ms, double dt) {
    const int steps = static_cast<int>(delay_ms / dt + 0.5);
    return steps < 1 ? 1 : steps;
}

// ----------------------------------------------------------------------------
// Correctly rounded division by a fixed positive divisor.
// ----------------------------------------------------------------------------
// The compiler's FP64 a / b is eleven instructions, one of them a
// quarter-rate reciprocal. For a divisor b > 0 that is reused across many
// divisions, four fused operations give the same correctly rounded quotient,
// every bit of it including the sign of a zero, from four precomputed values
//   y = RN(1/b),  yd = RD(1/b),  yld = RN(1/b - yd) >= 0,  ny = -y:
//   q0 = RN(a*yd + RN(a*yld))  is within half an ulp plus a relative 2^-104
//                              of a/b, so one of the two doubles around it;
//   rn = RN(b*q0 - a)          is exact, because q0 is within one ulp;
//   q  = RN(q0 + rn*ny)        is RN(a/b): with y within half an ulp of 1/b
//                              and q0 within one ulp of a/b, the exact
//                              q0 + (a - b*q0)*y lies nearer a/b than any
//                              rounding midpoint (Markstein's theorem, IBM
//                              J. Res. Dev. 34(1), 1990).
// For a = +-0 every intermediate is a zero of a's sign (yld >= 0, and the
// residual is formed as b*q0 - a), which is the sign of a/b for b > 0. ny is
// stored negated so that no fused result is negated: a compiler may fold
// such a negation into the fused operation, which is exact except for the
// sign of a zero. Valid while no step overflows or underflows (|a| between
// 2^-960 and 2^1000, or zero, and b normal).
struct FixedDivisor {
    double b, yd, yld, ny;
};

AS_HD FixedDivisor make_fixed_divisor(double b) {
    FixedDivisor d;
    const double y = 1.0 / b;
    double yd = y;
    double e = __builtin_fma(-b, y, 1.0);  // 1 - b*y, exact
    if (e < 0.0) {                         // y above 1/b: the double below it
        yd = as_dm_double(as_dm_bits(y) - 1);
        e = __builtin_fma(-b, yd, 1.0);
    }
    d.b = b;
    d.yd = yd;
    d.yld = e / b;
    d.ny = -y;
    return d;
}

// a / b, bit for bit, for the divisor d was made from.
AS_HD double divide(double a, const FixedDivisor& d) {
    const double q0 = __builtin_fma(a, d.yd, a * d.yld);
    const double rn = __builtin_fma(d.b, q0, -a);
    return __builtin_fma(rn, d.ny, q0);
}

// Li-Rinzel constants. Reference: src/astrosim_fused.cu:222-226.
struct AstroConst {
    double Kd_IP3_1, Kd_IP3_2, Kd_act, Kd_inh, Km_SERCA;
    double k_IP3R, rate_IP3R, rate_L, rate_SERCA, ratio_ER_cyt;
    double SIC_scale, SIC_th;
    FixedDivisor ratio_ER_cyt_div;
};

// AdEx constants of one population. Reference: src/astrosim_fused.cu:228-232.
struct NeuronConst {
    double C_m, g_L, E_L, V_th, Delta_T, a, tau_w;
    double V_peak, t_ref, E_ex, E_in, tau_syn_ex, tau_syn_in, I_e;
    double psc_init_ex, psc_init_in;  // e / tau_syn
    FixedDivisor Delta_T_div, C_m_div, tau_w_div;
    int refractory_steps;  // (int)(t_ref / h_step + 0.5), src/astrosim_fused.cu:549
};

AstroConst make_astro_const(const Biology& bio);
NeuronConst make_neuron_const(const Biology& bio, bool excitatory, double h_step);

// ----------------------------------------------------------------------------
// Per-cell FP64 arithmetic.
// ----------------------------------------------------------------------------

// Li-Rinzel right-hand side. Reference: src/astrosim_fused.cu:236-253.
// tau_IP3 is the cell's divisor made from its tau_IP3.
AS_HD void astro_rhs(const AstroConst& c, double Ca_tot, double IP3_0, const FixedDivisor& tau_IP3,
                     double Ca, double IP3, double h, double noise, double& dCa, double& dIP3,
                     double& dh) {
    const double Ca_ER = divide(Ca_tot - Ca, c.ratio_ER_cyt_div);
    const double m_inf = IP3 / (IP3 + c.Kd_IP3_1);
    const double n_inf = Ca / (Ca + c.Kd_act);
    const double m3 = m_inf * m_inf * m_inf;
    const double n3 = n_inf * n_inf * n_inf;
    const double h3 = h * h * h;
    const double J_channel = c.ratio_ER_cyt * c.rate_IP3R * m3 * n3 * h3 * (Ca_ER - Ca);
    const double J_pump = c.rate_SERCA * Ca * Ca / (c.Km_SERCA * c.Km_SERCA + Ca * Ca);
    const double J_leak = c.ratio_ER_cyt * c.rate_L * (Ca_ER - Ca);
    const double alpha_h = c.k_IP3R * c.Kd_inh * (IP3 + c.Kd_IP3_1) / (IP3 + c.Kd_IP3_2);
    const double beta_h = c.k_IP3R * Ca;
    dCa = J_channel - J_pump + J_leak + noise;
    dIP3 = divide(IP3_0 - IP3, tau_IP3);
    dh = alpha_h * (1.0 - h) - beta_h * h;
}

// IP3 jump, then `substeps` classic RK4 steps, then clamps. h_sixth is
// h_step / 6.0. Reference: src/astrosim_fused.cu:255-278.
AS_HD void astro_advance(const AstroConst& c, double Ca_tot, double IP3_0, const FixedDivisor& tau_IP3,
                         double delta_IP3, double ip3_input, double noise, double h_step,
                         double h_sixth, int substeps, double& Ca, double& IP3, double& h) {
    if (ip3_input != 0.0) IP3 += delta_IP3 * ip3_input;
    for (int s = 0; s < substeps; ++s) {
        double k1C, k1I, k1h, k2C, k2I, k2h, k3C, k3I, k3h, k4C, k4I, k4h;
        astro_rhs(c, Ca_tot, IP3_0, tau_IP3, Ca, IP3, h, noise, k1C, k1I, k1h);
        astro_rhs(c, Ca_tot, IP3_0, tau_IP3, Ca + 0.5 * h_step * k1C, IP3 + 0.5 * h_step * k1I,
                  h + 0.5 * h_step * k1h, noise, k2C, k2I, k2h);
        astro_rhs(c, Ca_tot, IP3_0, tau_IP3, Ca + 0.5 * h_step * k2C, IP3 + 0.5 * h_step * k2I,
                  h + 0.5 * h_step * k2h, noise, k3C, k3I, k3h);
        astro_rhs(c, Ca_tot, IP3_0, tau_IP3, Ca + h_step * k3C, IP3 + h_step * k3I,
                  h + h_step * k3h, noise, k4C, k4I, k4h);
        Ca += h_sixth * (k1C + 2.0 * k2C + 2.0 * k3C + k4C);
        IP3 += h_sixth * (k1I + 2.0 * k2I + 2.0 * k3I + k4I);
        h += h_sixth * (k1h + 2.0 * k2h + 2.0 * k3h + k4h);
        Ca = Ca < 0.0 ? 0.0 : (Ca > Ca_tot ? Ca_tot : Ca);
        h = h < 0.0 ? 0.0 : (h > 1.0 ? 1.0 : h);
        IP3 = IP3 < 0.0 ? 0.0 : IP3;
    }
}

// SIC factor of an astrocyte: SIC_scale*log((Ca-SIC_th)*1000) above 1, else 0.
// Reference: src/astrosim_fused.cu:416-417.
AS_HD double sic_factor(const AstroConst& c, double Ca) {
    const double y = (Ca - c.SIC_th) * 1000.0;
    return (y <= 1.0) ? 0.0 : c.SIC_scale * as_log(y);
}

// AdEx right-hand side with V clamped at V_peak. Reference: src/astrosim_fused.cu:281-290.
AS_HD void neuron_rhs(const NeuronConst& p, double V, double w, double g_ex, double g_in,
                      double I_ext, double& dV, double& dw) {
    const double Vc = V < p.V_peak ? V : p.V_peak;
    // x is never NaN (the clamp above maps a NaN V to V_peak) and at most
    // (V_peak - V_th) / Delta_T, which make_neuron_const holds to <= 709, so
    // the clamped exp returns as_exp's bits.
    const double x = divide(Vc - p.V_th, p.Delta_T_div);
    const double I_spike = p.g_L * p.Delta_T * as_exp_upto709(x);
    const double I_syn_ex = g_ex * (Vc - p.E_ex);
    const double I_syn_in = g_in * (Vc - p.E_in);
    dV = divide(-p.g_L * (Vc - p.E_L) + I_spike - I_syn_ex - I_syn_in - w + p.I_e + I_ext, p.C_m_div);
    dw = divide(p.a * (Vc - p.E_L) - w, p.tau_w_div);
}

// Registers of one neuron across a step.
struct NeuronRegs {
    double V, w, g_ex, dg_ex, g_in, dg_in;
    int refractory;
};

// A neuron's reset potential and spike-triggered adaptation increment.
struct ResetPair {
    double V_reset, b;
};

// One communication step of one neuron after its inputs are known: the
// substep loop with the exact alpha-cascade propagator, the refractory clamp
// (Euler on w only), RK4 on (V, w), spike detection and reset.
// `ex_decay`/`in_decay` are as_exp(-h_step/tau_syn_*) (population constants),
// h_sixth is h_step / 6.0. `reset()` returns the neuron's ResetPair; it is
// called only on the refractory clamp and on a spike, the two branches that
// use them, so a caller can load them lazily.
// Returns true if the neuron spiked. Reference: src/astrosim_fused.cu:508-553.
template <class Reset>
AS_HD bool neuron_substeps_lazy(const NeuronConst& p, NeuronRegs& r, double I_ext, Reset&& reset,
                                double h_step, double h_sixth, int substeps, double ex_decay,
                                double in_decay) {
    bool spiked = false;
    for (int sub = 0; sub < substeps; ++sub) {
        const double g_ex_next = ex_decay * (r.g_ex + h_step * r.dg_ex);
        const double dg_ex_next = ex_decay * r.dg_ex;
        const double g_in_next = in_decay * (r.g_in + h_step * r.dg_in);
        const double dg_in_next = in_decay * r.dg_in;
        if (r.refractory > 0) {
            r.V = reset().V_reset;
            double dV, dw;
            neuron_rhs(p, r.V, r.w, r.g_ex, r.g_in, I_ext, dV, dw);
            r.w += h_step * dw;
        } else {
            const double g_ex_mid = 0.5 * (r.g_ex + g_ex_next);
            const double g_in_mid = 0.5 * (r.g_in + g_in_next);
            double k1V, k1w, k2V, k2w, k3V, k3w, k4V, k4w;
            neuron_rhs(p, r.V, r.w, r.g_ex, r.g_in, I_ext, k1V, k1w);
            neuron_rhs(p, r.V + 0.5 * h_step * k1V, r.w + 0.5 * h_step * k1w, g_ex_mid, g_in_mid,
                       I_ext, k2V, k2w);
            neuron_rhs(p, r.V + 0.5 * h_step * k2V, r.w + 0.5 * h_step * k2w, g_ex_mid, g_in_mid,
                       I_ext, k3V, k3w);
            neuron_rhs(p, r.V + h_step * k3V, r.w + h_step * k3w, g_ex_next, g_in_next, I_ext,
                       k4V, k4w);
            r.V += h_sixth * (k1V + 2.0 * k2V + 2.0 * k3V + k4V);
            r.w += h_sixth * (k1w + 2.0 * k2w + 2.0 * k3w + k4w);
        }
        r.g_ex = g_ex_next;
        r.dg_ex = dg_ex_next;
        r.g_in = g_in_next;
        r.dg_in = dg_in_next;
        if (r.refractory > 0) {
            --r.refractory;
        } else if (r.V >= p.V_peak) {
            const ResetPair rb = reset();
            r.V = rb.V_reset;
            r.w += rb.b;
            r.refractory = p.refractory_steps;
            spiked = true;
        }
    }
    return spiked;
}

// The same with the reset values passed in.
AS_HD bool neuron_substeps(const NeuronConst& p, NeuronRegs& r, double I_ext, double V_reset,
                           double b, double h_step, double h_sixth, int substeps, double ex_decay,
                           double in_decay) {
    return neuron_substeps_lazy(
        p, r, I_ext, [&]() { return ResetPair{V_reset, b}; }, h_step, h_sixth, substeps, ex_decay,
        in_decay);
}

// Per-presynaptic-neuron Tsodyks-Markram update on a spike at t_now; returns
// the release multiplier x*u carried by every outgoing synapse. (Plain
// divisions: this runs on the rare spike path, where the fixed-divisor
// constants would cost more scalar registers than the divisions cost time.)
// Reference: src/astrosim_fused.cu:565-576.
AS_HD double stp_on_spike(double& x, double& u, double& t_last, double t_now, double U,
                          double tau_rec, double tau_fac) {
    const double elapsed = t_now - t_last;
    const double x_decay = as_exp(-elapsed / tau_rec);
    const double u_decay = tau_fac < 1e-10 ? 0.0 : as_exp(-elapsed / tau_fac);
    x = 1.0 + (x - x * u - 1.0) * x_decay;
    u = U + u * (1.0 - U) * u_decay;
    t_last = t_now;
    return x * u;
}

}  // namespace astrosim
