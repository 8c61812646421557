"""Known-answer checks of the oracle's step arithmetic, through its per-cell
entry points (oracle/oracle.h), against what the mathematics fixes.

Every constant below is typed from the reference's Biology table and
constants (src/astrosim_fused.cu:744-769, 944-951, 1003-1005), not read from
the oracle, so a mistyped constant in the oracle fails here. Each check is
chosen so that a plausible transcription mistake fails it: a swapped Kd, a
square for a cube, a dropped term, RK4 stages evaluated at the wrong point or
weighted wrongly, the exact conductance midpoint in place of the reference's
average, a wrong propagator, a refractory branch that integrates V, an STP
update in the wrong order, a gather summed out of row order.
"""
import math
from fractions import Fraction as F

import numpy as np
import pytest

import pyoracle as O

# Li-Rinzel (src/astrosim_fused.cu:744-750)
CA_TOT, IP3_0, TAU_IP3, DELTA_IP3 = 1.8958264782990153, 0.010925583871830699, 1058.7589568400845, 0.04911592657233997
KD_IP3_1, KD_IP3_2, KD_ACT, KD_INH, KM_SERCA = 0.13, 0.9434, 0.08234, 1.049, 0.1
SIC_TH, K_IP3R, RATE_IP3R, RATE_L, RATE_SERCA, RATIO = 0.19669, 0.0002, 0.006, 0.00011, 0.0009, 0.185
# AdEx (src/astrosim_fused.cu:753-762): excitatory, inhibitory
EXC = dict(DT=2.0, a=4.0, tau_w=450.0, C_m=130.0, g_L=18.0, E_L=-58.0, V_T=-50.0, V_peak=0.0, E_ex=0.0,
           E_in=-85.0, tau_ex=0.2, tau_in=2.0)
INH = dict(DT=2.0, a=-0.8, tau_w=264.0, C_m=104.0, g_L=4.3, E_L=-65.0, V_T=-52.0, V_peak=0.0, E_ex=0.0,
           E_in=-85.0, tau_ex=0.2, tau_in=2.0)
# STP (src/astrosim_fused.cu:768): U, tau_rec; tau_fac = 0
STP_U, STP_TAU_REC = 0.5, 800.0


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def rhs(Ca, IP3, h, noise=0.0, Ca_tot=CA_TOT):
    return O.astro_rhs(Ca_tot, IP3_0, TAU_IP3, Ca, IP3, h, noise)


# ---------------------------------------------------------------- A6: Li-Rinzel right-hand side

def test_ip3_relaxes_to_its_resting_value():
    # dIP3 = (IP3_0 - IP3) / tau_IP3: zero exactly at IP3_0 whatever Ca and h
    for Ca, h in ((0.0, 0.0), (0.07, 0.79), (1.2, 1.0)):
        assert rhs(Ca, IP3_0, h)[1] == 0.0
    assert rel(rhs(0.1, 0.0, 0.5)[1], IP3_0 / TAU_IP3) < 1e-15
    assert rel(rhs(0.1, 3 * IP3_0, 0.5)[1], -2 * IP3_0 / TAU_IP3) < 1e-15


@pytest.mark.parametrize("noise", [0.0, 0.3, -2e-4])
def test_calcium_rate_without_calcium_or_ip3_is_the_leak(noise):
    # Ca = IP3 = 0: no channel flux (m_inf = 0), no pump (Ca^2 = 0); the leak
    # ratio*rate_L*(Ca_ER - 0) with Ca_ER = Ca_tot/ratio is rate_L*Ca_tot
    for h in (0.0, 0.5, 1.0):
        assert rel(rhs(0.0, 0.0, h, noise)[0], RATE_L * CA_TOT + noise) < 1e-14


@pytest.mark.parametrize("c", [0.05, 0.1, 0.3, 1.0])
def test_pump_and_leak_without_ip3(c):
    # IP3 = 0 removes the channel term for every h; what is left is
    # rate_L*(Ca_tot - c) - ratio*rate_L*c - rate_SERCA*c^2/(Km^2 + c^2)
    fc = F(c)
    want = (F(RATE_L) * (F(CA_TOT) - fc) - F(RATIO) * F(RATE_L) * fc
            - F(RATE_SERCA) * fc * fc / (F(KM_SERCA) ** 2 + fc * fc))
    for h in (0.0, 0.7):
        assert rel(rhs(c, 0.0, h)[0], float(want)) < 1e-12


def test_channel_flux_cubes_and_its_half_activation_constants():
    # At IP3 = Kd_IP3_1 and Ca = Kd_act, m_inf = n_inf = 1/2, so the channel
    # flux (the only h-dependent part of dCa) is
    # ratio*rate_IP3R*(1/8)(1/8)*h^3*(Ca_ER - Ca)
    ca = KD_ACT
    d1 = rhs(ca, KD_IP3_1, 1.0)[0] - rhs(ca, KD_IP3_1, 0.0)[0]
    want = F(RATE_IP3R) / 64 * (F(CA_TOT) - F(ca) - F(RATIO) * F(ca))
    assert rel(d1, float(want)) < 1e-9
    dh = rhs(ca, KD_IP3_1, 0.5)[0] - rhs(ca, KD_IP3_1, 0.0)[0]
    assert rel(dh, float(want) / 8) < 1e-9


def test_h_gate_rates():
    # alpha_h = k_IP3R*Kd_inh*(IP3 + Kd_IP3_1)/(IP3 + Kd_IP3_2), beta_h = k_IP3R*Ca:
    # at h = 0 and Ca = 0 only alpha_h acts; at h = 1 only -beta_h
    assert rel(rhs(0.0, 0.0, 0.0)[2], K_IP3R * KD_INH * KD_IP3_1 / KD_IP3_2) < 1e-15
    assert rel(rhs(0.0, 1.0, 0.0)[2], K_IP3R * KD_INH * (1.0 + KD_IP3_1) / (1.0 + KD_IP3_2)) < 1e-15
    for c in (0.05, 0.4):
        assert rel(rhs(c, 0.3, 1.0)[2], -K_IP3R * c) < 1e-15


def test_sic_factor_threshold_and_log():
    # SIC_scale*log((Ca - SIC_th)*1000) when that argument exceeds 1, else 0 (src:416-417)
    assert O.sic_factor(0.0) == 0.0
    assert O.sic_factor(SIC_TH) == 0.0
    assert O.sic_factor(SIC_TH + 0.0009) == 0.0
    assert abs(O.sic_factor(SIC_TH + 0.002) - math.log(2.0)) < 1e-12
    assert abs(O.sic_factor(SIC_TH + 0.001 * math.e ** 3) - 3.0) < 1e-11


# ---------------------------------------------------------------- A7: astro_advance

def advance(state, **kw):
    return O.astro_advance(CA_TOT, IP3_0, TAU_IP3, DELTA_IP3, state, **kw)


def test_astro_rk4_is_fourth_order():
    # no noise, no clamping: halving the step divides the error by ~16
    st0, T = [0.15, 0.6, 0.5], 2000.0
    ref = advance(st0, h_step=T / 8192, substeps=8192)
    errs = [np.max(np.abs(advance(st0, h_step=T / k, substeps=k) - ref) / np.abs(ref)) for k in (20, 40, 80)]
    for a, b in zip(errs, errs[1:]):
        assert 14.0 < a / b < 19.0, errs


def test_ip3_jump_precedes_the_step():
    # with a zero step the RK4 increments vanish: only the jump delta_IP3*input
    st = advance([0.1, 0.02, 0.6], ip3_input=3.0, h_step=0.0)
    assert st[0] == 0.1 and st[2] == 0.6
    assert st[1] == 0.02 + DELTA_IP3 * 3.0
    assert np.array_equal(advance([0.1, 0.02, 0.6], ip3_input=0.0, h_step=0.0), [0.1, 0.02, 0.6])


def test_astro_clamps():
    # Ca in [0, Ca_tot], h in [0, 1], IP3 >= 0 after each substep
    assert np.array_equal(advance([-0.1, -1.0, 1.5], h_step=0.0), [0.0, 0.0, 1.0])
    assert np.array_equal(advance([CA_TOT + 0.5, 0.3, -0.5], h_step=0.0), [CA_TOT, 0.3, 0.0])
    # the per-cell Ca_tot bounds calcium, not the mean
    st = O.astro_advance(1.0, IP3_0, TAU_IP3, DELTA_IP3, [1.2, 0.01, 0.5], 0.0, 0.0, 0.0, 1)
    assert st[0] == 1.0


# ---------------------------------------------------------------- A8: AdEx right-hand side

@pytest.mark.parametrize("exc,p", [(True, EXC), (False, INH)])
def test_adex_at_threshold_rest_and_peak(exc, p):
    # at V = V_T the spike current is g_L*Delta_T*exp(0)
    dV, dw = O.neuron_rhs(exc, p["V_T"], 0.0)
    assert rel(dV, (-p["g_L"] * (p["V_T"] - p["E_L"]) + p["g_L"] * p["DT"]) / p["C_m"]) < 1e-14
    assert rel(dw, p["a"] * (p["V_T"] - p["E_L"]) / p["tau_w"]) < 1e-14
    # adaptation relaxes with tau_w at V = E_L
    assert rel(O.neuron_rhs(exc, p["E_L"], 1.0)[1], -1.0 / p["tau_w"]) < 1e-15
    # two Delta_T above threshold the spike current is e^2 times larger
    V = p["V_T"] + 2 * p["DT"]
    dV = O.neuron_rhs(exc, V, 0.0)[0]
    assert rel(dV * p["C_m"] + p["g_L"] * (V - p["E_L"]), p["g_L"] * p["DT"] * math.e ** 2) < 1e-12
    # V is clamped at V_peak inside the right-hand side
    assert np.array_equal(O.neuron_rhs(exc, 7.0, 3.0, 0.2, 0.1, 5.0), O.neuron_rhs(exc, p["V_peak"], 3.0, 0.2, 0.1, 5.0))


@pytest.mark.parametrize("exc,p", [(True, EXC), (False, INH)])
def test_adex_synaptic_and_external_currents(exc, p):
    V = -60.0
    base = O.neuron_rhs(exc, V, 2.0)
    assert rel(O.neuron_rhs(exc, V, 2.0, g_ex=1.0)[0] - base[0], -(V - p["E_ex"]) / p["C_m"]) < 1e-12
    assert rel(O.neuron_rhs(exc, V, 2.0, g_in=1.0)[0] - base[0], -(V - p["E_in"]) / p["C_m"]) < 1e-12
    assert rel(O.neuron_rhs(exc, V, 2.0, I_ext=13.0)[0] - base[0], 13.0 / p["C_m"]) < 1e-12
    assert rel(O.neuron_rhs(exc, V, 3.0)[0] - base[0], -1.0 / p["C_m"]) < 1e-12  # -w


def test_adex_rk4_fourth_order_without_conductances():
    def run(k):
        st, _, sp = O.neuron_substeps(True, [-70.0, 10.0, 0, 0, 0, 0], I_ext=100.0, h_step=10.0 / k, substeps=k)
        assert not sp
        return st[0]
    ref = run(20480)
    e = [abs(run(k) - ref) for k in (20, 40, 80)]
    for a, b in zip(e, e[1:]):
        assert 14.0 < a / b < 19.0, e


def test_adex_conductance_midpoint_is_the_average():
    # The reference evaluates RK4's midpoint stages at 0.5*(g(t) + g(t+h))
    # (src:528-537), an O(h^2) approximation of g(t + h/2): with decaying
    # conductances the step is second order (the exact midpoint would make it
    # fourth order).
    def run(k):
        st, _, sp = O.neuron_substeps(True, [-70.0, 10.0, 0.5, 0.3, 0.2, 0.05], I_ext=100.0, h_step=10.0 / k,
                                      substeps=k)
        assert not sp
        return st[0]
    ref = run(20480)
    e = [abs(run(k) - ref) for k in (20, 40, 80)]
    for a, b in zip(e, e[1:]):
        assert 3.0 < a / b < 5.0, e


# ---------------------------------------------------------------- A9: alpha cascade

def test_psc_init_is_e_over_tau():
    assert rel(O.psc_init(True, False), math.e / 0.2) < 1e-15
    assert rel(O.psc_init(True, True), math.e / 2.0) < 1e-15
    assert rel(O.psc_init(False, False), math.e / 0.2) < 1e-15


@pytest.mark.parametrize("syn,tau", [("ex", 0.2), ("in", 2.0)])
def test_alpha_cascade_peaks_at_one_after_tau(syn, tau):
    # a unit jump dg = e/tau; the exact propagator g' = e^{-h/tau}(g + h dg),
    # dg' = e^{-h/tau} dg gives g(t) = (t/tau) e^{1 - t/tau}: 1 at t = tau, less
    # before and after (h = 0.1 ms: 2 steps for tau = 0.2, 20 for tau = 2)
    h, n = 0.1, int(round(tau / 0.1))
    st = [-70.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    gi = 2 if syn == "ex" else 4
    st[gi + 1] = math.e / tau
    gs = []
    for _ in range(n + 2):
        st, _, _ = O.neuron_substeps(True, st, h_step=h, substeps=1)
        gs.append(st[gi])
    assert abs(gs[n - 1] - 1.0) < 1e-13
    assert gs[n - 2] < gs[n - 1] and gs[n] < gs[n - 1]
    for k, g in enumerate(gs, start=1):
        t = k * h
        assert rel(g, (t / tau) * math.exp(1.0 - t / tau)) < 1e-12
    assert rel(st[gi + 1], math.e / tau * math.exp(-(n + 2) * h / tau)) < 1e-12


# ---------------------------------------------------------------- A10: spike, reset, refractory clamp

def test_spike_resets_v_and_adds_b():
    st0 = [-1.0, 20.0, 0.0, 0.0, 0.0, 0.0]
    s0, r0, sp0 = O.neuron_substeps(True, st0, I_ext=1e5, V_reset=-55.5, b=0.0)
    s1, r1, sp1 = O.neuron_substeps(True, st0, I_ext=1e5, V_reset=-55.5, b=300.0)
    assert sp0 and sp1
    assert s1[0] == -55.5
    assert s1[1] == s0[1] + 300.0
    assert r1 == 0  # t_ref = 0 (src:754, 758): int(0/h + 0.5)
    # below V_peak nothing resets
    s2, _, sp2 = O.neuron_substeps(True, [-65.0, 0.0, 0, 0, 0, 0], V_reset=-55.5, b=300.0)
    assert not sp2 and s2[0] != -55.5


def test_refractory_clamps_v_and_steps_w_by_euler():
    V_reset, h = -52.0, 0.1
    st, r, sp = O.neuron_substeps(True, [-40.0, 7.0, 0.3, 0.1, 0.2, 0.0], refractory=2, I_ext=1e5, V_reset=V_reset,
                                  b=300.0, h_step=h)
    assert not sp and r == 1
    assert st[0] == V_reset
    dw = (EXC["a"] * (V_reset - EXC["E_L"]) - 7.0) / EXC["tau_w"]
    assert rel(st[1], 7.0 + h * dw) < 1e-15
    # the conductances still follow the propagator
    assert rel(st[2], math.exp(-h / 0.2) * (0.3 + h * 0.1)) < 1e-14


# ---------------------------------------------------------------- A11: STP

def test_stp_first_spike_and_recursion():
    # from x = 1, u = U, t_last = -1e12 (src:1003-1005): the first spike
    # releases x*u = U and leaves x = 1 (e^{-huge} = 0); u stays U (tau_fac = 0)
    out, st = O.stp_on_spike([1.0, STP_U, -1e12], 5.0)
    assert out == STP_U and st[0] == 1.0 and st[1] == STP_U and st[2] == 5.0
    # the second, 100 ms later: x = 1 + (x - x*u - 1) e^{-100/800}
    out, st = O.stp_on_spike(st, 105.0)
    x = 1.0 - STP_U * math.exp(-100.0 / STP_TAU_REC)
    assert rel(st[0], x) < 1e-15 and rel(out, STP_U * x) < 1e-15
    # a regular train converges to the fixed point x* = (1 - E)/(1 - (1 - U) E)
    E = math.exp(-50.0 / STP_TAU_REC)
    t = 105.0
    for _ in range(400):
        t += 50.0
        out, st = O.stp_on_spike(st, t)
    assert rel(st[0], (1.0 - E) / (1.0 - (1.0 - STP_U) * E)) < 1e-12
    # x recovers with the release fraction u the synapse had before the spike,
    # then u is updated (src:571-572); from u = 0.2 the new u is U again
    out, st = O.stp_on_spike([0.9, 0.2, 0.0], 40.0)
    E = math.exp(-40.0 / STP_TAU_REC)
    assert rel(st[0], 1.0 + (0.9 - 0.9 * 0.2 - 1.0) * E) < 1e-15
    assert st[1] == STP_U and rel(out, st[0] * STP_U) < 1e-15


# ---------------------------------------------------------------- A12: gathers

def test_gather_sums_in_row_order_and_ignores_zero_terms():
    ring = np.zeros(10)
    ring[[1, 4, 7]] = [1e16, 1.0, 1.0]
    # row order 1, 4, 7: (1e16 + 1) + 1 = 1e16 (each +1 rounds away); a
    # different order would give 1e16 + 2
    rs, src = [0, 3], [1, 4, 7]
    assert O.gather(rs, src, ring, 0) == 1e16
    assert O.gather([0, 3], [4, 7, 1], ring, 0) == 1e16 + 2.0
    # zero-valued sources anywhere in the row change nothing, bit for bit
    for row in ([0, 1, 4, 7], [1, 0, 2, 4, 3, 7, 9], [1, 4, 5, 6, 7, 8]):
        assert O.gather([0, len(row)], row, ring, 0) == 1e16
    # rows of zero-valued sources sum to +0.0 in any order; an empty row too
    for row in ([0, 2, 3], [3, 2, 0], []):
        v = O.gather([0, len(row)], row, ring, 0)
        assert v == 0.0 and math.copysign(1.0, v) == 1.0
    # the row is selected by the target's row_start pair
    assert O.gather([0, 2, 5, 5], [0, 2, 1, 4, 7], ring, 1) == 1e16
    assert O.gather([0, 2, 5, 5], [0, 2, 1, 4, 7], ring, 2) == 0.0


# ---------------------------------------------------------------- A16: delays and ring depth

@pytest.mark.parametrize("dt,d,R", [(0.1, 10, 11), (0.05, 20, 21), (0.3, 3, 4), (0.15, 7, 8), (0.4, 3, 4),
                                    (1.0, 1, 2), (12.0, 1, 2)])
def test_delays_and_ring_depth(dt, d, R):
    # delay_to_steps = max(1, int(1.0/dt + 0.5)) (src:783-786); R = max(d) + 1
    # (src:957); the noise index advances every llround(10*dt/dt) = 10 steps
    net = O.OracleNet(5, 1)
    sim = O.OracleSim(net, dt=dt, threads=1)
    assert O.sim_params(sim) == {"d_e": d, "d_i": d, "d_a2n": d, "R": R, "steps_per_noise": 10}
