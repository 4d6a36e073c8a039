// Hand-written adjoint (vector-Jacobian product) of one AnisoMPM substep.
// Included by pmpm.hip after the forward kernels; uses StepArgs, the bucket
// schedule and the weight helpers defined there.
//
// Given the adjoints of the step's outputs (x', v', F', C', d', lap', D'),
// the backward step produces the adjoints of its inputs (x, v, F, C, d, lap,
// D_hist) and accumulates those of the per-particle parameters mu, lam, A0,
// sigma_c.
// The grid of the step is recomputed first (forward binning, P2G and grid
// pass, with m and sum w exported), then:
//   k_adj_scatter  G2P adjoint: v_bar and gd_bar onto the grid, one
//                  wavefront per bucket, lane per cell, LDS frames, no atomics
//   k_adj_grid     grid adjoint: node v_bar, gd_bar -> (mv_bar, m_bar,
//                  sum_wd_bar, sum_w_bar)
//   k_adj_gather   per particle (sorted order): P2G adjoint, G2P position
//                  terms, affine, advection, particle BC -> tau_bar, d_new bar
//                  (the affine matrix and d_new come from the recompute's
//                  P2G records)
//   k_adj_const    per particle: constitutive VJP and the F update
// Supported: the strain limit (f_clamp), device colliders (sticky and
// slip with friction), velocity BCs, damage_every = 1; not the directional
// (interface) stress.
// Subgradients follow torch's conventions so the result matches autograd of
// the reference step: relu'(0) = 0, clamp passes the gradient on the closed
// interval, maximum splits it 1/2 : 1/2 at ties.
#pragma once

// a[k][axis] for a uniform runtime k, by selects (a dynamic index into a
// private array would put it in scratch memory)
template <int N, typename T> __device__ __forceinline__ T pick(const T (&a)[N][3], int k, int axis) {
    T r = a[0][axis];
#pragma unroll
    for (int j = 1; j < N; ++j) r = k == j ? a[j][axis] : r;
    return r;
}

template <typename T> struct AdjArgs {
    // upstream adjoints (of the step's outputs) and the step's outputs needed
    const T *gx, *gv, *gF, *gC, *gd, *glap, *gD;
    const T *nv, *nC;                   // v', C' of the forward step
    // adjoints written (of the step's inputs) and parameter adjoints (+=)
    T *ox, *ov, *oF, *oC, *od, *olap, *oD, *omu, *olam, *oA0, *osc;
    // grid: forward m, sum w (node-major pairs) and node adjoints (6 per node)
    const T* gmw;
    T* gadj;
    T* pbar;                            // per particle: tau_bar (9), d_new bar from the cubic scatter (1)
    int dmg_on;                         // the forward step ran the damage update
};

template <typename T> __device__ __forceinline__ T ddot3(const M3<T>& a, const M3<T>& b) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) s += a.m[i][j] * b.m[i][j];
    return s;
}
template <typename T> __device__ __forceinline__ M3<T> transpose(const M3<T>& a) {
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}

// torch.maximum backward: the larger input takes the gradient, ties split it.
template <typename T> __device__ __forceinline__ void max_grad(T a, T b, T g, T& ga, T& gb) {
    if (a > b) { ga = g; gb = T(0); }
    else if (a < b) { ga = T(0); gb = g; }
    else { ga = T(0.5) * g; gb = T(0.5) * g; }
}

// Adjoint of collider_resolve (anisompm.py:811-825) at input velocity v:
// g holds the adjoint of the resolved velocity on entry and that of v on
// exit (the collider velocity and normal do not depend on the state).
// torch's conventions: clamp(max=0) of the normal part and clamp(min=0) of
// the friction factor pass the gradient on their closed intervals; the norm
// of a zero tangential part passes none.
template <typename T>
__device__ __forceinline__ void collider_resolve_vjp(const T v[3], const T vc[3], const T n[3], int mode, T friction,
                                                     T g[3]) {
    if (mode == 0) {
        g[0] = g[1] = g[2] = T(0);      // sticky: the result is the collider velocity
        return;
    }
    const T vr[3] = {v[0] - vc[0], v[1] - vc[1], v[2] - vc[2]};
    const T vn = vr[0] * n[0] + vr[1] * n[1] + vr[2] * n[2];
    const T inward = minp(vn, T(0));
    const T vt[3] = {vr[0] - vn * n[0], vr[1] - vn * n[1], vr[2] - vn * n[2]};
    // v' = vc + vt' + (vn - inward) n
    const T keep_bar = g[0] * n[0] + g[1] * n[1] + g[2] * n[2];
    T vn_bar = keep_bar, inward_bar = -keep_bar;
    T vt_bar[3];
    if (friction > T(0)) {
        const T mag = sqrt(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
        const T den = mag + T(1e-9);
        const T raw = T(1) + friction * inward / den;
        const T sc = maxp(raw, T(0));
        const T sc_bar = g[0] * vt[0] + g[1] * vt[1] + g[2] * vt[2];
#pragma unroll
        for (int a = 0; a < 3; ++a) vt_bar[a] = g[a] * sc;
        const T raw_bar = raw >= T(0) ? sc_bar : T(0);
        inward_bar += raw_bar * friction / den;
        const T mag_bar = -raw_bar * friction * inward / (den * den);
        if (mag > T(0)) {
#pragma unroll
            for (int a = 0; a < 3; ++a) vt_bar[a] += mag_bar * vt[a] / mag;
        }
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) vt_bar[a] = g[a];
    }
    if (vn <= T(0)) vn_bar += inward_bar;
    vn_bar -= vt_bar[0] * n[0] + vt_bar[1] * n[1] + vt_bar[2] * n[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) g[a] = vt_bar[a] + vn_bar * n[a];
}

// Constitutive VJP of one particle (no f_clamp, no interface).  Recomputes
// the forward of anisompm.py:614-691 in registers, then back-propagates
// tau_bar (degraded Kirchhoff stress), dn_bar (post-update damage) and
// Dn_bar (post-update history drive) to F, d, lap, D_hist, mu, lam, A0.
// With the strain limit, F is the step's input and Fk the limited F the
// stress and the F update use; Fk_bar_in is the adjoint reaching Fk from the
// F update, and F_bar comes out with respect to the input F.
template <typename T>
__device__ void constitutive_vjp(const StepArgs<T>& A, long long p, const M3<T>& F, bool dmg_on,
                                 const M3<T>& tau_bar, T dn_bar, T Dn_bar, const M3<T>& Fk_bar_in,
                                 M3<T>& F_bar, T& d_bar, T& lap_bar, T& Dh_bar, T& mu_bar, T& lam_bar,
                                 M3<T>& A0_bar, T& sc_bar) {
    // ---- forward
    M3<T> Cm = mul_at(F, F);
    T lc[3];
    M3<T> V;
    sym_eig3<T, true>(Cm, lc, V);
    T S[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) S[i] = sqrt(maxp(lc[i], T(1e-12)));
    M3<T> W = mul(F, V), Wh;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const T inv = T(1) / (sqrt(W.m[0][j] * W.m[0][j] + W.m[1][j] * W.m[1][j] + W.m[2][j] * W.m[2][j]) + T(1e-12));
#pragma unroll
        for (int i = 0; i < 3; ++i) Wh.m[i][j] = W.m[i][j] * inv;
    }
    const M3<T> R = mul_bt(Wh, V);
    // the strain limit (anisompm.py:622-636): Fk = F V diag(clamp(S)/S) V^T, Sk = clamp(S)
    M3<T> Fk = F;
    T Sk[3] = {S[0], S[1], S[2]};
    if (A.use_fclamp) fclamp_apply(A, V, Fk, Sk);
    const T J0 = Sk[0] * Sk[1] * Sk[2];
    const T J = maxp(J0, T(1e-6));
    const T mu = A.mu[p], lam = A.lam[p];
    M3<T> FmR;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) FmR.m[i][j] = Fk.m[i][j] - R.m[i][j];
    const M3<T> P0 = mul_bt(FmR, Fk);      // (Fk - R) Fk^T
    M3<T> tdev;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) tdev.m[i][j] = T(2) * mu * P0.m[i][j];
    const T tvol = lam * (J - T(1)) * J;
    const T d = A.d[p], Dh = A.D_hist[p];
    T dn = d, Dn = Dh;
    // damage intermediates
    M3<T> sig, Q, sp, Am, A0m;
    T ev[3] = {T(0), T(0), T(0)}, fp[3] = {T(0), T(0), T(0)};
    T phi = T(0), Dz = T(0), u = T(0), y = T(0), dtr = T(0), sc = T(1), l0 = T(0), kd = T(0);
    if (dmg_on) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) sig.m[i][j] = (tdev.m[i][j] + (i == j ? tvol : T(0))) / J;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = i + 1; j < 3; ++j) {
                const T s = T(0.5) * (sig.m[i][j] + sig.m[j][i]);
                sig.m[i][j] = s; sig.m[j][i] = s;
            }
        sym_eig3<T, true>(sig, ev, Q);
        M3<T> QD;
#pragma unroll
        for (int i = 0; i < 3; ++i) fp[i] = maxp(ev[i], T(0));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) QD.m[i][j] = Q.m[i][j] * fp[j];
        sp = mul_bt(QD, Q);
        A0m = load_m3(A.A0, p);
        Am = mul_bt(mul(R, A0m), R);
        const M3<T> M = mul(Am, sp);
        T tr = T(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) tr += M.m[i][0] * M.m[0][i] + M.m[i][1] * M.m[1][i] + M.m[i][2] * M.m[2][i];
        sc = A.sigma_c[p];
        phi = tr / (sc * sc);
        Dz = A.zeta[p] * relu(phi - T(1));
        Dn = maxp(Dh, Dz);
        l0 = A.l0[p];
        u = (T(1) - d) * Dn - (d - (l0 * l0) * A.lap[p]);
        kd = A.dt_dmg / A.eta[p];
        y = d + kd * relu(u);
        dtr = minp(maxp(y, T(0)), T(1));
        dn = A.allow[p] ? maxp(dtr, d) : d;
    }
    const T res = A.residual[p];
    const T g = (T(1) - dn) * (T(1) - dn) * (T(1) - res) + res;
    const bool tens = J >= T(1);
    const T vdeg = tens ? g : T(1);

    // ---- backward
    const T trb = tau_bar.m[0][0] + tau_bar.m[1][1] + tau_bar.m[2][2];
    M3<T> tdev_bar;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) tdev_bar.m[i][j] = g * tau_bar.m[i][j];
    T tvol_bar = vdeg * trb;
    const T g_bar = ddot3(tau_bar, tdev) + (tens ? trb * tvol : T(0));
    T dn_tot = dn_bar + g_bar * (T(-2) * (T(1) - dn) * (T(1) - res));
    d_bar = T(0); lap_bar = T(0); Dh_bar = T(0); sc_bar = T(0);
    T J_bar = T(0);
    M3<T> R_bar;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) { R_bar.m[i][j] = T(0); A0_bar.m[i][j] = T(0); }
    if (dmg_on) {
        T dtr_bar = T(0);
        if (A.allow[p]) {
            T gd_;
            max_grad(dtr, d, dn_tot, dtr_bar, gd_);
            d_bar += gd_;
        } else {
            d_bar += dn_tot;
        }
        const T y_bar = (y >= T(0) && y <= T(1)) ? dtr_bar : T(0);
        d_bar += y_bar;
        const T u_bar = u > T(0) ? y_bar * kd : T(0);
        d_bar += u_bar * (-Dn - T(1));
        const T Dn_tot = Dn_bar + u_bar * (T(1) - d);
        lap_bar += u_bar * (l0 * l0);
        T Dz_bar;
        max_grad(Dh, Dz, Dn_tot, Dh_bar, Dz_bar);
        const T phi_bar = (phi - T(1)) > T(0) ? A.zeta[p] * Dz_bar : T(0);
        const T k2 = phi_bar * T(2) / (sc * sc);
        sc_bar = phi_bar * (T(-2) * phi / sc);          // phi = tr(...) / sc^2
        // phi = tr(A s+ A s+)/sc^2:  d/ds+ = 2 A s+ A,  d/dA = 2 s+ A s+
        const M3<T> ASA = mul(mul(Am, sp), Am), SAS = mul(mul(sp, Am), sp);
        M3<T> sp_bar, Am_bar;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                sp_bar.m[i][j] = k2 * ASA.m[j][i];
                Am_bar.m[i][j] = k2 * SAS.m[j][i];
            }
        // A = R A0 R^T
        const M3<T> t1 = mul(mul(Am_bar, R), transpose(A0m));
        const M3<T> t2 = mul(mul(transpose(Am_bar), R), A0m);
        const M3<T> a0b = mul(mul_at(R, Am_bar), R);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                R_bar.m[i][j] += t1.m[i][j] + t2.m[i][j];
                A0_bar.m[i][j] = a0b.m[i][j];
            }
        // s+ = Q diag(max(ev,0)) Q^T: Daleckii-Krein divided differences,
        // f'(ev) = [ev >= 0] at ties (torch clamp(min=0) passes ev == 0)
        M3<T> spb_s;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) spb_s.m[i][j] = T(0.5) * (sp_bar.m[i][j] + sp_bar.m[j][i]);
        M3<T> Sb = mul(mul_at(Q, spb_s), Q);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                T K;
                if (ev[i] != ev[j]) K = (fp[i] - fp[j]) / (ev[i] - ev[j]);
                else K = ev[i] >= T(0) ? T(1) : T(0);
                Sb.m[i][j] *= K;
            }
        const M3<T> sig_bar = mul_bt(mul(Q, Sb), Q);
        // sig = sym(tfull) / J
        T sgs = T(0);
        M3<T> tfull_bar;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const T sb = T(0.5) * (sig_bar.m[i][j] + sig_bar.m[j][i]);
                tfull_bar.m[i][j] = sb / J;
                sgs += sig_bar.m[i][j] * sig.m[i][j];
            }
        J_bar -= sgs / J;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) tdev_bar.m[i][j] += tfull_bar.m[i][j];
        tvol_bar += tfull_bar.m[0][0] + tfull_bar.m[1][1] + tfull_bar.m[2][2];
    } else {
        d_bar += dn_tot;
        Dh_bar += Dn_bar;
    }
    // tvol = lam (J - 1) J
    lam_bar = tvol_bar * (J - T(1)) * J;
    J_bar += tvol_bar * lam * (T(2) * J - T(1));
    // tdev = 2 mu (Fk - R) Fk^T
    mu_bar = T(2) * ddot3(tdev_bar, P0);
    const M3<T> a1 = mul(tdev_bar, Fk), a2 = mul_at(tdev_bar, FmR);
    M3<T> Fk_bar;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Fk_bar.m[i][j] = T(2) * mu * (a1.m[i][j] + a2.m[i][j]) + Fk_bar_in.m[i][j];
            R_bar.m[i][j] -= T(2) * mu * a1.m[i][j];
            F_bar.m[i][j] = T(0);
        }
    // J = max(Sk0 Sk1 Sk2, 1e-6), Sk_i = clamp(S_i) (or S_i), S_i = sqrt(max(lambda_i(F^T F), 1e-12)):
    // lambda_bar_i = J_bar (prod of the other Sk) [S_i in the clamp range] / (2 S_i), F_bar += 2 F V diag(lambda_bar) V^T
    const T J0_bar = J0 >= T(1e-6) ? J_bar : T(0);
    if (J0_bar != T(0)) {
        T lb[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const T others = Sk[(i + 1) % 3] * Sk[(i + 2) % 3];
            const bool pass = !A.use_fclamp || (S[i] >= A.fc_lo && S[i] <= A.fc_hi);
            lb[i] = (pass && lc[i] >= T(1e-12)) ? J0_bar * others / (T(2) * S[i]) : T(0);
        }
        M3<T> VD;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) VD.m[i][j] = V.m[i][j] * lb[j];
        const M3<T> dJ = mul(F, mul_bt(VD, V));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) F_bar.m[i][j] += T(2) * dJ.m[i][j];
    }
    // polar rotation (of the input F): dR = R Omega, (V^T Omega V)_ij = (X_ij - X_ji)/(s_i + s_j),
    // X = V^T R^T dF V  =>  F_bar += R V H V^T, H_ij = (G_ij - G_ji)/(s_i + s_j), G = V^T R^T R_bar V
    const M3<T> Gp = mul(mul_at(V, mul_at(R, R_bar)), V);
    M3<T> H;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const T den = S[i] + S[j];
            H.m[i][j] = den > T(0) ? (Gp.m[i][j] - Gp.m[j][i]) / den : T(0);
        }
    const M3<T> fr = mul_bt(mul(mul(R, V), H), V);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F_bar.m[i][j] += fr.m[i][j];
    if (!A.use_fclamp) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) F_bar.m[i][j] += Fk_bar.m[i][j];
        return;
    }
    // Fk = F Hk, Hk = g(F^T F) = V diag(g(lambda)) V^T, g(lambda) = clamp(S)/max(S, 1e-12):
    // F_bar += Fk_bar Hk + 2 F C_bar, C_bar = V (Gamma o V^T sym(F^T Fk_bar) V) V^T with the
    // divided differences Gamma_ij = (g_i - g_j)/(lambda_i - lambda_j), g'(lambda_i) at ties
    T gv[3], gd[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const T Sm = maxp(S[i], T(1e-12));
        gv[i] = Sk[i] / Sm;
        const bool pass = S[i] >= A.fc_lo && S[i] <= A.fc_hi;
        const T drdS = (pass ? T(1) : T(0)) / Sm - (S[i] >= T(1e-12) ? Sk[i] / (Sm * Sm) : T(0));
        gd[i] = lc[i] >= T(1e-12) ? drdS / (T(2) * S[i]) : T(0);
    }
    M3<T> VG;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) VG.m[i][j] = V.m[i][j] * gv[j];
    const M3<T> Hk = mul_bt(VG, V);
    const M3<T> t3 = mul(Fk_bar, Hk);
    const M3<T> Hb = mul_at(F, Fk_bar);
    M3<T> Hs;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Hs.m[i][j] = T(0.5) * (Hb.m[i][j] + Hb.m[j][i]);
    M3<T> K = mul(mul_at(V, Hs), V);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) K.m[i][j] *= lc[i] != lc[j] ? (gv[i] - gv[j]) / (lc[i] - lc[j]) : gd[i];
    const M3<T> Cb = mul_bt(mul(V, K), V);
    const M3<T> t4 = mul(F, Cb);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F_bar.m[i][j] += t3.m[i][j] + T(2) * t4.m[i][j];
}

// v'_bar and C'_bar seen by the transfer: the advection x' = clamp(x + dt v')
// and F' = (I + dt C') F routes part of x'_bar and F'_bar into them (F: the
// strain-limited F when f_clamp is on).
template <typename T>
__device__ __forceinline__ void transfer_bars(const StepArgs<T>& A, const AdjArgs<T>& B, long long p, const T xp[3],
                                              const M3<T>& F, T vbar[3], M3<T>& Cbar, bool xin[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T xn = xp[a] + A.dt * B.nv[3 * p + a];
        xin[a] = xn >= A.pad_lo && xn <= A.pad_hi;
        vbar[a] = B.gv[3 * p + a] + (xin[a] ? A.dt * B.gx[3 * p + a] : T(0));
    }
    const M3<T> Fb = load_m3(B.gF, p);
    const M3<T> FbFt = mul_bt(Fb, F);
    const M3<T> Cb = load_m3(B.gC, p);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Cbar.m[i][j] = Cb.m[i][j] + A.dt * FbFt.m[i][j];
}

// G2P adjoint onto the grid: per node v_bar += w (v'_bar + kappa C'_bar (x_i - x_p)),
// gd_bar += L lap'_bar.  Same schedule and frames as k_p2g; 4 channels.
template <typename T>
__global__ void __launch_bounds__(P2G_WG) k_adj_scatter(StepArgs<T> A, AdjArgs<T> B) {
    constexpr int NW = sizeof(T) == 4 ? 4 : 2;
    __shared__ __align__(16) T s_q[NW][AR3][4];
    const int nact = A.misc[0];
    const int nb = A.nb, ng = A.ng;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (BucketIter it(nact); it.i < it.end; it.i += it.step) {
        const int bk = A.alist[it.i];
        const int scene = (int)(bk / A.NBs), loc = (int)(bk - scene * A.NBs);
        const int bx = loc / (nb * nb), by = (loc / nb) % nb, bz = loc % nb;
        const int org[3] = {BLK * bx - 1, BLK * by - 1, BLK * bz - 1};
        for (int l = threadIdx.x; l < NW * 4 * AR3; l += P2G_WG) (&s_q[0][0][0])[l] = T(0);
        __syncthreads();
        const long long cell = ((long long)bk << 6) + lane;
        const int c0 = A.start[cell], cnt = A.start[cell + 1] - c0;
        int rounds = cnt;
        for (int o = 32; o > 0; o >>= 1) rounds = max(rounds, __shfl_xor(rounds, o));
        for (int r = wv; r < rounds; r += NW) {
            if (wv >= NW || r >= cnt) continue;
            const long long p = A.perm[c0 + r];
            const T xp[3] = {A.x[3 * p], A.x[3 * p + 1], A.x[3 * p + 2]};
            const M3<T> F = fclamped(A, load_m3(A.F, p));
            T vb[3];
            M3<T> Cb;
            bool xin[3];
            transfer_bars(A, B, p, xp, F, vb, Cb, xin);
            const T lb = B.glap[p];
            const T Xp[3] = {scaled(xp[0], A.inv_dx), scaled(xp[1], A.inv_dx), scaled(xp[2], A.inv_dx)};
            int bq[3], bcb[3];
            T fx[3], w[3][3], wc[4][3], w2[4][3];
            quad_weights(Xp, ng, bq, fx, w);
            cubic_weights(Xp, ng, bcb, wc, w2);
            int lp[3][4];
            T wq[4][3], dq[4][3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int sh = bq[a] - bcb[a];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    lp[a][k] = clampi(bcb[a] + k - org[a], 0, AR - 1);
                    const int j = k - sh;
                    wq[k][a] = j == 0 ? w[0][a] : (j == 1 ? w[1][a] : (j == 2 ? w[2][a] : T(0)));
                    dq[k][a] = (T(j) - fx[a]) * A.dx;
                }
            }
            T* acc = &s_q[wv < NW ? wv : 0][0][0];
#pragma unroll
            for (int ox = 0; ox < 4; ++ox)
#pragma unroll
                for (int oy = 0; oy < 4; ++oy)
#pragma unroll
                    for (int oz = 0; oz < 4; ++oz) {
                        const int l = (lp[0][ox] * AR + lp[1][oy]) * AR + lp[2][oz];
                        const T wk = wq[ox][0] * wq[oy][1] * wq[oz][2];
                        const T L = A.lap_scale * (w2[ox][0] * wc[oy][1] * wc[oz][2] + wc[ox][0] * w2[oy][1] * wc[oz][2] +
                                                   wc[ox][0] * wc[oy][1] * w2[oz][2]);
                        const T dp[3] = {dq[ox][0], dq[oy][1], dq[oz][2]};
                        T* q = acc + 4 * l;
                        T cv[4] = {q[0], q[1], q[2], q[3]};
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            const T cd = Cb.m[a][0] * dp[0] + Cb.m[a][1] * dp[1] + Cb.m[a][2] * dp[2];
                            cv[a] += wk * (vb[a] + A.c_scale * cd);
                        }
                        cv[3] += L * lb;
                        q[0] = cv[0]; q[1] = cv[1]; q[2] = cv[2]; q[3] = cv[3];
                        __asm__ __volatile__("" ::: "memory");
                    }
        }
        __syncthreads();
        T* dst = A.arena + (size_t)it.i * (4 * AR3);
        for (int l = threadIdx.x; l < 4 * AR3; l += P2G_WG) {
            const int c = l / AR3, node = l - c * AR3;
            T v = s_q[0][node][c];
            for (int w = 1; w < NW; ++w) v += s_q[w][node][c];
            dst[l] = v;
        }
        __syncthreads();
    }
}

// Grid adjoint per marked node block: gather the adjoint arenas (with the
// edge fold of k_grid), then
//   v = (mv/m + dt g) * damp  (active)   ->  mv_bar = damp v_bar / m,
//                                            m_bar = -damp v_bar . (mv/m) / m
//   gd = S/W (W > 1e-12)                 ->  S_bar = gd_bar / W, W_bar = -gd_bar gd / W
template <typename T>
__global__ void __launch_bounds__(WG) k_adj_grid(StepArgs<T> A, AdjArgs<T> B) {
    __shared__ pmpm_collider s_col[PMPM_MAX_COLLIDERS];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < PMPM_MAX_COLLIDERS; ++k) s_col[k] = A.btab ? A.btab->colliders[k] : A.col[k];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * (unsigned)WG + threadIdx.x) >> 6);
    const int nwaves = gridDim.x * (WG / 64);
    const int nblocks = A.misc[1];
    const int nb = A.nb, ng = A.ng;
    for (int i = wave; i < nblocks; i += nwaves) {
        const int blk = A.blist[i];
        const int scene = (int)(blk / A.NBs), loc = (int)(blk - scene * A.NBs);
        const int bx = loc / (nb * nb), by = (loc / nb) % nb, bz = loc % nb;
        const int g[3] = {BLK * bx + (lane >> 4), BLK * by + ((lane >> 2) & 3), BLK * bz + (lane & 3)};
        if (g[0] >= ng || g[1] >= ng || g[2] >= ng) continue;
        T acc[4] = {T(0), T(0), T(0), T(0)};
        const int* sl = A.slotof + scene * A.NBs;
        for (int k = 0; k < 27; ++k) {
            const int b3[3] = {bx - 1 + k / 9, by - 1 + (k / 3) % 3, bz - 1 + k % 3};
            if (b3[0] < 0 || b3[0] >= nb || b3[1] < 0 || b3[1] >= nb || b3[2] < 0 || b3[2] >= nb) continue;
            const int slot = sl[(b3[0] * nb + b3[1]) * nb + b3[2]];
            if (slot < 0) continue;
            int lo[3], hi[3];
            bool ok = true;
            for (int a = 0; a < 3; ++a) {
                const int o = BLK * b3[a] - 1;
                const int plo = g[a] == 0 ? o : g[a];
                const int phi = g[a] == ng - 1 ? o + AR - 1 : g[a];
                lo[a] = (plo > o ? plo : o) - o;
                hi[a] = (phi < o + AR - 1 ? phi : o + AR - 1) - o;
                ok = ok && lo[a] <= hi[a];
            }
            if (!ok) continue;
            const T* base = A.arena + (size_t)slot * (4 * AR3);
            for (int lx = lo[0]; lx <= hi[0]; ++lx)
                for (int ly = lo[1]; ly <= hi[1]; ++ly)
                    for (int lz = lo[2]; lz <= hi[2]; ++lz) {
                        const T* src = base + (lx * AR + ly) * AR + lz;
#pragma unroll
                        for (int c = 0; c < 4; ++c) acc[c] += src[c * AR3];
                    }
        }
        const long long node = scene * A.Gs + ((long long)g[0] * ng + g[1]) * ng + g[2];
        const T* gw = B.gmw + GMW * node;
        const T m = gw[0], Wsum = gw[1];
        const T* gv = A.gv + 4 * node;
        T out[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        if (m > T(1e-12)) {
            const T q[3] = {gw[2], gw[3], gw[4]};          // mv / m
            T vb[3] = {acc[0], acc[1], acc[2]};             // adjoint of the node's final velocity
            if (A.n_col > 0) {
                // back through the colliders in reverse order, each from the
                // velocity that entered it (the grid pass: gravity, damping,
                // then the colliders in list order, anisompm.py:730-746)
                T v0[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    v0[a] = q[a] + A.dtg[a];
                    if (A.damp != T(1)) v0[a] = v0[a] * A.damp;
                }
                const T pos[3] = {T(g[0]) * A.dx, T(g[1]) * A.dx, T(g[2]) * A.dx};
                for (int k = A.n_col - 1; k >= 0; --k) {
                    T vk[3] = {v0[0], v0[1], v0[2]};
                    for (int j = 0; j < k; ++j) {
                        T vc[3], nj[3];
                        if (collider_setup<T>(s_col[j], pos, vc, nj)) collider_resolve(vk, vc, nj, s_col[j].mode, T(s_col[j].friction));
                    }
                    T vc[3], nk[3];
                    if (collider_setup<T>(s_col[k], pos, vc, nk))
                        collider_resolve_vjp(vk, vc, nk, s_col[k].mode, T(s_col[k].friction), vb);
                }
            }
            T vbar_dot = T(0);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const T vbd = A.damp * vb[a];
                out[a] = vbd / m;
                vbar_dot += vbd * q[a];
            }
            out[3] = -vbar_dot / m;
        }
        if (Wsum > T(1e-12)) {
            out[4] = acc[3] / Wsum;
            out[5] = -acc[3] * gv[3] / Wsum;
        }
        T* dst = B.gadj + 6 * node;
#pragma unroll
        for (int c = 0; c < 6; ++c) dst[c] = out[c];
    }
}

// Per-particle adjoint: transfer and constitutive VJP (see the file header).
template <typename T>
__global__ void __launch_bounds__(WG) k_adj_gather(StepArgs<T> A, AdjArgs<T> B) {
    // particles in sorted order: neighbouring lanes share grid nodes in the caches
    const long long i = (long long)blockIdx.x * WG + threadIdx.x;
    if (i >= A.n) return;
    const long long p = A.perm[i];
    const int ng = A.ng;
    // scene of the particle's grid (single scene in the adjoint)
    const long long gbase = 0;
    const T xp[3] = {A.x[3 * p], A.x[3 * p + 1], A.x[3 * p + 2]};
    const M3<T> F = fclamped(A, load_m3(A.F, p));
    T vb[3];
    M3<T> Cb;
    bool xin[3];
    transfer_bars(A, B, p, xp, F, vb, Cb, xin);
    const T lb = B.glap[p];
    // forward quantities of the transfer
    const T mass = A.mass[p];
    T vin[3] = {A.v[3 * p], A.v[3 * p + 1], A.v[3 * p + 2]};
    const int bc = A.bc_id ? A.bc_id[p] : 0;
    bc_velocity(A, bc, vin);
    // the affine matrix (sk tau + m C) and the post-update d of the forward
    // step, from the P2G record the recompute just wrote at the particle's
    // sorted position
    const T* rr = A.rec + (size_t)P2G_REC * i;
    M3<T> af;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) af.m[a][b] = rr[3 + 3 * a + b];
    const T dnew = A.dsrt[i];
    T sk = T(-1) * A.dt * A.vol0[p];
    sk = sk * T(4);
    sk = sk * A.inv_dx;
    sk = sk * A.inv_dx;

    const T Xp[3] = {scaled(xp[0], A.inv_dx), scaled(xp[1], A.inv_dx), scaled(xp[2], A.inv_dx)};
    int bq[3], bcb[3];
    T fx[3], w[3][3], wc[4][3], w2[4][3];
    quad_weights(Xp, ng, bq, fx, w);
    cubic_weights(Xp, ng, bcb, wc, w2);
    // weight derivatives with respect to the scaled coordinate
    T dw[3][3], dwc[4][3], dw2[4][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        dw[0][a] = fx[a] - T(1.5);
        dw[1][a] = T(-2) * (fx[a] - T(1));
        dw[2][a] = fx[a] - T(0.5);
        const T f = Xp[a] - floor(Xp[a]), o = T(1) - f;
        dwc[0][a] = T(-0.5) * o * o;
        dwc[1][a] = T(-2) * f + T(1.5) * f * f;
        dwc[2][a] = T(2) * o - T(1.5) * o * o;
        dwc[3][a] = T(0.5) * f * f;
        dw2[0][a] = T(-1); dw2[1][a] = T(3); dw2[2][a] = T(-3); dw2[3][a] = T(1);
    }
    T xb[3] = {T(0), T(0), T(0)};
    T vinb[3] = {T(0), T(0), T(0)};
    M3<T> Ab;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) Ab.m[a][b] = T(0);
    T dnb_cubic = T(0);
    // quadratic stencil: G2P (v', C') and P2G (m, mv) terms; the x loop stays
    // rolled (a fully unrolled stencil keeps too many weight products live)
#pragma unroll 1
    for (int ox = 0; ox < 3; ++ox) {
        const T wx = pick<3>(w, ox, 0), dwx = pick<3>(dw, ox, 0);
#pragma unroll
        for (int oy = 0; oy < 3; ++oy)
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
                const long long node = gbase + node_lin(bq[0] + ox, bq[1] + oy, bq[2] + oz, ng);
                const T* gvn = A.gv + 4 * node;
                const T* ga = B.gadj + 6 * node;
                const T vi[3] = {gvn[0], gvn[1], gvn[2]};
                const T mvb[3] = {ga[0], ga[1], ga[2]};
                const T mb = ga[3];
                const T wk = wx * w[oy][1] * w[oz][2];
                const T gw[3] = {A.inv_dx * dwx * w[oy][1] * w[oz][2], A.inv_dx * wx * dw[oy][1] * w[oz][2],
                                 A.inv_dx * wx * w[oy][1] * dw[oz][2]};
                const T dp[3] = {(T(ox) - fx[0]) * A.dx, (T(oy) - fx[1]) * A.dx, (T(oz) - fx[2]) * A.dx};
                // G2P: v' = sum w v_i, C' = kappa sum w v_i dp^T
                T vCd = T(0), CtV[3] = {T(0), T(0), T(0)};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const T cd = Cb.m[a][0] * dp[0] + Cb.m[a][1] * dp[1] + Cb.m[a][2] * dp[2];
                    vCd += vi[a] * cd;
#pragma unroll
                    for (int b = 0; b < 3; ++b) CtV[b] += Cb.m[a][b] * vi[a];
                }
                const T sg = vi[0] * vb[0] + vi[1] * vb[1] + vi[2] * vb[2] + A.c_scale * vCd;
                // P2G: m_i = sum w m, mv_i = sum w (m v + af dp)
                T mom[3], mvbm = T(0), AtM[3] = {T(0), T(0), T(0)};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    mom[a] = mass * vin[a] + (af.m[a][0] * dp[0] + af.m[a][1] * dp[1] + af.m[a][2] * dp[2]);
                    mvbm += mvb[a] * mom[a];
#pragma unroll
                    for (int b = 0; b < 3; ++b) AtM[b] += af.m[a][b] * mvb[a];
                }
                const T sp2 = mass * mb + mvbm;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    xb[a] += gw[a] * (sg + sp2) - wk * (A.c_scale * CtV[a] + AtM[a]);
                    vinb[a] += wk * mass * mvb[a];
#pragma unroll
                    for (int b = 0; b < 3; ++b) Ab.m[a][b] += wk * mvb[a] * dp[b];
                }
            }
    }
    // cubic stencil: Laplacian gather and damage scatter (x loop rolled)
#pragma unroll 1
    for (int ox = 0; ox < 4; ++ox) {
        const T wcx = pick<4>(wc, ox, 0), dwcx = pick<4>(dwc, ox, 0), w2x = pick<4>(w2, ox, 0),
                dw2x = pick<4>(dw2, ox, 0);
#pragma unroll
        for (int oy = 0; oy < 4; ++oy)
#pragma unroll
            for (int oz = 0; oz < 4; ++oz) {
                const long long node = gbase + node_lin(bcb[0] + ox, bcb[1] + oy, bcb[2] + oz, ng);
                const T gd = A.gv[4 * node + 3];
                const T* ga = B.gadj + 6 * node;
                const T Sb = ga[4], Wb = ga[5];
                const T wck = wcx * wc[oy][1] * wc[oz][2];
                dnb_cubic += wck * Sb;
                const T gwc[3] = {dwcx * wc[oy][1] * wc[oz][2], wcx * dwc[oy][1] * wc[oz][2],
                                  wcx * wc[oy][1] * dwc[oz][2]};
                const T gL[3] = {
                    dw2x * wc[oy][1] * wc[oz][2] + dwcx * w2[oy][1] * wc[oz][2] + dwcx * wc[oy][1] * w2[oz][2],
                    w2x * dwc[oy][1] * wc[oz][2] + wcx * dw2[oy][1] * wc[oz][2] + wcx * dwc[oy][1] * w2[oz][2],
                    w2x * wc[oy][1] * dwc[oz][2] + wcx * w2[oy][1] * dwc[oz][2] + wcx * wc[oy][1] * dw2[oz][2]};
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    xb[a] += A.inv_dx * (gwc[a] * (dnew * Sb + Wb) + A.lap_scale * gL[a] * gd * lb);
            }
    }
    // affine: af = sk tau + m C
    M3<T> tau_bar, Cbar_in;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            tau_bar.m[a][b] = sk * Ab.m[a][b];
            Cbar_in.m[a][b] = mass * Ab.m[a][b];
        }
    T* pb = B.pbar + 10 * p;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) pb[3 * a + b] = tau_bar.m[a][b];
    pb[9] = dnb_cubic;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        B.ox[3 * p + a] = (xin[a] ? B.gx[3 * p + a] : T(0)) + xb[a];
        B.ov[3 * p + a] = bc > 0 ? T(0) : vinb[a];
    }
    store_m3(B.oC, p, Cbar_in);
}

// Constitutive VJP per particle, and the F update F' = (I + dt C') F.
template <typename T>
__global__ void __launch_bounds__(WG) k_adj_const(StepArgs<T> A, AdjArgs<T> B) {
    const long long p = (long long)blockIdx.x * WG + threadIdx.x;
    if (p >= A.n) return;
    const M3<T> F = load_m3(A.F, p);
    const T* pb = B.pbar + 10 * p;
    M3<T> tau_bar;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) tau_bar.m[a][b] = pb[3 * a + b];
    const T dnb_cubic = pb[9];
    // F' = (I + dt C') Fk: its adjoint reaches the limited F
    const M3<T> Fbn = load_m3(B.gF, p);
    const M3<T> Cn = load_m3(B.nC, p);
    M3<T> Ipc;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) Ipc.m[a][b] = (a == b ? T(1) : T(0)) + A.dt * Cn.m[a][b];
    const M3<T> Fk_bar = mul_at(Ipc, Fbn);
    M3<T> Fb, A0b;
    T d_b, lap_b, Dh_b, mu_b, lam_b, sc_b;
    const bool dmg_on = B.dmg_on != 0;
    constitutive_vjp(A, p, F, dmg_on, tau_bar, B.gd[p] + dnb_cubic, B.gD[p], Fk_bar, Fb, d_b, lap_b, Dh_b, mu_b, lam_b,
                     A0b, sc_b);
    B.od[p] = d_b;
    B.olap[p] = lap_b;
    B.oD[p] = Dh_b;
    store_m3(B.oF, p, Fb);
    B.omu[p] += mu_b;
    B.olam[p] += lam_b;
    B.osc[p] += sc_b;
    M3<T> a0 = load_m3((const T*)B.oA0, p);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) a0.m[a][b] += A0b.m[a][b];
    store_m3(B.oA0, p, a0);
}
