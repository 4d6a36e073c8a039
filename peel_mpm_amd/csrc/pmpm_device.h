// Per-particle and per-node device math of the AnisoMPM step, held in
// registers.  Every function mirrors a piece of
// anisompm_py/src/anisompm.py (cited per function) with the same operation
// order where that fixes rounding; the code is written for 64-lane CDNA4
// wavefronts, one particle (or node) per lane.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmpm {

template <typename T> struct Lim;
template <> struct Lim<float> {
    static __device__ __forceinline__ float maxv() { return 3.402823466e38f; }
};
template <> struct Lim<double> {
    static __device__ __forceinline__ double maxv() { return 1.7976931348623157e308; }
};

template <typename T> __device__ __forceinline__ bool isnan_(T a) { return a != a; }

// torch.maximum / clamp / relu propagate NaN; fmax would drop it.
template <typename T> __device__ __forceinline__ T maxp(T a, T b) {
    return (isnan_(a) || isnan_(b)) ? (a + b) : (a > b ? a : b);
}
template <typename T> __device__ __forceinline__ T minp(T a, T b) {
    return (isnan_(a) || isnan_(b)) ? (a + b) : (a < b ? a : b);
}
template <typename T> __device__ __forceinline__ T relu(T a) { return isnan_(a) ? a : (a > T(0) ? a : T(0)); }

// torch.nan_to_num defaults: NaN -> 0, +-inf -> +-max finite.
template <typename T> __device__ __forceinline__ T nan_to_num(T a) {
    if (isnan_(a)) return T(0);
    if (a > Lim<T>::maxv()) return Lim<T>::maxv();
    if (a < -Lim<T>::maxv()) return -Lim<T>::maxv();
    return a;
}

// ---------------------------------------------------------------- grid index
// Xp = x * inv_dx and the floors that pick base cells must round exactly as
// the reference's separate tensor ops (anisompm.py:548-549, :573-574): no
// contraction of the multiply into the subtract.
template <typename T> __device__ __forceinline__ T scaled(T x, T inv_dx) {
#pragma clang fp contract(off)
    return x * inv_dx;
}
template <typename T> __device__ __forceinline__ T sub_exact(T a, T b) {
#pragma clang fp contract(off)
    return a - b;
}

// floor() of a scaled coordinate as an int, limited to [-4, ng+4]: beyond
// that range every stencil node clamps to the same edge node, so the limit
// changes no index and keeps the int conversion in range (NaN -> -4, which
// clamps to node 0 like the reference's int64 conversion of NaN).
template <typename T> __device__ __forceinline__ int floor_idx(T v, int ng) {
    T f = floor(v);
    f = f > T(-4) ? f : T(-4);
    f = f < T(ng + 4) ? f : T(ng + 4);
    return (int)f;
}

__device__ __forceinline__ int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

// Linear node index with clamp-to-edge (anisompm.py:586-590).
__device__ __forceinline__ int node_lin(int i, int j, int k, int ng) {
    return (clampi(i, 0, ng - 1) * ng + clampi(j, 0, ng - 1)) * ng + clampi(k, 0, ng - 1);
}

// Quadratic B-spline weights (anisompm.py:544-557).  w[node][axis].
template <typename T>
__device__ __forceinline__ void quad_weights(const T Xp[3], int ng, int base[3], T fx[3], T w[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        T bm = sub_exact(Xp[a], T(0.5));
        base[a] = floor_idx(bm, ng);
        T bf = floor(bm);                       // exact value of the unclamped floor
        fx[a] = sub_exact(Xp[a], bf);
        T u = T(1.5) - fx[a], c = fx[a] - T(1), e = fx[a] - T(0.5);
        w[0][a] = T(0.5) * (u * u);
        w[1][a] = T(0.75) - c * c;
        w[2][a] = T(0.5) * (e * e);
    }
}

// Cubic B-spline weights and second derivatives in grid units (anisompm.py:559-584).
template <typename T>
__device__ __forceinline__ void cubic_weights(const T Xp[3], int ng, int base[3], T wc[4][3], T w2[4][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        T fl = floor(Xp[a]);
        base[a] = floor_idx(Xp[a], ng) - 1;
        T f = sub_exact(Xp[a], fl);
        T one = T(1) - f;
        T f2 = f * f, o2 = one * one;
        wc[0][a] = T(1.0 / 6.0) * (o2 * one);
        wc[1][a] = (T(2.0 / 3.0) - f2) + T(0.5) * (f2 * f);
        wc[2][a] = (T(2.0 / 3.0) - o2) + T(0.5) * (o2 * one);
        wc[3][a] = T(1.0 / 6.0) * (f2 * f);
        w2[0][a] = one;
        w2[1][a] = T(3) * f - T(2);
        w2[2][a] = T(1) - T(3) * f;
        w2[3][a] = f;
    }
}

// ---------------------------------------------------------------- 3x3 algebra
template <typename T> struct M3 {
    T m[3][3];
};

template <typename T> __device__ __forceinline__ M3<T> mul(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
// a @ b^T
template <typename T> __device__ __forceinline__ M3<T> mul_bt(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[j][0] + a.m[i][1] * b.m[j][1] + a.m[i][2] * b.m[j][2];
    return r;
}
// a^T @ b
template <typename T> __device__ __forceinline__ M3<T> mul_at(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[0][i] * b.m[0][j] + a.m[1][i] * b.m[1][j] + a.m[2][i] * b.m[2][j];
    return r;
}

// Reciprocal and reciprocal square root of the rotation angle.  fp32 uses
// the hardware v_rcp_f32 / v_rsq_f32 (about 1 ulp): the rotation only has to
// be a rotation to rounding, and c, s come out of correctly rounded division
// with 0.5 ulp error anyway.  fp64 keeps IEEE division and sqrt.
__device__ __forceinline__ float rot_rcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double rot_rcp(double v) { return 1.0 / v; }
__device__ __forceinline__ float rot_rsq(float v) { return __builtin_amdgcn_rsqf(v); }
__device__ __forceinline__ double rot_rsq(double v) { return 1.0 / sqrt(v); }
__device__ __forceinline__ float rot_sqrt(float v) { return __builtin_amdgcn_sqrtf(v); }
__device__ __forceinline__ double rot_sqrt(double v) { return sqrt(v); }

// One Jacobi rotation zeroing a[p][q] (anisompm.py:74-91), p<q, r the third
// index.  keep: leave a and V as they are (the lane has converged).
template <int p, int q, typename T>
__device__ __forceinline__ void jacobi_rot(T (&a)[3][3], T (&V)[3][3], bool keep = false) {
    constexpr int r = 3 - p - q;
    const T app = a[p][p], aqq = a[q][q], apq = a[p][q], apr = a[p][r], aqr = a[q][r];
    const T zeta = (aqq - app) * rot_rcp(T(2) * apq);
    const T sgn = zeta >= T(0) ? T(1) : T(-1);
    T t = sgn * rot_rcp(fabs(zeta) + rot_sqrt(T(1) + zeta * zeta));
    t = fabs(apq) > T(1e-30) ? t : T(0);
    t = keep ? T(0) : t;
    const T c = keep ? T(1) : rot_rsq(T(1) + t * t);
    const T s = c * t;
    a[p][p] = c * c * app - T(2) * s * c * apq + s * s * aqq;
    a[q][q] = s * s * app + T(2) * s * c * apq + c * c * aqq;
    a[p][q] = keep ? apq : T(0); a[q][p] = a[p][q];
    const T npr = c * apr - s * aqr, nqr = s * apr + c * aqr;
    a[p][r] = npr; a[r][p] = npr;
    a[q][r] = nqr; a[r][q] = nqr;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const T vp = V[i][p], vq = V[i][q];
        V[i][p] = c * vp - s * vq;
        V[i][q] = s * vp + c * vq;
    }
}

// Off-diagonal size, relative to the diagonal, below which a Jacobi sweep
// no longer changes a: its rotations have |t| <= ~tol / (relative gap), which
// moves eigenvalues by ~tol^2 and eigenvectors by about one rounding unit
// (or, for equal eigenvalues, within their common eigenspace).
template <typename T> __device__ __forceinline__ T jacobi_tol();
template <> __device__ __forceinline__ float jacobi_tol<float>() { return 1e-9f; }
template <> __device__ __forceinline__ double jacobi_tol<double>() { return 1e-18; }

// Symmetric 3x3 eigendecomposition by 5 cyclic Jacobi sweeps
// (anisompm.py:48-95): symmetrise, sanitise, rotate, sort ascending; V holds
// the eigenvectors as columns.  Ties keep their diagonal order.
// CONVERGED: a lane whose off-diagonal part has fallen below jacobi_tol of
// its diagonal after a sweep takes no further rotations (a per-lane rule, so
// the result does not depend on which particles share a wavefront), and a
// wavefront whose lanes have all converged skips the remaining sweeps.
template <typename T, bool CONVERGED = false>
__device__ __forceinline__ void sym_eig3(const M3<T>& A, T ev[3], M3<T>& Vout) {
    T a[3][3], V[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            a[i][j] = nan_to_num(T(0.5) * (A.m[i][j] + A.m[j][i]));
            V[i][j] = i == j ? T(1) : T(0);
        }
    bool done = false;
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        if constexpr (CONVERGED) {
            if (s > 0) {
                const T off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
                const T dia = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
                done = done || off <= jacobi_tol<T>() * dia;
                if (__all(done)) break;
            }
        }
        jacobi_rot<0, 1>(a, V, done);
        jacobi_rot<0, 2>(a, V, done);
        jacobi_rot<1, 2>(a, V, done);
    }
    T e[3] = {a[0][0], a[1][1], a[2][2]};
    int o[3] = {0, 1, 2};
    // three-compare stable sorting network
    if (e[1] < e[0]) { T t = e[0]; e[0] = e[1]; e[1] = t; int u = o[0]; o[0] = o[1]; o[1] = u; }
    if (e[2] < e[1]) { T t = e[1]; e[1] = e[2]; e[2] = t; int u = o[1]; o[1] = o[2]; o[2] = u; }
    if (e[1] < e[0]) { T t = e[0]; e[0] = e[1]; e[1] = t; int u = o[0]; o[0] = o[1]; o[1] = u; }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ev[c] = e[c];
#pragma unroll
        for (int i = 0; i < 3; ++i) Vout.m[i][c] = o[c] == 0 ? V[i][0] : (o[c] == 1 ? V[i][1] : V[i][2]);
    }
}

// ---------------------------------------------------------------- colliders
// Sticky / slip(+friction) projection (anisompm.py:811-825).
template <typename T>
__device__ __forceinline__ void collider_resolve(T v[3], const T vc[3], const T n[3], int mode, T friction) {
    if (mode == 0) {
        v[0] = vc[0]; v[1] = vc[1]; v[2] = vc[2];
        return;
    }
    T vr[3] = {v[0] - vc[0], v[1] - vc[1], v[2] - vc[2]};
    T vn = vr[0] * n[0] + vr[1] * n[1] + vr[2] * n[2];
    T inward = minp(vn, T(0));
    T vt[3] = {vr[0] - vn * n[0], vr[1] - vn * n[1], vr[2] - vn * n[2]};
    if (friction > T(0)) {
        T mag = sqrt(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
        T sc = maxp(T(1) + friction * inward / (mag + T(1e-9)), T(0));
        vt[0] *= sc; vt[1] *= sc; vt[2] *= sc;
    }
    T keep = vn - inward;
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = vc[a] + vt[a] + keep * n[a];
}

}  // namespace pmpm
