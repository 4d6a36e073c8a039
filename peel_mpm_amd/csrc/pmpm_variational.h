// Option A directional damage, the reference's variationally consistent
// pair (anisompm_py/src/anisompm.py:129-187): per particle the energy
// density
//   Psi = mu |F - R|^2 + lam/2 (J - 1)^2
//       + gI(d) gamma_I/2 <lam_n - 1>^2 + gII(d) gamma_II/2 (I5 - I4^2),
//   I4 = |F n0|^2 = lam_n^2,  I5 = |C n0|^2,  C = F^T F,  J = det F,
//   gI(d) = gII(d) = (1 - d)^2 + 1e-4,
// and its first Piola-Kirchhoff stress P = dPsi/dF
//   P = 2 mu (F - R) + lam (J - 1) J F^-T
//     + gI  gamma_I (<lam_n - 1> / lam_n) (F n0) n0^T
//     + gII gamma_II ((F n0)(C n0)^T + (F C n0) n0^T - 2 I4 (F n0) n0^T).
// R is the rotation of the polar decomposition (the reference takes U V^T of
// an SVD: the same matrix for det F != 0), from the Jacobi eigenvectors of
// C as in the step; J F^-T is the cofactor matrix of F.  One thread per
// particle.  Included by pmpm.hip inside its anonymous namespace.

template <typename T> __device__ __forceinline__ T option_a_g(T d) {
    const T om = T(1) - d;
    return om * om + T(1e-4);
}

template <typename T>
__device__ __forceinline__ T param_at(const T* __restrict__ a, T s, long long p) {
    return a ? a[p] : s;
}

template <typename T>
__global__ void __launch_bounds__(WG) k_option_a(long long n, const T* __restrict__ Fp, const T* __restrict__ n0p,
                                                 const T* __restrict__ dp, const T* __restrict__ mu_a, T mu_s,
                                                 const T* __restrict__ lam_a, T lam_s, const T* __restrict__ gI_a,
                                                 T gI_s, const T* __restrict__ gII_a, T gII_s, T* __restrict__ P,
                                                 T* __restrict__ psi) {
    const long long p = (long long)blockIdx.x * WG + threadIdx.x;
    if (p >= n) return;
    const M3<T> F = load_m3(Fp, p);
    const T n0[3] = {n0p[3 * p], n0p[3 * p + 1], n0p[3 * p + 2]};
    const T d = dp[p];
    const T mu = param_at(mu_a, mu_s, p), lam = param_at(lam_a, lam_s, p);
    const T gam1 = param_at(gI_a, gI_s, p), gam2 = param_at(gII_a, gII_s, p);
    // polar rotation R = (F V S^-1) V^T with V, S^2 the eigenpairs of F^T F
    const M3<T> C = mul_at(F, F);
    T lc[3];
    M3<T> V;
    sym_eig3<T, true>(C, lc, V);
    M3<T> W = mul(F, V), Wh;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const T inv = T(1) / sqrt(W.m[0][j] * W.m[0][j] + W.m[1][j] * W.m[1][j] + W.m[2][j] * W.m[2][j]);
#pragma unroll
        for (int i = 0; i < 3; ++i) Wh.m[i][j] = W.m[i][j] * inv;
    }
    const M3<T> R = mul_bt(Wh, V);
    // J and the cofactor matrix (J F^-T)
    M3<T> cof;
    cof.m[0][0] = F.m[1][1] * F.m[2][2] - F.m[1][2] * F.m[2][1];
    cof.m[0][1] = F.m[1][2] * F.m[2][0] - F.m[1][0] * F.m[2][2];
    cof.m[0][2] = F.m[1][0] * F.m[2][1] - F.m[1][1] * F.m[2][0];
    cof.m[1][0] = F.m[0][2] * F.m[2][1] - F.m[0][1] * F.m[2][2];
    cof.m[1][1] = F.m[0][0] * F.m[2][2] - F.m[0][2] * F.m[2][0];
    cof.m[1][2] = F.m[0][1] * F.m[2][0] - F.m[0][0] * F.m[2][1];
    cof.m[2][0] = F.m[0][1] * F.m[1][2] - F.m[0][2] * F.m[1][1];
    cof.m[2][1] = F.m[0][2] * F.m[1][0] - F.m[0][0] * F.m[1][2];
    cof.m[2][2] = F.m[0][0] * F.m[1][1] - F.m[0][1] * F.m[1][0];
    const T J = F.m[0][0] * cof.m[0][0] + F.m[0][1] * cof.m[0][1] + F.m[0][2] * cof.m[0][2];
    // invariants
    T Fn[3], Cn[3], FCn[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        Fn[i] = F.m[i][0] * n0[0] + F.m[i][1] * n0[1] + F.m[i][2] * n0[2];
        Cn[i] = C.m[i][0] * n0[0] + C.m[i][1] * n0[1] + C.m[i][2] * n0[2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) FCn[i] = F.m[i][0] * Cn[0] + F.m[i][1] * Cn[1] + F.m[i][2] * Cn[2];
    const T I4 = Fn[0] * Fn[0] + Fn[1] * Fn[1] + Fn[2] * Fn[2];
    const T I5 = Cn[0] * Cn[0] + Cn[1] * Cn[1] + Cn[2] * Cn[2];
    const T lam_n = sqrt(I4);
    const T opening = maxp(lam_n - T(1), T(0));
    const T gI = option_a_g(d), gII = option_a_g(d);
    if (psi) {
        T fr = T(0);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const T e = F.m[i][j] - R.m[i][j];
                fr += e * e;
            }
        psi[p] = mu * fr + T(0.5) * lam * (J - T(1)) * (J - T(1)) + gI * (T(0.5) * gam1 * opening * opening) +
                 gII * (T(0.5) * gam2 * (I5 - I4 * I4));
    }
    if (P) {
        const T vol = lam * (J - T(1));
        const T a1 = gI * gam1 * (opening / lam_n), a2 = gII * gam2;
        M3<T> out;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                out.m[i][j] = T(2) * mu * (F.m[i][j] - R.m[i][j]) + vol * cof.m[i][j] + a1 * Fn[i] * n0[j] +
                              a2 * (Fn[i] * Cn[j] + FCn[i] * n0[j] - T(2) * I4 * Fn[i] * n0[j]);
        store_m3(P, p, out);
    }
}

template <typename T>
int launch_option_a(long long n, const void* F, const void* n0, const void* d, const void* mu_a, double mu,
                    const void* lam_a, double lam, const void* gI_a, double gam1, const void* gII_a, double gam2, void* P,
                    void* psi, hipStream_t st) {
    hipLaunchKernelGGL(k_option_a<T>, dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, st, n, (const T*)F, (const T*)n0,
                       (const T*)d, (const T*)mu_a, (T)mu, (const T*)lam_a, (T)lam, (const T*)gI_a, (T)gam1,
                       (const T*)gII_a, (T)gam2, (T*)P, (T*)psi);
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}
