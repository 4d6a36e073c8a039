// Gaussian carry: every splat rides on the affine motion of the particle it
// is bound to (anisompm_py/src/run_orange.py:116-128):
//   pos   = sim_to_world(x_p) + F_p off0,   sim_to_world(y) = (y / grid_lim - 0.5) / scale_to_sim + center
//   cov   = (F_p cov0) F_p^T
//   color = color0 (1 - crack_tint d_p)      (only when crack_tint > 0)
// One thread per splat: it gathers its particle's x, F (and d) through the
// binding and streams its own off0, cov0, color in and pos, cov, color out.
// Included by pmpm.hip inside its anonymous namespace.

template <typename T>
__global__ void __launch_bounds__(WG) k_gaussian_carry(long long G, const long long* __restrict__ bind,
                                                       const T* __restrict__ x, const T* __restrict__ F,
                                                       const T* __restrict__ d, const T* __restrict__ off0,
                                                       const T* __restrict__ cov0, const T* __restrict__ color,
                                                       T grid_lim, T scale_to_sim, T c0, T c1, T c2, T tint,
                                                       T* __restrict__ pos, T* __restrict__ cov,
                                                       T* __restrict__ ocolor) {
    const long long g = (long long)blockIdx.x * WG + threadIdx.x;
    if (g >= G) return;
    const long long p = bind[g];
    const M3<T> Fp = load_m3(F, p);
    const T xp[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
    const T o[3] = {off0[3 * g], off0[3 * g + 1], off0[3 * g + 2]};
    const M3<T> S = load_m3(cov0, g);
    const T cen[3] = {c0, c1, c2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // the reference's elementwise order: divide, subtract, divide, add
        const T w = (xp[a] / grid_lim - T(0.5)) / scale_to_sim + cen[a];
        pos[3 * g + a] = w + (Fp.m[a][0] * o[0] + Fp.m[a][1] * o[1] + Fp.m[a][2] * o[2]);
    }
    store_m3(cov, g, mul_bt(mul(Fp, S), Fp));
    if (ocolor) {
        const T k = tint > T(0) ? T(1) - tint * d[p] : T(1);
#pragma unroll
        for (int a = 0; a < 3; ++a) ocolor[3 * g + a] = tint > T(0) ? color[3 * g + a] * k : color[3 * g + a];
    }
}

template <typename T>
int launch_gaussian_carry(long long G, const long long* bind, const void* x, const void* F, const void* d, const void* off0,
                          const void* cov0, const void* color, double grid_lim, double scale_to_sim, const double center[3],
                          double crack_tint, void* pos, void* cov, void* ocolor, hipStream_t st) {
    hipLaunchKernelGGL(k_gaussian_carry<T>, dim3((unsigned)((G + WG - 1) / WG)), dim3(WG), 0, st, G, bind, (const T*)x,
                       (const T*)F, (const T*)d, (const T*)off0, (const T*)cov0, (const T*)color, T(grid_lim),
                       T(scale_to_sim), T(center[0]), T(center[1]), T(center[2]), T(crack_tint), (T*)pos, (T*)cov,
                       (T*)ocolor);
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}
