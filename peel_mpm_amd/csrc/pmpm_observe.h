// Peel-test observables on the device: the per-frame reductions of the
// reference's frame loop (anisompm_py/src/peel_test.py:271-310) without a
// host synchronisation per frame.
//   k_obs_frame    per particle: running maxima phi_pp = max(phi_pp, phi),
//                  psh_pp = max(psh_pp, phi_shear); per-work-group partials
//                  of the region reductions
//   k_obs_reduce   one work-group: the partials in a fixed order -> peaks,
//                  routing, the onset test, the frame's log row
//   k_obs_snapshot per particle, only on the onset frame: the onset
//                  snapshot (d, phi_pp, psh_pp, tau_last, J_last, x)
// Maxima propagate NaN as torch's max and maximum do.  Included by pmpm.hip
// inside its anonymous namespace.

constexpr int OBS_WG_COUNT = 1024;     // work-groups of k_obs_frame = partials per frame
// partial slots per work-group
enum { OBS_PHI_INT, OBS_PHI_FLESH, OBS_PHI_PEEL, OBS_PSH_INT, OBS_DSUM_INT, OBS_DMAX_INT, OBS_DBROKEN_INT,
       OBS_N_INT, OBS_DMAX_FLESH, OBS_DMAX_PEEL, OBS_CAPY_SUM, OBS_N_CAP, OBS_NAN, OBS_NSLOT };

__device__ __forceinline__ double nanmax(double a, double b) { return (a != a || b != b) ? (a != a ? a : b) : (a > b ? a : b); }
__device__ __forceinline__ bool obs_is_sum(int k) {
    return k == OBS_DSUM_INT || k == OBS_DBROKEN_INT || k == OBS_N_INT || k == OBS_CAPY_SUM || k == OBS_N_CAP || k == OBS_NAN;
}

template <typename T>
__global__ void __launch_bounds__(WG) k_obs_frame(long long n, const T* __restrict__ phi, const T* __restrict__ psh,
                                                  const T* __restrict__ d, const T* __restrict__ x,
                                                  const int8_t* __restrict__ region, const uint8_t* __restrict__ cap,
                                                  T* __restrict__ phi_pp, T* __restrict__ psh_pp,
                                                  double* __restrict__ part) {
    const double NINF = -__builtin_inf();
    double v[OBS_NSLOT] = {NINF, NINF, NINF, NINF, 0.0, NINF, 0.0, 0.0, NINF, NINF, 0.0, 0.0, 0.0};
    for (long long p = (long long)blockIdx.x * WG + threadIdx.x; p < n; p += (long long)gridDim.x * WG) {
        const int r = region[p];
        if (phi) {
            const T f = phi[p];
            const T m = phi_pp[p];
            phi_pp[p] = (f != f || m != m) ? (f != f ? f : m) : (f > m ? f : m);
            // compile-time slots only (a runtime index would put v in scratch)
            const double fd = (double)f;
            if (r == 1) v[OBS_PHI_INT] = nanmax(v[OBS_PHI_INT], fd);
            else if (r == 0) v[OBS_PHI_FLESH] = nanmax(v[OBS_PHI_FLESH], fd);
            else v[OBS_PHI_PEEL] = nanmax(v[OBS_PHI_PEEL], fd);
        }
        if (psh) {
            const T f = psh[p];
            const T m = psh_pp[p];
            psh_pp[p] = (f != f || m != m) ? (f != f ? f : m) : (f > m ? f : m);
            if (r == 1) v[OBS_PSH_INT] = nanmax(v[OBS_PSH_INT], (double)f);
        }
        const double dp = (double)d[p];
        if (r == 1) {
            v[OBS_DSUM_INT] += dp;
            v[OBS_DMAX_INT] = nanmax(v[OBS_DMAX_INT], dp);
            v[OBS_DBROKEN_INT] += dp > 0.5 ? 1.0 : 0.0;
            v[OBS_N_INT] += 1.0;
        } else if (r == 0) {
            v[OBS_DMAX_FLESH] = nanmax(v[OBS_DMAX_FLESH], dp);
        } else {
            v[OBS_DMAX_PEEL] = nanmax(v[OBS_DMAX_PEEL], dp);
        }
        const T x0 = x[3 * p], x1 = x[3 * p + 1], x2 = x[3 * p + 2];
        if (cap[p]) { v[OBS_CAPY_SUM] += (double)x1; v[OBS_N_CAP] += 1.0; }
        if (x0 != x0 || x1 != x1 || x2 != x2) v[OBS_NAN] = 1.0;
    }
    // work-group reduction in a fixed order (wave shuffles, then waves in order)
    __shared__ double s_v[WG / 64][OBS_NSLOT];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < OBS_NSLOT; ++k) {
        const bool sum = obs_is_sum(k);
        double a = v[k];
        for (int o = 32; o > 0; o >>= 1) {
            const double b = __shfl_xor(a, o);
            a = sum ? a + b : nanmax(a, b);
        }
        if (lane == 0) s_v[w][k] = a;
    }
    __syncthreads();
    if (threadIdx.x < OBS_NSLOT) {
        const int k = threadIdx.x;
        const bool sum = obs_is_sum(k);
        double a = s_v[0][k];
        for (int j = 1; j < WG / 64; ++j) a = sum ? a + s_v[j][k] : nanmax(a, s_v[j][k]);
        part[(size_t)blockIdx.x * OBS_NSLOT + k] = a;
    }
}

// peaks: [phi_peak interface, flesh, peel, phi_shear_peak, peak_routing];
// onset: [onset frame (-1 before onset), snapshot on this frame]; row: the
// frame's 10 log values (anisompm_py/src/peel_test.py:295-298).  IS32: the
// step runs in fp32, so the onset fraction and the routing ratio are formed
// in fp32 as the reference's fp32 tensor operations form them.  The
// partials are combined in a fixed order: thread t takes partials t, t+256,
// ... in turn, then a tree over the threads.
constexpr int OBS_RED = 256;
__global__ void __launch_bounds__(OBS_RED) k_obs_reduce(const double* __restrict__ part, int nparts, int frame, int has_phi,
                                                        int has_psh, int is32, double* __restrict__ peaks,
                                                        int* __restrict__ onset, double* __restrict__ row) {
    __shared__ double s_v[OBS_NSLOT][OBS_RED];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < OBS_NSLOT; ++k) {
        double a = obs_is_sum(k) ? 0.0 : -__builtin_inf();
        for (int j = t; j < nparts; j += OBS_RED) {
            const double b = part[(size_t)j * OBS_NSLOT + k];
            a = obs_is_sum(k) ? a + b : nanmax(a, b);
        }
        s_v[k][t] = a;
    }
    __syncthreads();
    for (int h = OBS_RED / 2; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int k = 0; k < OBS_NSLOT; ++k)
                s_v[k][t] = obs_is_sum(k) ? s_v[k][t] + s_v[k][t + h] : nanmax(s_v[k][t], s_v[k][t + h]);
        }
        __syncthreads();
    }
    if (t != 0) return;
    double tot[OBS_NSLOT];
#pragma unroll
    for (int k = 0; k < OBS_NSLOT; ++k) tot[k] = s_v[k][0];
    const double NaN = __builtin_nan("");
    double phf[3] = {NaN, NaN, NaN};
    if (has_phi) {
        phf[0] = tot[OBS_PHI_INT]; phf[1] = tot[OBS_PHI_FLESH]; phf[2] = tot[OBS_PHI_PEEL];
        for (int k = 0; k < 3; ++k) peaks[k] = nanmax(peaks[k], phf[k]);
        double r;
        if (is32) r = (double)((float)phf[0] / ((float)nanmax(phf[1], phf[2]) + 1e-9f));
        else r = phf[0] / (nanmax(phf[1], phf[2]) + 1e-9);
        peaks[4] = nanmax(peaks[4], r);
    }
    if (has_psh) peaks[3] = nanmax(peaks[3], tot[OBS_PSH_INT]);
    const double n_int = tot[OBS_N_INT];
    double frac;
    if (is32) frac = (double)((float)tot[OBS_DBROKEN_INT] / (float)n_int);
    else frac = tot[OBS_DBROKEN_INT] / n_int;
    onset[1] = 0;
    if (onset[0] < 0 && frac > 0.05) { onset[0] = frame; onset[1] = 1; }
    row[0] = tot[OBS_DSUM_INT] / n_int;
    row[1] = tot[OBS_DMAX_INT];
    row[2] = frac;
    row[3] = tot[OBS_DMAX_FLESH];
    row[4] = tot[OBS_DMAX_PEEL];
    row[5] = phf[0]; row[6] = phf[1]; row[7] = phf[2];
    row[8] = tot[OBS_CAPY_SUM] / tot[OBS_N_CAP];
    row[9] = tot[OBS_NAN] > 0.0 ? 1.0 : 0.0;
}

template <typename T>
__global__ void __launch_bounds__(WG) k_obs_snapshot(long long n, const int* __restrict__ onset, const T* __restrict__ d,
                                                     const T* __restrict__ x, const T* __restrict__ tau,
                                                     const T* __restrict__ J, const T* __restrict__ phi_pp,
                                                     const T* __restrict__ psh_pp, T* __restrict__ d_e,
                                                     T* __restrict__ x_e, T* __restrict__ tau_e, T* __restrict__ J_e,
                                                     T* __restrict__ phipp_e, T* __restrict__ pshpp_e) {
    if (onset[1] == 0) return;
    for (long long p = (long long)blockIdx.x * WG + threadIdx.x; p < n; p += (long long)gridDim.x * WG) {
        d_e[p] = d[p];
        phipp_e[p] = phi_pp[p];
        pshpp_e[p] = psh_pp[p];
#pragma unroll
        for (int a = 0; a < 3; ++a) x_e[3 * p + a] = x[3 * p + a];
        if (tau_e) {
#pragma unroll
            for (int k = 0; k < 9; ++k) tau_e[9 * p + k] = tau[9 * p + k];
        }
        if (J_e) J_e[p] = J[p];
    }
}

template <typename T>
int launch_observe(long long n, int frame, const pmpm_frame_in* in, const pmpm_frame_obs* o, double* part, hipStream_t st) {
    const unsigned G = (unsigned)std::min<long long>(OBS_WG_COUNT, (n + WG - 1) / WG);
    hipLaunchKernelGGL(k_obs_frame<T>, dim3(G), dim3(WG), 0, st, n, (const T*)in->phi, (const T*)in->phi_shear,
                       (const T*)in->d, (const T*)in->x, in->region, in->cap, (T*)o->phi_pp, (T*)o->psh_pp, part);
    hipLaunchKernelGGL(k_obs_reduce, dim3(1), dim3(OBS_RED), 0, st, (const double*)part, (int)G, frame, in->phi != nullptr,
                       in->phi_shear != nullptr, (int)(sizeof(T) == 4), o->peaks, o->onset, o->log + (size_t)frame * 10);
    hipLaunchKernelGGL(k_obs_snapshot<T>, dim3(G), dim3(WG), 0, st, n, (const int*)o->onset, (const T*)in->d, (const T*)in->x,
                       (const T*)in->tau_last, (const T*)in->J_last, (const T*)o->phi_pp, (const T*)o->psh_pp, (T*)o->d_e,
                       (T*)o->x_e, (T*)o->tau_e, (T*)o->J_e, (T*)o->phipp_e, (T*)o->pshpp_e);
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}
