// Scene builders on the device: the reference's particle-set constructors,
// bit-exact with their NumPy / SciPy results.
//   ball_particles  jittered lattice points inside a ball
//                   (anisompm_py/src/peel_test.py:128-137)
//   grid_block      the cell-centred lattice of a box (test_block.py:41-47)
//   fill_interior   voxelised point cloud, closed, holes filled, one
//                   jittered particle per solid voxel (orange_setup.py:65-92)
// The jitter draws are NumPy's default_rng stream (PCG64, XSL-RR output,
// doubles from the top 53 bits), reproduced per element by LCG jump-ahead,
// so every thread produces its own draws independently.  Arithmetic follows
// NumPy's operation order with contraction off (NumPy rounds every product
// and sum).  Included by pmpm.hip inside its anonymous namespace.

typedef unsigned __int128 u128;

// PCG64's 128-bit LCG multiplier
__host__ __device__ constexpr u128 pcg_mult() {
    return ((u128)0x2360ED051FC65DA4ULL << 64) | (u128)0x4385DF649FCCF645ULL;
}

struct Pcg {
    u128 state, inc;
};

// The generator delta steps ahead: s -> A^delta s + C(delta), by squaring
// the step map (Brown, "Random number generation with arbitrary strides",
// 1994).
__device__ __forceinline__ Pcg pcg_advance(Pcg r, unsigned long long delta) {
    u128 acc_mult = 1, acc_plus = 0, cur_mult = pcg_mult(), cur_plus = r.inc;
    while (delta) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    r.state = acc_mult * r.state + acc_plus;
    return r;
}

// Next double in [0, 1): step, XSL-RR output, top 53 bits.
__device__ __forceinline__ double pcg_next_double(Pcg& r) {
    r.state = r.state * pcg_mult() + r.inc;
    const unsigned long long v = (unsigned long long)(r.state >> 64) ^ (unsigned long long)r.state;
    const unsigned rot = (unsigned)(r.state >> 122);
    const unsigned long long out = (v >> rot) | (v << ((64u - rot) & 63u));
    return (double)(out >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ Pcg pcg_from(const pmpm_pcg64& s) {
    Pcg r;
    r.state = ((u128)s.state_hi << 64) | (u128)s.state_lo;
    r.inc = ((u128)s.inc_hi << 64) | (u128)s.inc_lo;
    return r;
}

// ------------------------------------------------------------ stable compaction
// Keep-flags over N candidates in index order: per work-group chunk of
// SC_CHUNK candidates a count (k_sc_count), one work-group's exclusive scan
// of the counts (k_sc_scan), then every chunk recomputes its flags, ranks
// them inside the work-group and emits the kept candidates at their global
// rank (the caller's emit kernel, through sc_rank).
constexpr int SC_ITEMS = 4;
constexpr int SC_CHUNK = WG * SC_ITEMS;

template <typename Flag>
__global__ void __launch_bounds__(WG) k_sc_count(long long N, Flag flag, int* __restrict__ wgcount) {
    const long long b0 = (long long)blockIdx.x * SC_CHUNK;
    int c = 0;
#pragma unroll
    for (int t = 0; t < SC_ITEMS; ++t) {
        const long long i = b0 + (long long)t * WG + threadIdx.x;
        c += (i < N && flag(i)) ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    __shared__ int s[WG / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < WG / 64; ++w) t += s[w];
        wgcount[blockIdx.x] = t;
    }
}

// Exclusive scan of nwg counts by one work-group (in place); *total = the sum.
__global__ void __launch_bounds__(WG) k_sc_scan(int* __restrict__ wgcount, long long nwg, long long* __restrict__ total) {
    __shared__ long long s_carry;
    __shared__ long long s_w[WG / 64];
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (long long b = 0; b < nwg; b += WG) {
        const long long i = b + threadIdx.x;
        const long long v = i < nwg ? wgcount[i] : 0;
        long long x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const long long y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_w[w] = x;
        __syncthreads();
        long long pre = s_carry;
        for (int k = 0; k < w; ++k) pre += s_w[k];
        if (i < nwg) wgcount[i] = (int)(pre + x - v);
        __syncthreads();
        if (threadIdx.x == WG - 1) s_carry = pre + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = s_carry;
}

// Global rank of candidate i (kept) within the work-group's chunk, for item
// t of this thread; keep[t] are this thread's flags.  Work-group uniform.
__device__ __forceinline__ void sc_rank(const bool keep[SC_ITEMS], const int* __restrict__ wgoff, long long rank[SC_ITEMS]) {
    __shared__ int s_w[SC_ITEMS][WG / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int below[SC_ITEMS];
#pragma unroll
    for (int t = 0; t < SC_ITEMS; ++t) {
        const unsigned long long bal = __ballot(keep[t]);
        below[t] = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) s_w[t][w] = __popcll(bal);
    }
    __syncthreads();
    long long base = wgoff[blockIdx.x];
#pragma unroll
    for (int t = 0; t < SC_ITEMS; ++t) {
        long long pre = base;
        for (int k = 0; k < w; ++k) pre += s_w[t][k];
        rank[t] = pre + below[t];
        for (int k = 0; k < WG / 64; ++k) base += s_w[t][k];
    }
}

// ------------------------------------------------------------ ball_particles
// Candidate i = (a*n + b)*n + c of the lattice g x g x g (meshgrid "ij",
// raveled in C order) is kept when sqrt(x^2 + y^2 + z^2) < R.
struct BallFlag {
    const double* g;
    long long n;
    double R;
    __device__ __forceinline__ bool operator()(long long i) const {
#pragma clang fp contract(off)
        const double x = g[i / (n * n)], y = g[(i / n) % n], z = g[i % n];
        const double s = (x * x + y * y) + z * z;
        return __builtin_sqrt(s) < R;
    }
};

// The k-th kept point: (p + (u - 0.5) * js) + center with u = draws 3k..3k+2,
// rounded to fp32 (peel_test.py:136-137).
__global__ void __launch_bounds__(WG) k_ball_emit(BallFlag f, long long N, const int* __restrict__ wgoff, double js, double c0,
                                                  double c1, double c2, pmpm_pcg64 seed, float* __restrict__ out) {
#pragma clang fp contract(off)
    const long long b0 = (long long)blockIdx.x * SC_CHUNK;
    bool keep[SC_ITEMS];
    long long idx[SC_ITEMS], rank[SC_ITEMS];
#pragma unroll
    for (int t = 0; t < SC_ITEMS; ++t) {
        idx[t] = b0 + (long long)t * WG + threadIdx.x;
        keep[t] = idx[t] < N && f(idx[t]);
    }
    sc_rank(keep, wgoff, rank);
    const Pcg r0 = pcg_from(seed);
    const double c[3] = {c0, c1, c2};
#pragma unroll
    for (int t = 0; t < SC_ITEMS; ++t) {
        if (!keep[t]) continue;
        const long long i = idx[t], n = f.n;
        const double p[3] = {f.g[i / (n * n)], f.g[(i / n) % n], f.g[i % n]};
        Pcg r = pcg_advance(r0, 3ull * (unsigned long long)rank[t]);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double u = pcg_next_double(r);
            const double q = p[a] + (u - 0.5) * js;
            out[3 * rank[t] + a] = (float)(q + c[a]);
        }
    }
}

// ------------------------------------------------------------ grid_block
__global__ void __launch_bounds__(WG) k_grid_block(const double* __restrict__ xs, const double* __restrict__ ys,
                                                   const double* __restrict__ zs, long long nx, long long ny, long long nz,
                                                   double* __restrict__ out) {
    const long long i = (long long)blockIdx.x * WG + threadIdx.x;
    if (i >= nx * ny * nz) return;
    out[3 * i] = xs[i / (ny * nz)];
    out[3 * i + 1] = ys[(i / nz) % ny];
    out[3 * i + 2] = zs[i % nz];
}

// ------------------------------------------------------------ fill_interior
// sim-frame position of point p on axis a: (w - c) * scale + 0.5
// (orange_setup.py:74), in the type CT NumPy's promotion gives the caller's
// inputs (fp32 for fp32 points, centre and scale), from points of type W.
template <typename W, typename CT>
__device__ __forceinline__ CT fill_sim(const W* __restrict__ w, long long p, int a, const CT c[3], CT scale) {
#pragma clang fp contract(off)
    return ((CT)w[3 * p + a] - c[a]) * scale + CT(0.5);
}

// Per-work-group minima of the sim-frame positions (exact in any order).
template <typename W, typename CT>
__global__ void __launch_bounds__(WG) k_fill_min(const W* __restrict__ w, long long n, double c0, double c1, double c2,
                                                 double scale, double* __restrict__ part) {
    const CT c[3] = {(CT)c0, (CT)c1, (CT)c2};
    double m[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    for (long long p = (long long)blockIdx.x * WG + threadIdx.x; p < n; p += (long long)gridDim.x * WG)
#pragma unroll
        for (int a = 0; a < 3; ++a) m[a] = fmin(m[a], (double)fill_sim<W, CT>(w, p, a, c, (CT)scale));
    __shared__ double s[3][WG / 64];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int o = 32; o > 0; o >>= 1) m[a] = fmin(m[a], __shfl_xor(m[a], o));
        if ((threadIdx.x & 63) == 0) s[a][threadIdx.x >> 6] = m[a];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = s[threadIdx.x][0];
        for (int k = 1; k < WG / 64; ++k) t = fmin(t, s[threadIdx.x][k]);
        part[3 * blockIdx.x + threadIdx.x] = t;
    }
}

// Voxel of point p: floor((sim - lo) / dx) per axis (orange_setup.py:78),
// in CT (dx rounded to CT first, as NumPy does with a Python float).
template <typename W, typename CT>
__device__ __forceinline__ long long fill_voxel(const W* __restrict__ w, long long p, int a, const CT c[3], CT scale,
                                                const CT lo[3], CT dx) {
#pragma clang fp contract(off)
    return (long long)floor((fill_sim<W, CT>(w, p, a, c, scale) - lo[a]) / dx);
}

struct FillFrame {
    double c[3], scale, lo[3], dx;
};

// Per-work-group maxima of the voxel indices.
template <typename W, typename CT>
__global__ void __launch_bounds__(WG) k_fill_max(const W* __restrict__ w, long long n, FillFrame f, long long* __restrict__ part) {
    const CT c[3] = {(CT)f.c[0], (CT)f.c[1], (CT)f.c[2]}, lo[3] = {(CT)f.lo[0], (CT)f.lo[1], (CT)f.lo[2]};
    long long m[3] = {0, 0, 0};
    for (long long p = (long long)blockIdx.x * WG + threadIdx.x; p < n; p += (long long)gridDim.x * WG)
#pragma unroll
        for (int a = 0; a < 3; ++a) m[a] = max(m[a], fill_voxel<W, CT>(w, p, a, c, (CT)f.scale, lo, (CT)f.dx));
    __shared__ long long s[3][WG / 64];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int o = 32; o > 0; o >>= 1) m[a] = max(m[a], __shfl_xor(m[a], o));
        if ((threadIdx.x & 63) == 0) s[a][threadIdx.x >> 6] = m[a];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        long long t = s[threadIdx.x][0];
        for (int k = 1; k < WG / 64; ++k) t = max(t, s[threadIdx.x][k]);
        part[3 * blockIdx.x + threadIdx.x] = t;
    }
}

struct Dims {
    long long x, y, z;
    __device__ __forceinline__ long long count() const { return x * y * z; }
};

template <typename W, typename CT>
__global__ void __launch_bounds__(WG) k_fill_occ(const W* __restrict__ w, long long n, FillFrame f, Dims D,
                                                 unsigned char* __restrict__ occ) {
    const long long p = (long long)blockIdx.x * WG + threadIdx.x;
    if (p >= n) return;
    const CT c[3] = {(CT)f.c[0], (CT)f.c[1], (CT)f.c[2]}, lo[3] = {(CT)f.lo[0], (CT)f.lo[1], (CT)f.lo[2]};
    const CT sc = (CT)f.scale, dx = (CT)f.dx;
    const long long i = fill_voxel<W, CT>(w, p, 0, c, sc, lo, dx), j = fill_voxel<W, CT>(w, p, 1, c, sc, lo, dx),
                    k = fill_voxel<W, CT>(w, p, 2, c, sc, lo, dx);
    // lo and dims from pmpm_fill_bounds put every point inside; others (a
    // caller's mismatched frame, NaN points) are dropped, never written
    if (i < 0 || j < 0 || k < 0 || i >= D.x || j >= D.y || k >= D.z) return;
    occ[(i * D.y + j) * D.z + k] = 1;
}

// The 6-neighbourhood plus the voxel itself (the cross structuring element
// of SciPy's generate_binary_structure(3, 1)), voxels outside the array
// reading `outside`.  DILATE: any of them set; else (erode): all set.
template <bool DILATE>
__global__ void __launch_bounds__(WG) k_fill_morph(const unsigned char* __restrict__ src, Dims D, unsigned char outside,
                                                   unsigned char* __restrict__ dst) {
    const long long v = (long long)blockIdx.x * WG + threadIdx.x;
    if (v >= D.count()) return;
    const long long i = v / (D.y * D.z), j = (v / D.z) % D.y, k = v % D.z;
    auto at = [&](long long a, long long b, long long c) -> bool {
        if (a < 0 || b < 0 || c < 0 || a >= D.x || b >= D.y || c >= D.z) return outside != 0;
        return src[(a * D.y + b) * D.z + c] != 0;
    };
    const bool nb[7] = {at(i, j, k), at(i - 1, j, k), at(i + 1, j, k), at(i, j - 1, k), at(i, j + 1, k), at(i, j, k - 1),
                        at(i, j, k + 1)};
    bool r = nb[0];
#pragma unroll
    for (int t = 1; t < 7; ++t) r = DILATE ? (r || nb[t]) : (r && nb[t]);
    dst[v] = r ? 1 : 0;
}

// One sweep of the flood from the array's faces through the background
// (SciPy's binary_fill_holes: a dilation from an empty start with border
// value 1, restricted to the background, to convergence): a background voxel
// becomes reached when a 6-neighbour is reached or lies outside the array.
// Updates in place (the set only grows, so reading a neighbour's new or old
// value both converge to the same set); *changed set when a voxel joined.
__global__ void __launch_bounds__(WG) k_fill_flood(const unsigned char* __restrict__ solid, Dims D,
                                                   unsigned char* reach, int* __restrict__ changed) {
    const long long v = (long long)blockIdx.x * WG + threadIdx.x;
    bool joined = false;
    if (v < D.count() && solid[v] == 0 && reach[v] == 0) {
        const long long i = v / (D.y * D.z), j = (v / D.z) % D.y, k = v % D.z;
        auto at = [&](long long a, long long b, long long c) -> bool {
            if (a < 0 || b < 0 || c < 0 || a >= D.x || b >= D.y || c >= D.z) return true;
            return ((volatile unsigned char*)reach)[(a * D.y + b) * D.z + c] != 0;
        };
        if (at(i - 1, j, k) || at(i + 1, j, k) || at(i, j - 1, k) || at(i, j + 1, k) || at(i, j, k - 1) || at(i, j, k + 1)) {
            reach[v] = 1;
            joined = true;
        }
    }
    if (__any(joined) && (threadIdx.x & 63) == 0) changed[0] = 1;
}

__global__ void k_fill_flag_clear(int* __restrict__ f) { f[0] = 0; }

// A voxel is solid when the flood did not reach it (it is in the closed
// shell or enclosed by it).
struct SolidFlag {
    const unsigned char* reach;
    __device__ __forceinline__ bool operator()(long long v) const { return reach[v] == 0; }
};

// The k-th solid voxel (C order): lo + (ijk + 0.5) dx, plus (u - 0.5) js
// with u = draws 3k..3k+2 when jittered, rounded to fp32
// (orange_setup.py:88-91).
__global__ void __launch_bounds__(WG) k_fill_emit(SolidFlag f, Dims D, const int* __restrict__ wgoff, double lo0, double lo1,
                                                  double lo2, double dx, int jitter, double js, pmpm_pcg64 seed,
                                                  float* __restrict__ out) {
#pragma clang fp contract(off)
    const long long b0 = (long long)blockIdx.x * SC_CHUNK;
    const long long N = D.count();
    bool keep[SC_ITEMS];
    long long idx[SC_ITEMS], rank[SC_ITEMS];
#pragma unroll
    for (int t = 0; t < SC_ITEMS; ++t) {
        idx[t] = b0 + (long long)t * WG + threadIdx.x;
        keep[t] = idx[t] < N && f(idx[t]);
    }
    sc_rank(keep, wgoff, rank);
    const Pcg r0 = pcg_from(seed);
    const double lo[3] = {lo0, lo1, lo2};
#pragma unroll
    for (int t = 0; t < SC_ITEMS; ++t) {
        if (!keep[t]) continue;
        const long long v = idx[t];
        const long long ijk[3] = {v / (D.y * D.z), (v / D.z) % D.y, v % D.z};
        Pcg r = r0;
        if (jitter) r = pcg_advance(r0, 3ull * (unsigned long long)rank[t]);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double q = lo[a] + ((double)ijk[a] + 0.5) * dx;
            if (jitter) q = q + (pcg_next_double(r) - 0.5) * js;
            out[3 * rank[t] + a] = (float)q;
        }
    }
}

// ------------------------------------------------------------ host side
// Compaction scratch: nwg chunk counts (then offsets) and the total.
struct ScScratch {
    int* wgc;
    long long* total;
};
inline size_t sc_bytes(long long N) { return al((size_t)((N + SC_CHUNK - 1) / SC_CHUNK + 1) * sizeof(int)) + al(sizeof(long long)); }
inline ScScratch sc_scratch(char* ws, long long N) {
    return ScScratch{(int*)ws, (long long*)(ws + al((size_t)((N + SC_CHUNK - 1) / SC_CHUNK + 1) * sizeof(int)))};
}

// Counts and offsets of the kept candidates; *kept (host) = their number.
template <typename Flag>
int sc_prepare(long long N, const Flag& f, const ScScratch& s, long long* kept, hipStream_t st) {
    *kept = 0;
    if (N == 0) return PMPM_OK;
    const long long nwg = (N + SC_CHUNK - 1) / SC_CHUNK;
    hipLaunchKernelGGL(k_sc_count<Flag>, dim3((unsigned)nwg), dim3(WG), 0, st, N, f, s.wgc);
    hipLaunchKernelGGL(k_sc_scan, dim3(1), dim3(WG), 0, st, s.wgc, nwg, s.total);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(kept, s.total, sizeof(long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return PMPM_OK;
}

int launch_ball(const double* g, long long n, double R, double js, const double* c, const pmpm_pcg64& rng, float* out,
                long long cap, long long* count, char* ws, hipStream_t st) {
    const long long N = n * n * n;
    const BallFlag f{g, n, R};
    const ScScratch s = sc_scratch(ws, N);
    long long kept = 0;
    int rc = sc_prepare(N, f, s, &kept, st);
    if (rc) return rc;
    *count = kept;
    if (kept > cap) return fail(PMPM_ESIZE, "%soutput holds too few points, need %lld", "", kept);
    if (kept == 0) return PMPM_OK;
    hipLaunchKernelGGL(k_ball_emit, dim3((unsigned)((N + SC_CHUNK - 1) / SC_CHUNK)), dim3(WG), 0, st, f, N, (const int*)s.wgc, js,
                       c[0], c[1], c[2], rng, out);
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}

constexpr int FILL_PARTS = 1024;      // work-groups of the bound reductions
constexpr int FILL_SWEEPS = 32;       // flood sweeps between convergence checks

template <typename W, typename CT>
int launch_fill_bounds(const W* w, long long n, const double c[3], double scale, double dx, double lo[3], long long dims[3],
                       char* ws, hipStream_t st) {
    const unsigned G = (unsigned)std::min<long long>(FILL_PARTS, (n + WG - 1) / WG);
    double* pmin = (double*)ws;
    long long* pmax = (long long*)(ws + al(3 * FILL_PARTS * sizeof(double)));
    std::vector<double> hmin(3 * G);
    std::vector<long long> hmax(3 * G);
    hipLaunchKernelGGL((k_fill_min<W, CT>), dim3(G), dim3(WG), 0, st, w, n, c[0], c[1], c[2], scale, pmin);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hmin.data(), pmin, 3 * G * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int a = 0; a < 3; ++a) {
        lo[a] = hmin[a];
        for (unsigned b = 1; b < G; ++b) lo[a] = std::min(lo[a], hmin[3 * b + a]);
    }
    FillFrame f{{c[0], c[1], c[2]}, scale, {lo[0], lo[1], lo[2]}, dx};
    hipLaunchKernelGGL((k_fill_max<W, CT>), dim3(G), dim3(WG), 0, st, w, n, f, pmax);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hmax.data(), pmax, 3 * G * sizeof(long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int a = 0; a < 3; ++a) {
        long long m = hmax[a];
        for (unsigned b = 1; b < G; ++b) m = std::max(m, hmax[3 * b + a]);
        dims[a] = m + 3;
    }
    return PMPM_OK;
}

// Scratch of the interior fill: three voxel byte grids (marks / closed
// shell, the dilation, the flood), the compaction scratch, the flood flag.
struct FillScratch {
    unsigned char *occ, *tmp, *reach;
    ScScratch sc;
    int* flag;
};
inline size_t fill_bytes(long long nvox) { return 3 * al((size_t)nvox) + sc_bytes(nvox) + al(sizeof(int)); }
inline FillScratch fill_scratch(char* ws, long long nvox) {
    FillScratch s;
    s.occ = (unsigned char*)ws;
    s.tmp = s.occ + al((size_t)nvox);
    s.reach = s.tmp + al((size_t)nvox);
    char* q = (char*)(s.reach + al((size_t)nvox));
    s.sc = sc_scratch(q, nvox);
    s.flag = (int*)(q + sc_bytes(nvox));
    return s;
}

template <typename W, typename CT>
int launch_fill_interior(const W* w, long long n, const double c[3], double scale, double dx, const double lo[3],
                         const long long dims[3], double js, const pmpm_pcg64* rng, float* out, long long cap,
                         long long* count, char* ws, hipStream_t st) {
    const Dims D{dims[0], dims[1], dims[2]};
    const long long nvox = dims[0] * dims[1] * dims[2];
    const FillScratch s = fill_scratch(ws, nvox);
    const unsigned VG = (unsigned)((nvox + WG - 1) / WG);
    const unsigned ZG = (unsigned)std::max<long long>(1, (nvox / 16 + WG - 1) / WG);
    hipLaunchKernelGGL(k_zero_bytes, dim3(ZG), dim3(WG), 0, st, s.occ, nvox);
    hipLaunchKernelGGL(k_zero_bytes, dim3(ZG), dim3(WG), 0, st, s.reach, nvox);
    const FillFrame f{{c[0], c[1], c[2]}, scale, {lo[0], lo[1], lo[2]}, dx};
    if (n > 0) hipLaunchKernelGGL((k_fill_occ<W, CT>), dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, st, w, n, f, D, s.occ);
    // binary closing: dilation, then erosion, both with 0 outside the array
    hipLaunchKernelGGL(k_fill_morph<true>, dim3(VG), dim3(WG), 0, st, (const unsigned char*)s.occ, D, (unsigned char)0, s.tmp);
    hipLaunchKernelGGL(k_fill_morph<false>, dim3(VG), dim3(WG), 0, st, (const unsigned char*)s.tmp, D, (unsigned char)0, s.occ);
    HIPCHK(hipGetLastError());
    // flood from the faces to convergence: every productive sweep adds a
    // voxel, so nvox sweeps bound the loop
    for (long long done = 0;; done += FILL_SWEEPS) {
        if (done > nvox + FILL_SWEEPS) return fail(PMPM_EHIP, "%sflood fill did not converge", "");
        hipLaunchKernelGGL(k_fill_flag_clear, dim3(1), dim3(1), 0, st, s.flag);
        for (int k = 0; k < FILL_SWEEPS; ++k)
            hipLaunchKernelGGL(k_fill_flood, dim3(VG), dim3(WG), 0, st, (const unsigned char*)s.occ, D, s.reach, s.flag);
        HIPCHK(hipGetLastError());
        int changed = 0;
        HIPCHK(hipMemcpyAsync(&changed, s.flag, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (!changed) break;
    }
    const SolidFlag sf{s.reach};
    long long kept = 0;
    int rc = sc_prepare(nvox, sf, s.sc, &kept, st);
    if (rc) return rc;
    *count = kept;
    if (kept > cap) return fail(PMPM_ESIZE, "%soutput holds too few points, need %lld", "", kept);
    if (kept == 0) return PMPM_OK;
    const pmpm_pcg64 zero = {0, 0, 0, 0};
    hipLaunchKernelGGL(k_fill_emit, dim3((unsigned)((nvox + SC_CHUNK - 1) / SC_CHUNK)), dim3(WG), 0, st, sf, D,
                       (const int*)s.sc.wgc, lo[0], lo[1], lo[2], dx, rng ? 1 : 0, js, rng ? *rng : zero, out);
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}
