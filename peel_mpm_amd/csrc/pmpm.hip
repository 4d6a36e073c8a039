// libpmpm: MI355X (gfx950) kernels of the AnisoMPM substep and their C ABI.
//
// One step (anisompm_py/src/anisompm.py:593-803) runs as:
//   k_bin_count   key = cell (bucket-major) of each particle, its slot in the
//                 cell (wave-aggregated atomics over runs of equal keys),
//                 flags on the touched 4x4x4-cell buckets
//   k_bucket_tiles / k_bucket_apply
//                 bucket totals of the flagged buckets, exclusive scan over
//                 buckets: bucket and cell offsets, ordered list of
//                 non-empty buckets, marks on the grid blocks their
//                 stencils reach
//   k_bin_scatter perm[start[key]+slot] = particle and its inverse inv; its
//                 first work-groups list the marked grid blocks (clearing
//                 the marks)
//                 (+ k_bin_stabilize: ascending particle ids per cell)
//   k_const       per particle, in particle order: stress/damage/degradation
//                 in registers, then the particle's P2G record (x, affine
//                 momentum, mass*v, mass) and d at its sorted position
//   k_p2g         per non-empty bucket: quadratic momentum and cubic damage
//                 scatter of the bucket's records into 7x7x7-node LDS arenas
//                 (lane per cell, plain read-modify-write, two particles per
//                 update), stored to the bucket's HBM arena slot with plain
//                 coalesced stores -- no atomics
//   [PCG damage mode: k_pcg_* between binning and k_const]
//   k_grid        per marked 4x4x4 node block: each node sums the <= 8
//                 bucket arenas that cover it in a fixed order, then
//                 velocity, gravity, damping, colliders, cubic normalisation
//   k_g2p         per non-empty bucket: LDS-staged arena gather, velocity,
//                 affine C, cubic Laplacian, advection, F update, clamp
// Particle arrays stay in the caller's order; kernels visit them through perm.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pmpm.h"
#include "pmpm_device.h"

using namespace pmpm;

namespace {

constexpr int BLK = 4;                 // cells per bucket edge (= nodes per grid-block edge)
constexpr int AR = BLK + 3;            // arena edge: nodes [4b-1, 4b+5]
constexpr int AR3 = AR * AR * AR;      // 343
// Padded LDS arenas of the lane-per-cell scatters: node (x, y, z) at entry
// x PAD_XS + 7 y + z, and lane l owning cell (x, y, z) = ((l >> 2) & 3,
// l >> 4, l & 3) of its bucket (pad_cell).  The 8 lanes of one 128-byte LDS
// cycle of a 16-byte access (fixed y, x in {2k, 2k+1}, z 0..3) then reach 8
// distinct entries mod 8, and the 16 lanes of an 8-byte access 16 distinct
// entries mod 16, so a frame offset's update of the lanes' own cells runs
// without bank conflicts (the dense stride of 49, in cell order, conflicts
// two ways).
constexpr int PAD_XS = AR * AR + 3;
constexpr int PAD3 = AR * PAD_XS;        // entries of one padded arena
__host__ __device__ constexpr int pad_cell(int l) { return (((l >> 2) & 3) << 4) | ((l >> 4) << 2) | (l & 3); }
__host__ __device__ constexpr int pad_entry(int node) { return (node / (AR * AR)) * PAD_XS + node % (AR * AR); }
constexpr int ARZ8 = 2 * BLK + 3;      // z edge of an 8-cell-long block arena: nodes [8c-1, 8c+9]
constexpr int AR3B = AR * AR * ARZ8;   // 539
constexpr int WG = 256;                // threads per work-group
constexpr int PERSIST_WG = 2048;       // work-groups of the persistent bucket kernels
constexpr int BTILE = WG;              // buckets per bucket-scan tile
constexpr int P2G_GRID = 2048;         // work-groups of the persistent P2G kernel
constexpr int GRID_WAVES = 8192;       // waves of the grid kernel
constexpr int GRID_WG_COUNT = GRID_WAVES / 4;   // its 256-thread work-groups
constexpr int G2P_PREFETCH = 2;        // particles per G2P thread fetched before the node staging
constexpr int P2G_REC = 16;            // values per P2G record: x(3), affine(9), mass*v(3), mass
constexpr int GMW = 5;                 // values per node kept for the adjoint: m, sum w, mv/m (3)
// PCG damage scalars (doubles in the pcgr section after the partial dots)
enum { PCG_RZ = 0, PCG_BB, PCG_ALPHA, PCG_BETA, PCG_REL, PCG_ITER, PCG_DONE, PCG_NSCAL = 8 };

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, const char* a = "", long long b = 0) {
    snprintf(g_err, sizeof g_err, fmt, a, b);
    return code;
}
#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) return fail(PMPM_EHIP, "%s (code %lld)", hipGetErrorString(e_), (long long)e_); \
    } while (0)

inline int nblk(int ng) { return (ng + BLK - 1) / BLK; }

// Optional live timing: HIP events recorded on the launch stream around each
// kernel group, summed on read.  Groups: 0 binning, 1 p2g, 2 grid, 3 g2p,
// 4 the PCG damage solve (drive, assembly, iterations, interpolation),
// 5 one CG iteration of it (operator, fold, reductions, update).
constexpr int NTIMED = 7;
struct Timing {
    bool on = false;
    std::vector<hipEvent_t> ev[NTIMED];   // start/stop pairs
    std::vector<hipEvent_t> pool;
    hipEvent_t get() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    void reset() {
        for (auto& v : ev) { for (auto e : v) pool.push_back(e); v.clear(); }
    }
};
Timing g_timing;
thread_local bool t_untimed = false;    // set while a no-allocation step enqueues (no event creation)
struct TimedScope {
    int group;
    hipStream_t st;
    hipEvent_t stop = nullptr;
    TimedScope(int g, hipStream_t s) : group(g), st(s) {
        if (!g_timing.on || t_untimed) return;
        hipEvent_t start = g_timing.get();
        stop = g_timing.get();
        if (!start || !stop) { stop = nullptr; return; }
        // a failed record surfaces when the interval is read (pmpm_timing_read)
        (void)hipEventRecord(start, st);
        g_timing.ev[group].push_back(start);
    }
    ~TimedScope() {
        if (!stop) return;
        (void)hipEventRecord(stop, st);
        g_timing.ev[group].push_back(stop);
    }
};
inline size_t al(size_t v) { return (v + 255) & ~size_t(255); }

// Process-wide launch shape (pmpm_set_launch_config); defaults below.
pmpm_launch_config g_launch = {PERSIST_WG, GRID_WG_COUNT, 2, 256, 2, 256, 4, 1, 1};

// Workspace layout (byte offsets).  The step keeps count, bflag, mark and
// misc at zero between steps: whoever reads them clears them.  count/start
// are per cell (bucket-major; start is kept for the cells of non-empty
// buckets), btot/bstart/bflag/alist/slotof/mark/blist per bucket or grid
// block.  arena holds one 6 x 343 slot per non-empty bucket (m, mv, sum w d,
// sum w), in active-list order.
//
// Every section whose size does not depend on the particle count comes
// first, so a workspace laid out for max_n particles serves any n <= max_n
// at the same offsets for them -- in particular for the zero-state sections.
// The sections sized by n (arena, then the per-particle ones) follow; the
// step rewrites them in full every call.
struct Layout {
    size_t arena, gv, active, count, count2, start, btot, bstart, bflag, bflag2, tiles, mark, alist, slotof, blist, key, slot,
        perm, inv, rec, dsrt, misc, stats, pcgd, pcgr, gmw, total;
    long long NB, NBs, NC, ntiles, G, Gs, max_slots;
    // ns scenes, each with its own n_grid^3 grid and buckets (scene-major)
    Layout(int64_t n, int ng, int dtype, int ns = 1) {
        size_t ts = dtype == PMPM_F64 ? 8 : 4;
        if (ns < 1) ns = 1;
        const size_t np = (size_t)(n > 0 ? n : 1);
        Gs = (long long)ng * ng * ng;
        G = Gs * ns;
        NBs = (long long)nblk(ng) * nblk(ng) * nblk(ng);
        NB = NBs * ns;
        NC = NB * 64;                            // cells, bucket-major
        ntiles = (NB + BTILE - 1) / BTILE;       // bucket-scan tiles
        max_slots = NB < (long long)np ? NB : (long long)np;
        size_t o = 0;
        // ---- sections independent of n
        misc = o;   o = al(o + 64);          // [0] n_active buckets, [1] n marked blocks, [2] grid partials,
                                             // [3] some stencil leaves the grid (this step), [4] its raw flag,
                                             // [5] the raw flag of the carried count (G2P's)
        count = o;  o = al(o + (size_t)NC * 4);
        bflag = o;  o = al(o + (size_t)NB * 4);
        count2 = o; o = al(o + (size_t)NC * 4);     // the carried count: G2P's, for the next step
        bflag2 = o; o = al(o + (size_t)NB * 4);
        mark = o;   o = al(o + (size_t)NB * 4);
        start = o;  o = al(o + (size_t)(NC + 1) * 4);
        btot = o;   o = al(o + (size_t)NB * 4);
        bstart = o; o = al(o + (size_t)NB * 4);
        tiles = o;  o = al(o + (size_t)ntiles * 2 * 4);
        alist = o;  o = al(o + (size_t)NB * 4);
        slotof = o; o = al(o + (size_t)NB * 4);
        blist = o;  o = al(o + (size_t)NB * 4);
        stats = o;  o = al(o + (size_t)GRID_WG_COUNT * 16);   // per grid work-group: active q, c nodes
        pcgr = o;   o = al(o + (size_t)(GRID_WG_COUNT * 3 + PCG_NSCAL) * 8);   // PCG partial dots + scalars
        gv = o;     o = al(o + (size_t)G * 4 * ts);
        active = o; o = al(o + (size_t)G);
        pcgd = o;   o = al(o + (size_t)G * 2 * ts);                // PCG damage: per node (Jacobi diagonal, lumped mass)
        gmw = o;    o = al(o + (size_t)G * GMW * ts);              // per node (m, sum w, mv/m) kept for the adjoint
        // ---- sections sized by n
        // compact per-bucket arenas (343 nodes), or dense 4x4x8-cell block arenas (539 nodes)
        const size_t NB8 = (size_t)ns * nblk(ng) * nblk(ng) * ((nblk(ng) + 1) / 2);
        arena = o;  o = al(o + std::max((size_t)max_slots * AR3, NB8 * AR3B) * 6 * ts);
        key = o;    o = al(o + np * 4);
        slot = o;   o = al(o + np * 4);
        perm = o;   o = al(o + np * 4);
        inv = o;    o = al(o + np * 4);                   // sorted position of each particle
        rec = o;    o = al(o + np * P2G_REC * ts);        // P2G records, sorted order
        dsrt = o;   o = al(o + np * ts);                  // post-update d, sorted order
        total = o;
    }
    // the zero-state sections: one contiguous range at the front
    size_t zero_end() const { return start; }
};

// Workspaces set up by pmpm_workspace_init, by address: the shape each was
// laid out for.  A call checks its workspace here before it enqueues
// anything, so a buffer laid out for another grid, scene count or dtype, or
// for fewer particles, is refused instead of read at shifted offsets.
struct WsShape {
    int64_t max_n;
    int32_t ng, ns, dtype;
    size_t bytes;
};
std::mutex g_ws_mu;
std::unordered_map<uintptr_t, WsShape> g_ws;

void ws_register(const void* ws, const WsShape& s);

// The carried binning of each workspace, by address: set by a step whose
// G2P counted the next step's cells (bin_next), with the positions it wrote
// and the stream capture it was enqueued in (0: none).  A later step on the
// workspace takes it up (prebinned) only when its x is those positions, for
// the same particle count, in the same capture (or both outside one); any
// step or binning on the workspace consumes the record, since its scan
// clears the carried count.
struct Carry {
    const void* x;
    int64_t n;
    unsigned long long cap;
    bool pending;
    long long taken;                        // steps that started from a carried count
};
std::unordered_map<uintptr_t, Carry> g_carry;

unsigned long long capture_id(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    unsigned long long id = 0;
    if (hipStreamGetCaptureInfo(st, &cs, &id) != hipSuccess || cs != hipStreamCaptureStatusActive) return 0;
    return id ? id : 1;
}

// Whether this step on ws may start from the carried count (want: the
// caller vouches that x is unchanged since the step that wrote it); the
// record is consumed either way.
bool carry_take(const void* ws, bool want, const void* x, int64_t n, hipStream_t st) {
    if (!ws) return false;
    const unsigned long long cap = capture_id(st);
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_carry.find((uintptr_t)ws);
    if (it == g_carry.end()) return false;
    const Carry c = it->second;
    it->second.pending = false;
    const bool ok = want && c.pending && c.x == x && c.n == n && c.cap == cap && n > 0;
    if (ok) ++it->second.taken;
    return ok;
}

void carry_set(const void* ws, const void* x, int64_t n, hipStream_t st) {
    const unsigned long long cap = capture_id(st);
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Carry& c = g_carry[(uintptr_t)ws];
    c.x = x;
    c.n = n;
    c.cap = cap;
    c.pending = true;
}

long long carry_taken(const void* ws) {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_carry.find((uintptr_t)ws);
    return it == g_carry.end() ? 0 : it->second.taken;
}

void ws_register(const void* ws, const WsShape& s) {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    g_ws[(uintptr_t)ws] = s;
    g_carry.erase((uintptr_t)ws);           // the set-up clears the workspace
}

int ws_check(const void* ws, int64_t n, int ng, int ns, int dtype) {
    if (ns < 1) ns = 1;
    WsShape s;
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        auto it = g_ws.find((uintptr_t)ws);
        if (it == g_ws.end()) return fail(PMPM_EARG, "%sworkspace was not set up by pmpm_workspace_init", "");
        s = it->second;
    }
    if (s.ng != ng || s.ns != ns || s.dtype != dtype)
        return fail(PMPM_EARG, "%sworkspace was set up for another grid, scene count or dtype (n_grid %lld)", "", s.ng);
    if (n > s.max_n) return fail(PMPM_EARG, "%sworkspace was set up for at most %lld particles", "", (long long)s.max_n);
    return PMPM_OK;
}

// ------------------------------------------------------------------ binning
// Sort key of a particle: its cell, numbered bucket-major -- the 4x4x4-cell
// bucket (scene-major in a batch), then the cell inside the bucket:
//   key = (scene * NBs + bucket) * 64 + (cx & 3) * 16 + (cy & 3) * 4 + (cz & 3)
// with the cell c = floor(x * inv_dx) of the reference (anisompm.py:574)
// clamped to the grid's buckets.  A bucket's particles are contiguous and
// grouped by cell, which lets P2G put one cell on each lane.
template <typename T>
// off_grid: the particle's cubic stencil [c-1, c+2] (which holds its
// quadratic one) reaches past the grid on some axis.
__device__ __forceinline__ int cell_key(const T xp[3], T inv_dx, int ng, int nb, bool& off_grid) {
    int b[3], l[3];
    off_grid = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int c = floor_idx(scaled(xp[a], inv_dx), ng);
        off_grid = off_grid || c < 1 || c + 2 > ng - 1;
        b[a] = clampi(c >> 2, 0, nb - 1);
        l[a] = clampi(c - BLK * b[a], 0, BLK - 1);
    }
    return (((b[0] * nb + b[1]) * nb + b[2]) << 6) | (l[0] << 4) | (l[1] << 2) | l[2];
}

// Scene of particle p: the last s with off[s] <= p (off has ns+1 entries).
__device__ __forceinline__ int scene_of(const long long* __restrict__ off, int ns, long long p) {
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Run of equal values v among the live lanes [0, nlive) of a wavefront:
// whether this lane heads its run, the run's first lane and its length.
__device__ __forceinline__ void wave_run(int v, bool live, int lane, bool& head, int& first, int& len) {
    const int vprev = __shfl_up(v, 1);
    head = live && (lane == 0 || vprev != v);
    const unsigned long long heads = __ballot(head);
    const unsigned long long alive = __ballot(live);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1);
    first = 63 - __builtin_clzll((heads & upto) | 1ull);     // highest head <= lane
    const unsigned long long above = heads & ~upto;
    int end = above ? __builtin_ctzll(above) : 64;
    const int nlive = 64 - __builtin_clzll(alive | 1ull);   // live lanes are [0, nlive)
    if (end > nlive) end = nlive;
    len = end - first;
}

// Cell key and slot within the cell of each particle (one atomic per run
// of equal keys in the wavefront); each touched bucket is flagged in bflag
// (plain stores).  misc[4] is raised (plain store) when a particle's
// stencil reaches past the grid; k_bucket_apply publishes it as misc[3]
// for the step.
template <typename T>
__global__ void __launch_bounds__(WG) k_bin_count(const T* __restrict__ x, long long n, T inv_dx, int ng, int nb,
                                                  const long long* __restrict__ scene_off, int ns,
                                                  int* __restrict__ count, int* __restrict__ bflag, int* __restrict__ key,
                                                  int* __restrict__ slot, int* __restrict__ misc) {
    long long p = (long long)blockIdx.x * WG + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool live = p < n;
    int k = -1;
    bool off_grid = false;
    if (live) {
        const T xp[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
        k = cell_key(xp, inv_dx, ng, nb, off_grid);
        if (ns > 1) k += scene_of(scene_off, ns, p) * (nb * nb * nb * 64);
    }
    if (__ballot(off_grid) != 0 && lane == 0) misc[4] = 1;
    bool head;
    int first, len;
    wave_run(k, live, lane, head, first, len);
    int base = 0;
    if (head) base = atomicAdd(&count[k], len);
    base = __shfl(base, first);
    if (live) {
        key[p] = k;
        slot[p] = base + (lane - first);
    }
    const int kb = k >> 6, kbprev = __shfl_up(kb, 1);
    if (live && (lane == 0 || kbprev != kb)) bflag[kb] = 1;
}

// Bucket scan, pass 1: one thread per bucket, a tile of BTILE buckets per
// work-group.  A flagged bucket sums its 64 cell counts (btot); per tile:
// particles and non-empty buckets.
__global__ void __launch_bounds__(WG) k_bucket_tiles(const int* __restrict__ bflag, const int* __restrict__ count,
                                                     long long NB, int* __restrict__ btot, int* __restrict__ tiles) {
    __shared__ int s_a[WG / 64], s_b[WG / 64];
    const long long b = (long long)blockIdx.x * BTILE + threadIdx.x;
    int t = 0;
    if (b < NB && bflag[b]) {
        const int4* c4 = reinterpret_cast<const int4*>(count + (b << 6));
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int4 v = c4[i];
            t += v.x + v.y + v.z + v.w;
        }
    }
    if (b < NB) btot[b] = t;
    int a = t, c = t > 0 ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        c += __shfl_xor(c, o);
    }
    if ((threadIdx.x & 63) == 0) { s_a[threadIdx.x >> 6] = a; s_b[threadIdx.x >> 6] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        int ta = 0, tc = 0;
        for (int w = 0; w < WG / 64; ++w) { ta += s_a[w]; tc += s_b[w]; }
        tiles[2 * blockIdx.x] = ta;
        tiles[2 * blockIdx.x + 1] = tc;
    }
}

// Bucket scan, pass 2 (same tiling): each work-group sums the totals of the
// tiles before it, then scans its own buckets -- bucket starts (bstart,
// every bucket), the ordered list of non-empty buckets (alist) and their
// slots (slotof, -1 when empty), the cell starts of non-empty buckets (and
// the end of each one's last cell), marks on the grid blocks their arenas
// reach.  Clears count and bflag of the buckets it read, and the flagged
// buckets of the other count set (bflag_o, count_o; null when none), so a
// carried count that no step took up is gone.  Work-group 0 also publishes
// the off-grid flags of the count and of the carried count (misc[3] =
// misc[4] | misc[5], clearing both: a stale carried flag only makes the
// grid pass look beyond the grid needlessly) and resets misc[1]; the last
// writes misc[0] = non-empty buckets and start[NC] = particles.
__global__ void __launch_bounds__(WG) k_bucket_apply(int* __restrict__ bflag, int* __restrict__ count, const int* __restrict__ btot,
                                                     long long NB, int nb, const int* __restrict__ tiles, int* __restrict__ bstart,
                                                     int* __restrict__ alist, int* __restrict__ slotof, int* __restrict__ mark,
                                                     int* __restrict__ start, int* __restrict__ misc, int* __restrict__ bflag_o,
                                                     int* __restrict__ count_o) {
    __shared__ int s_a[WG / 64], s_b[WG / 64], s_pa, s_pb;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // prefix of the earlier tiles (wave 0)
    if (w == 0) {
        int pa = 0, pb = 0;
        for (int t = lane; t < (int)blockIdx.x; t += 64) { pa += tiles[2 * t]; pb += tiles[2 * t + 1]; }
        for (int o = 32; o > 0; o >>= 1) {
            pa += __shfl_xor(pa, o);
            pb += __shfl_xor(pb, o);
        }
        if (lane == 0) { s_pa = pa; s_pb = pb; }
    }
    const long long b = (long long)blockIdx.x * BTILE + threadIdx.x;
    const int t = b < NB ? btot[b] : 0;
    int ia = t, ib = t > 0 ? 1 : 0;
    for (int o = 1; o < 64; o <<= 1) {
        const int ta = __shfl_up(ia, o), tb = __shfl_up(ib, o);
        if (lane >= o) { ia += ta; ib += tb; }
    }
    if (lane == 63) { s_a[w] = ia; s_b[w] = ib; }
    __syncthreads();
    int pa = s_pa, pb = s_pb;
    for (int j = 0; j < w; ++j) { pa += s_a[j]; pb += s_b[j]; }
    pa += ia - t;                       // this bucket's first particle
    pb += ib - (t > 0 ? 1 : 0);         // its slot if non-empty
    if (b < NB) {
        bstart[b] = pa;
        slotof[b] = t > 0 ? pb : -1;
        if (bflag[b]) bflag[b] = 0;
        if (bflag_o && bflag_o[b]) {
            int4* c4 = reinterpret_cast<int4*>(count_o + (b << 6));
#pragma unroll
            for (int i = 0; i < 16; ++i) c4[i] = make_int4(0, 0, 0, 0);
            bflag_o[b] = 0;
        }
        if (t > 0) {
            alist[pb] = (int)b;
            int4* c4 = reinterpret_cast<int4*>(count + (b << 6));
            int4* s4 = reinterpret_cast<int4*>(start + (b << 6));
            int a = pa;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int4 v = c4[i];
                int4 st;
                st.x = a; a += v.x;
                st.y = a; a += v.y;
                st.z = a; a += v.z;
                st.w = a; a += v.w;
                s4[i] = st;
                c4[i] = make_int4(0, 0, 0, 0);
            }
            start[(b + 1) << 6] = a;
            const long long nbs = (long long)nb * nb * nb;
            const long long sc = b / nbs;
            const int loc = (int)(b - sc * nbs);
            const int bx = loc / (nb * nb), by = (loc / nb) % nb, bz = loc % nb;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        const int X = bx + dx, Y = by + dy, Z = bz + dz;
                        if (X >= 0 && X < nb && Y >= 0 && Y < nb && Z >= 0 && Z < nb)
                            mark[sc * nbs + (X * nb + Y) * nb + Z] = 1;
                    }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        misc[3] = misc[4] | misc[5];
        misc[4] = 0;
        misc[5] = 0;
        misc[1] = 0;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == WG - 1) {
        misc[0] = pb + (t > 0 ? 1 : 0);
        start[NB << 6] = pa + t;
    }
}

// Compact the marked grid blocks into blist (any order) and clear the marks
// (work-group wg of the launch).
__device__ __forceinline__ void mark_compact(long long wg, int* __restrict__ mark, long long NB, int* __restrict__ blist,
                                             int* __restrict__ misc) {
    const long long b = wg * WG + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool mk = b < NB && mark[b] != 0;
    if (mk) mark[b] = 0;
    const unsigned long long bal = __ballot(mk);
    if (!bal) return;
    int base = 0;
    if (lane == __builtin_ctzll(bal)) base = atomicAdd(&misc[1], __popcll(bal));
    base = __shfl(base, __builtin_ctzll(bal));
    if (mk) blist[base + __popcll(bal & ((1ull << lane) - 1))] = (int)b;
}

// Cell starts of the empty buckets (all their cells start where the bucket
// does) -- only for pmpm_bin's dense cell_start; the step reads the starts
// of non-empty buckets' cells alone.
__global__ void __launch_bounds__(WG) k_dense_starts(const int* __restrict__ bstart, const int* __restrict__ slotof,
                                                     long long NC, int* __restrict__ start) {
    const long long c = (long long)blockIdx.x * WG + threadIdx.x;
    if (c < NC && slotof[c >> 6] < 0) start[c] = bstart[c >> 6];
}

// One launch for two independent passes: the first nmark work-groups compact
// the marked grid blocks, the others place each particle at start[key] +
// slot (perm and its inverse inv).
__global__ void __launch_bounds__(WG) k_bin_scatter(long long n, const int* __restrict__ key, const int* __restrict__ slot,
                                                    const int* __restrict__ start, int* __restrict__ perm,
                                                    int* __restrict__ inv, long long nmark, int* __restrict__ mark,
                                                    long long NB, int* __restrict__ blist, int* __restrict__ misc) {
    if ((long long)blockIdx.x < nmark) {
        mark_compact(blockIdx.x, mark, NB, blist, misc);
        return;
    }
    const long long p = ((long long)blockIdx.x - nmark) * WG + threadIdx.x;
    if (p < n) {
        const int pos = start[key[p]] + slot[p];
        perm[pos] = (int)p;
        inv[p] = pos;
    }
}

// Sort the particle ids of n <= 16 entries ascending: an odd-even
// transposition network in registers.
__device__ __forceinline__ void sort16(int v[16]) {
#pragma unroll
    for (int round = 0; round < 16; ++round)
#pragma unroll
        for (int i = round & 1; i + 1 < 16; i += 2) {
            const int lo = v[i] < v[i + 1] ? v[i] : v[i + 1];
            const int hi = v[i] < v[i + 1] ? v[i + 1] : v[i];
            v[i] = lo;
            v[i + 1] = hi;
        }
}

// Sort each cell's particle ids ascending (then perm is a stable argsort of
// key): one work-group of 64 threads per non-empty bucket.  The bucket's
// ids are read into LDS with coalesced loads, lane l sorts the ids of cell
// l there (a register network up to 16, insertion sort beyond), and the
// bucket is written back with coalesced stores, with inv of each of its
// particles.  A bucket of more than STAB_CAP particles sorts in place in
// global memory instead.
constexpr int STAB_CAP = 2048;
constexpr int STAB_WGS = 8192;          // work-groups of k_bin_stabilize (grid-stride beyond)
__global__ void __launch_bounds__(64) k_bin_stabilize(const int* __restrict__ start, const int* __restrict__ alist,
                                                      const int* __restrict__ misc, int* __restrict__ perm,
                                                      int* __restrict__ inv) {
    __shared__ int s_id[STAB_CAP];
    const int nact = misc[0];
    const int lane = threadIdx.x;
    for (int ai = blockIdx.x; ai < nact; ai += gridDim.x) {
        const long long c = ((long long)alist[ai] << 6) + lane;
        const int s0 = start[c], s1 = start[c + 1], n = s1 - s0;
        const int b0 = __shfl(s0, 0), b1 = __shfl(s1, 63);
        const int nbk = b1 - b0;
        if (nbk <= STAB_CAP) {
            for (int i = lane; i < nbk; i += 64) s_id[i] = perm[b0 + i];
            __syncthreads();
            int* seg = s_id + (s0 - b0);
            if (n > 1 && n <= 16) {
                int v[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) v[i] = i < n ? seg[i] : 0x7fffffff;
                sort16(v);
#pragma unroll
                for (int i = 0; i < 16; ++i)
                    if (i < n) seg[i] = v[i];
            } else if (n > 16) {
                for (int i = 1; i < n; ++i) {
                    const int e = seg[i];
                    int j = i - 1;
                    while (j >= 0 && seg[j] > e) {
                        seg[j + 1] = seg[j];
                        --j;
                    }
                    seg[j + 1] = e;
                }
            }
            __syncthreads();
            for (int i = lane; i < nbk; i += 64) {
                const int id = s_id[i];
                perm[b0 + i] = id;
                inv[id] = b0 + i;
            }
            __syncthreads();
            continue;
        }
        if (n <= 1) continue;
        if (n <= 16) {
            int v[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) v[i] = i < n ? perm[s0 + i] : 0x7fffffff;
            sort16(v);
#pragma unroll
            for (int i = 0; i < 16; ++i)
                if (i < n) {
                    perm[s0 + i] = v[i];
                    inv[v[i]] = s0 + i;
                }
            continue;
        }
        for (int i = s0 + 1; i < s1; ++i) {
            const int e = perm[i];
            int j = i - 1;
            while (j >= s0 && perm[j] > e) {
                perm[j + 1] = perm[j];
                --j;
            }
            perm[j + 1] = e;
        }
        for (int i = s0; i < s1; ++i) inv[perm[i]] = i;
    }
}

// ------------------------------------------------------------------ step kernels
template <typename T> struct StepArgs {
    // particles in
    const T *x, *v, *F, *C, *d, *lap, *D_hist, *vol0, *mass, *mu, *lam, *sigma_c, *eta, *zeta, *l0, *residual, *A0,
        *n_if;
    const uint8_t *allow, *imask, *bc_id;
    // uniform material fields (bit k of umask: field k is uval[k], A0 is uA0)
    unsigned umask;
    T uval[9];
    M3<T> uA0;
    // particles out
    T *ox, *ov, *oF, *oC, *olap, *od, *oD, *ophi, *ophs, *otau, *oJ;
    // workspace
    T *arena, *gv;
    T* gmw;                     // optional: per node (m, sum w, mv/m) of the grid pass
    T* pdiag;                   // PCG damage: per node (Jacobi diagonal, lumped mass)
    double *ppart, *pscal;      // PCG damage: per-work-group partial dots, scalars
    uint8_t* active;
    const int *start, *alist, *slotof, *blist, *perm, *misc, *key, *inv;
    T *rec, *dsrt;              // P2G records and post-update d at sorted positions
    int rec_pp;                 // 1: rec and dsrt in particle order instead (P2G reads them through perm)
    // the carried binning: with bin_next, G2P takes the next step's cell keys
    // and slots (ckey, cslot) and counts into the carried set (ccount, cbflag,
    // cmisc[5]) from the positions it writes
    int *ckey, *cslot, *ccount, *cbflag, *cmisc;
    int bin_next;
    unsigned long long* stats;
    // scalars
    long long n, NBs, Gs;
    long long NBs8;             // 4x4x8-cell blocks per scene (nb x nb x ceil(nb/2))
    int ng, nb, nscenes;
    int cb;                     // cells along z per P2G block (4 or 8; launch shape)
    const uint8_t* scene_dmg;   // per-scene damage gate, or null: all scenes run damage
    T dx, inv_dx, dt, lim, pad_lo, pad_hi, fc_lo, fc_hi, dt_dmg, damp, stress_k, c_scale, lap_scale;
    T dtg[3];
    int use_fclamp, export_active, n_col, n_bc;
    double pcg_tol;
    int pcg_max;
    pmpm_collider col[PMPM_MAX_COLLIDERS];
    T bcv[PMPM_MAX_VEL_BC][3];
    // optional device copy of the colliders and BC velocities, read at kernel
    // time instead of the by-value copies above (a replayed graph then sees
    // values the host wrote before the replay)
    const pmpm_boundary_table* btab;
};

// Prescribed velocity of a particle with velocity-BC slot bc (1-based; 0: none).
template <typename T> __device__ __forceinline__ void bc_velocity(const StepArgs<T>& A, int bc, T vv[3]) {
    if (bc <= 0) return;
    if (A.btab) {
        const double* t = A.btab->vel_bc[bc - 1];
        vv[0] = T(t[0]); vv[1] = T(t[1]); vv[2] = T(t[2]);
        return;
    }
    // compile-time indices only, so the by-value kernel arguments are never
    // copied to scratch for a dynamic index
#pragma unroll
    for (int k = 0; k < PMPM_MAX_VEL_BC; ++k)
        if (bc == k + 1) { vv[0] = A.bcv[k][0]; vv[1] = A.bcv[k][1]; vv[2] = A.bcv[k][2]; }
}

// Persistent-bucket schedule: the non-empty bucket list is cut into 8
// contiguous parts, one per XCD (work-groups are dealt to the 8 XCDs round
// robin), so neighbouring buckets -- which share halo nodes -- run on the
// same XCD and hit its L2.
struct BucketIter {
    int i, end, step;
    __device__ BucketIter(int nact) {
        const int G = gridDim.x, xcd = blockIdx.x & 7, lw = blockIdx.x >> 3, per = G >> 3;
        const long long lo = (long long)nact * xcd / 8, hi = (long long)nact * (xcd + 1) / 8;
        i = (int)lo + lw;
        end = (int)hi;
        step = per;
    }
};

// Material field K of particle p: the uniform value when the caller passed
// one, else the array's entry.
template <int K, typename T> __device__ __forceinline__ T fld(const StepArgs<T>& A, const T* __restrict__ a, long long p) {
    return ((A.umask >> K) & 1u) ? A.uval[K] : a[p];
}

template <typename T> __device__ __forceinline__ M3<T> load_m3(const T* __restrict__ a, long long p) {
    M3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = a[9 * p + 3 * i + j];
    return r;
}
template <typename T> __device__ __forceinline__ void store_m3(T* __restrict__ a, long long p, const M3<T>& r) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) a[9 * p + 3 * i + j] = r.m[i][j];
}

// One particle's constitutive inputs, loaded together (load_const_in) at the
// top of a kernel so that they are all in flight at once, ahead of the
// register-bound eigen and stress arithmetic that uses them.
template <typename T> struct ConstIn {
    M3<T> A0;
    T d, mu, lam, res, sc, zeta, Dh, l0, lap, eta;
    bool allow, iface;
};
template <typename T, bool DAMAGE>
__device__ __forceinline__ void load_const_in(const StepArgs<T>& A, long long p, ConstIn<T>& q) {
    q.d = A.d[p];
    q.mu = fld<PMPM_U_MU>(A, A.mu, p);
    q.lam = fld<PMPM_U_LAM>(A, A.lam, p);
    q.res = fld<PMPM_U_RESIDUAL>(A, A.residual, p);
    q.iface = A.imask[p] != 0;
    q.Dh = A.D_hist[p];
    if constexpr (DAMAGE) {
        q.A0 = ((A.umask >> PMPM_U_A0) & 1u) ? A.uA0 : load_m3(A.A0, p);
        q.sc = fld<PMPM_U_SIGMA_C>(A, A.sigma_c, p);
        q.zeta = fld<PMPM_U_ZETA>(A, A.zeta, p);
        q.l0 = fld<PMPM_U_L0>(A, A.l0, p);
        q.lap = A.lap[p];
        q.eta = fld<PMPM_U_ETA>(A, A.eta, p);
        q.allow = A.allow[p] != 0;
    }
}

// Degraded Kirchhoff stress from the undegraded deviatoric and volumetric
// parts: g(d) on the deviatoric part, on the volumetric part only in
// tension (anisompm.py:687-691), and for interface particles the split by
// the deformed normal F n0 (anisompm.py:190-227, :697-703).  gate: the
// tension gate (the mode-I degradation only where the normal traction
// pulls); the step always gates, as the reference's solver does.
template <typename T>
__device__ __forceinline__ M3<T> degrade(T res, const M3<T>& F, T d, T J, const M3<T>& tdev, T tvol, bool iface,
                                         const T n0[3], bool gate = true) {
    const T om = T(1) - d;
    const T g = om * om * (T(1) - res) + res;
    const T vdeg = J >= T(1) ? g : T(1);
    M3<T> tau;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) tau.m[i][j] = g * tdev.m[i][j] + (i == j ? vdeg * tvol : T(0));
    if (iface) {
        // projection split by the deformed normal (anisompm.py:190-227, :697-703)
        M3<T> tc = tdev;
#pragma unroll
        for (int i = 0; i < 3; ++i) tc.m[i][i] += tvol;
        T m[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) m[i] = F.m[i][0] * n0[0] + F.m[i][1] * n0[1] + F.m[i][2] * n0[2];
        const T mn = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) + T(1e-9);
#pragma unroll
        for (int i = 0; i < 3; ++i) m[i] = m[i] / mn;
        M3<T> P, Q;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                P.m[i][j] = m[i] * m[j];
                Q.m[i][j] = (i == j ? T(1) : T(0)) - P.m[i][j];
            }
        M3<T> Pt = mul(P, tc), Qt = mul(Q, tc);
        M3<T> tperp = mul(Pt, P), tpq = mul(Pt, Q), tqp = mul(Qt, P), tpar = mul(Qt, Q);
        T tm[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) tm[i] = tc.m[i][0] * m[0] + tc.m[i][1] * m[1] + tc.m[i][2] * m[2];
        const T snn = m[0] * tm[0] + m[1] * tm[1] + m[2] * tm[2];
        const T gI = (snn > T(0) || !gate) ? g : T(1);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) tau.m[i][j] = tpar.m[i][j] + gI * tperp.m[i][j] + g * (tpq.m[i][j] + tqp.m[i][j]);
    }
        return tau;
}

// The strain limit (anisompm.py:622-636): F <- F V diag(clamp(S)/S) V^T from
// the eigenvectors V of F^T F and the singular values S, then S <- clamp(S).
template <typename T> __device__ __forceinline__ void fclamp_apply(const StepArgs<T>& A, const M3<T>& V, M3<T>& F, T S[3]) {
    T sc[3], r[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        sc[i] = minp(maxp(S[i], A.fc_lo), A.fc_hi);
        r[i] = sc[i] / maxp(S[i], T(1e-12));
    }
    M3<T> VD;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) VD.m[i][j] = V.m[i][j] * r[j];
    F = mul(F, mul_bt(VD, V));
#pragma unroll
    for (int i = 0; i < 3; ++i) S[i] = sc[i];
}

// The strain-limited F of a particle as the constitutive update forms it.
template <typename T> __device__ __forceinline__ M3<T> fclamped(const StepArgs<T>& A, const M3<T>& F) {
    if (!A.use_fclamp) return F;
    M3<T> Cm = mul_at(F, F), V;
    T lc[3], S[3];
    sym_eig3<T, true>(Cm, lc, V);
#pragma unroll
    for (int i = 0; i < 3; ++i) S[i] = sqrt(maxp(lc[i], T(1e-12)));
    M3<T> Fc = F;
    fclamp_apply(A, V, Fc, S);
    return Fc;
}

// Stress, damage and degradation of one particle (anisompm.py:614-706).
// Returns the degraded Kirchhoff stress; F is replaced by its clamped value
// when f_clamp is on; d becomes the post-update damage.  DRIVE: stop after
// the drive (phi, phi_shear and the D_hist history max are written; Dh gets
// the new D_hist) for the implicit damage solve, which updates d itself.
template <typename T, bool DAMAGE, bool EXACT_SIGMA, bool DRIVE = false>
__device__ __forceinline__ M3<T> constitutive(const StepArgs<T>& A, long long p, const ConstIn<T>& q, M3<T>& F, T& d, T& J,
                                              bool dmg_on, T* Dh_out = nullptr) {
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
    M3<T> R = mul_bt(Wh, V);
    if (A.use_fclamp) fclamp_apply(A, V, F, S);
    J = maxp(S[0] * S[1] * S[2], T(1e-6));
    const T mu = q.mu, lam = q.lam;
    M3<T> FmR;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) FmR.m[i][j] = (T(2) * mu) * (F.m[i][j] - R.m[i][j]);
    M3<T> tdev = mul_bt(FmR, F);
    const T tvol = lam * (J - T(1)) * J;
    const bool iface = q.iface;
    T n0[3] = {T(0), T(0), T(0)};
    if (iface) {
        n0[0] = A.n_if[3 * p]; n0[1] = A.n_if[3 * p + 1]; n0[2] = A.n_if[3 * p + 2];
    }

    if (DAMAGE && !dmg_on) {
        // this scene skips damage this step: d and D_hist pass through
        A.od[p] = d;
        A.oD[p] = q.Dh;
    } else if (DAMAGE) {
        const T sc = q.sc;
        T phi, tt2 = T(0);
        if (EXACT_SIGMA) {
            // Cauchy stress, its own Jacobi, tensile part (anisompm.py:650-655)
            M3<T> sig;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) sig.m[i][j] = (tdev.m[i][j] + (i == j ? tvol : T(0))) / J;
            T ev[3];
            M3<T> Q;
            sym_eig3<T, true>(sig, ev, Q);
            M3<T> QD;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) QD.m[i][j] = Q.m[i][j] * maxp(ev[j], T(0));
            M3<T> sp = mul_bt(QD, Q);
            const M3<T>& A0 = q.A0;
            M3<T> Am = mul_bt(mul(R, A0), R);
            M3<T> M = mul(Am, sp);
            T tr = T(0);
#pragma unroll
            for (int i = 0; i < 3; ++i) tr += M.m[i][0] * M.m[0][i] + M.m[i][1] * M.m[1][i] + M.m[i][2] * M.m[2][i];
            phi = tr / (sc * sc);
            if (iface) {
                T Sn[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) Sn[i] = sp.m[i][0] * n0[0] + sp.m[i][1] * n0[1] + sp.m[i][2] * n0[2];
                T tn = Sn[0] * n0[0] + Sn[1] * n0[1] + Sn[2] * n0[2];
#pragma unroll
                for (int i = 0; i < 3; ++i) { T u = Sn[i] - tn * n0[i]; tt2 += u * u; }
            }
        } else {
            // F = R U with U = V diag(S) V^T, so tau = (R V) diag(2mu(S-1)S + tvol) (R V)^T:
            // sigma's eigenvectors are the columns of W-hat and its eigenvalues are closed form.
            T D[3];
            const T invJ = T(1) / J;
#pragma unroll
            for (int i = 0; i < 3; ++i) D[i] = maxp(((T(2) * mu) * (S[i] - T(1)) * S[i] + tvol) * invJ, T(0));
            // phi = tr(A s+ A s+) = sum_ij B_ij B_ji D_i D_j with B = V^T A0 V
            const M3<T>& A0 = q.A0;
            M3<T> B = mul_at(V, mul(A0, V));
            T tr = T(0);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) tr += B.m[i][j] * B.m[j][i] * (D[i] * D[j]);
            phi = tr / (sc * sc);
            if (iface) {
                T u[3], Sn[3];
#pragma unroll
                for (int j = 0; j < 3; ++j) u[j] = D[j] * (Wh.m[0][j] * n0[0] + Wh.m[1][j] * n0[1] + Wh.m[2][j] * n0[2]);
#pragma unroll
                for (int i = 0; i < 3; ++i) Sn[i] = Wh.m[i][0] * u[0] + Wh.m[i][1] * u[1] + Wh.m[i][2] * u[2];
                T tn = Sn[0] * n0[0] + Sn[1] * n0[1] + Sn[2] * n0[2];
#pragma unroll
                for (int i = 0; i < 3; ++i) { T w = Sn[i] - tn * n0[i]; tt2 += w * w; }
            }
        }
        if (A.ophi) {               // null: a frame's caller reads these after a later substep
            const T tt = sqrt(tt2) / sc;
            A.ophi[p] = phi;
            A.ophs[p] = tt * tt;
        }
        // history max of the drive, explicit phase-field update (anisompm.py:679-685)
        const T Dh = maxp(q.Dh, q.zeta * relu(phi - T(1)));
        A.oD[p] = Dh;
        if constexpr (DRIVE) {
            *Dh_out = Dh;
            return M3<T>{};
        }
        const T l0 = q.l0;
        const T resist = d - (l0 * l0) * q.lap;
        const T ddot = relu((T(1) - d) * Dh - resist);
        T dn = d + (A.dt_dmg / q.eta) * ddot;
        dn = minp(maxp(dn, T(0)), T(1));
        if (q.allow) d = maxp(dn, d);
        A.od[p] = d;
    }

    return degrade<T>(q.res, F, d, J, tdev, tvol, iface, n0);
}

// Stress, damage and degradation per particle, in particle order (the
// particle arrays are read coalesced), then the particle's P2G record --
// x, the affine momentum matrix, mass * v (velocity BC applied) and mass
// (anisompm.py:712-714) -- and its post-update d, stored at the particle's
// sorted position so that P2G reads a bucket's records contiguously.
template <typename T, bool DAMAGE, bool EXACT_SIGMA>
__global__ void __launch_bounds__(WG) k_const(StepArgs<T> A) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n) return;
    // every input of the particle is loaded here, before the arithmetic
    M3<T> F = load_m3(A.F, p);
    ConstIn<T> q;
    load_const_in<T, DAMAGE>(A, p, q);
    const M3<T> Cp = load_m3(A.C, p);
    T vv[3] = {A.v[3 * p], A.v[3 * p + 1], A.v[3 * p + 2]};
    const T xp[3] = {A.x[3 * p], A.x[3 * p + 1], A.x[3 * p + 2]};
    const T mass = fld<PMPM_U_MASS>(A, A.mass, p), vol0 = fld<PMPM_U_VOL0>(A, A.vol0, p);
    const int bc = A.bc_id ? A.bc_id[p] : 0;
    const long long i = A.rec_pp ? p : (long long)A.inv[p];
    // a batched step gates damage per scene; the scene is in the cell key
    const bool dmg_on = !DAMAGE || A.scene_dmg == nullptr || A.scene_dmg[(A.key[p] >> 6) / A.NBs] != 0;
    T d = q.d, J;
    M3<T> tau = constitutive<T, DAMAGE, EXACT_SIGMA>(A, p, q, F, d, J, dmg_on);
    if (A.use_fclamp) store_m3(A.oF, p, F);
    if (A.otau) {
        store_m3(A.otau, p, tau);
        A.oJ[p] = J;
    }
    T sk = T(-1) * A.dt * vol0;
    sk = sk * T(4);
    sk = sk * A.inv_dx;
    sk = sk * A.inv_dx;
    bc_velocity(A, bc, vv);
    T* r = A.rec + (size_t)P2G_REC * i;
    r[0] = xp[0]; r[1] = xp[1]; r[2] = xp[2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) r[3 + 3 * a + b] = sk * tau.m[a][b] + mass * Cp.m[a][b];
    r[12] = mass * vv[0]; r[13] = mass * vv[1]; r[14] = mass * vv[2];
    r[15] = mass;
    A.dsrt[i] = d;
}

// Channel sets of a P2G scatter: all six arena channels, the quadratic four
// (m, mv) or the cubic two (sum w d, sum w).
constexpr int CH_ALL = 0, CH_QUAD = 1, CH_CUBIC = 2;
// The balanced pair of channel sets: (mv_x, mv_y) and (m, mv_z, sum w d,
// sum w), about equal arithmetic per particle.
constexpr int CH_XY = 3, CH_MZC = 4;

// What a channel set scatters: the mass, momentum components [mlo, mhi),
// the cubic pair.
template <int CH> struct P2GChannels {
    static constexpr bool mass = CH == CH_ALL || CH == CH_QUAD || CH == CH_MZC;
    static constexpr int mlo = CH == CH_MZC ? 2 : 0;
    static constexpr int mhi = CH == CH_CUBIC ? 0 : (CH == CH_XY ? 2 : 3);
    static constexpr bool cubic = CH == CH_ALL || CH == CH_CUBIC || CH == CH_MZC;
};

// One particle's P2G data in its cell's 4x4x4 frame: per axis and frame
// node the quadratic weight W (zero-padded to the frame), D = W (x_i - x_p)
// for that axis, and the cubic weight wc; the affine matrix, mass * v, mass
// and d.
template <typename T> struct P2GPart {
    T W[4][3], D[4][3], wc[4][3], af[3][3], mv[3], mass, d;
    int bcb[3];
};

template <typename T, int CH = CH_ALL>
__device__ __forceinline__ void p2g_load(const StepArgs<T>& A, long long i, int ng, P2GPart<T>& q, int sh[3]) {
    // sorted position i; records in particle order are read through perm
    if (A.rec_pp) i = A.perm[i];
    const T* r = A.rec + (size_t)P2G_REC * i;
    const T Xp[3] = {scaled(r[0], A.inv_dx), scaled(r[1], A.inv_dx), scaled(r[2], A.inv_dx)};
    if constexpr (CH != CH_CUBIC) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) q.af[a][b] = r[3 + 3 * a + b];
        q.mv[0] = r[12]; q.mv[1] = r[13]; q.mv[2] = r[14];
        q.mass = r[15];
    }
    if constexpr (P2GChannels<CH>::cubic) q.d = A.dsrt[i];
    int bq[3];
    T fx[3], w[3][3], w2[4][3];
    quad_weights(Xp, ng, bq, fx, w);
    cubic_weights(Xp, ng, q.bcb, q.wc, w2);
    if constexpr (CH == CH_CUBIC) return;
    // the quadratic stencil [b, b+2] (b = c-1 or c) sits in the frame at shift sh
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sh[a] = bq[a] - q.bcb[a];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int j = k - sh[a];
            q.W[k][a] = j == 0 ? w[0][a] : (j == 1 ? w[1][a] : (j == 2 ? w[2][a] : T(0)));
            q.D[k][a] = q.W[k][a] * ((T(j) - fx[a]) * A.dx);
        }
    }
}

// Zero contribution (a pair's second slot when it has no partner).
template <typename T, int CH = CH_ALL> __device__ __forceinline__ void p2g_mute(P2GPart<T>& q) {
    if constexpr (CH != CH_CUBIC) q.mass = T(0);
    if constexpr (CH != CH_QUAD) q.d = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if constexpr (CH != CH_CUBIC) { q.W[k][a] = T(0); q.D[k][a] = T(0); }
            if constexpr (CH != CH_QUAD) q.wc[k][a] = T(0);
        }
}

// Scatter NP particles that share one frame into one wavefront's arena,
// whose node (x, y, z) is entry x XS + y AZ + z (XS = AR AZ dense, or
// padded).  Node (i, j, k) of the frame sits at entry l0 + i XS + j AZ + k
// (l0 the frame's first node; the offsets are compile-time constants), or,
// with CLAMP, at lp[0][i] XS + lp[1][j] AZ + lp[2][k] for a frame clamped
// onto the arena's edge (a particle beyond the grid).  A particle adds to node
// (i, j, k) of its frame
//   m Wi Wj Wk,   Wi Wj Wk mv + A (Di Wj Wk, Wi Dj Wk, Wi Wj Dk)   (momentum)
//   d wci wcj wck, wci wcj wck                                       (cubic)
// which separates: per (i, j) the row terms P = Wi Wj mv + A0 Di Wj + A1 Wi Dj
// and Q = A2 Wi Wj, then per k momentum += P Wk + Q Dk -- fused multiply-adds
// straight into the node's arena values.  Lanes own distinct cells, so for a
// fixed node offset the 64 lanes touch 64 distinct entries and no atomics
// are needed.  Different nodes of different lanes can be one entry, so the
// accesses must reach the LDS in program order, which executes one
// wavefront's accesses in order: a compiler barrier follows every write.
// Software pipeline: the next node's entry is read right after this node's
// write (so after it, in LDS order), and the next node's contributions are
// computed while that read is in flight; a scheduling barrier keeps that
// arithmetic behind the read.
template <typename T, int NP, int AZ = AR, int CH = CH_ALL, bool CLAMP = false, int XS = AR * AZ>
__device__ __forceinline__ void p2g_scatter(T (*sq)[4], T (*sc)[2], int l0, const int (*lp)[4],
                                            const P2GPart<T> (&q)[NP]) {
    using K = P2GChannels<CH>;
    constexpr bool QD = K::mass || K::mhi > K::mlo, CU = K::cubic;
    auto at = [&](int ox, int oy, int oz) {
        if constexpr (CLAMP) return lp[0][ox] * XS + lp[1][oy] * AZ + lp[2][oz];
        else return l0 + ox * XS + oy * AZ + oz;
    };
    // accumulators: c0 mass, c1..c3 momentum, f0 f1 cubic; the arena entry
    // holds the set's channels (CH_XY: (mv_x, mv_y) in sc; CH_MZC: (m,
    // mv_z, sum w d, sum w) in sq)
    T c0, c1, c2, c3, f0, f1;
    auto load = [&](int l) {
        if constexpr (CH == CH_XY) { c1 = sc[l][0]; c2 = sc[l][1]; }
        else if constexpr (CH == CH_MZC) { c0 = sq[l][0]; c3 = sq[l][1]; f0 = sq[l][2]; f1 = sq[l][3]; }
        else {
            if constexpr (QD) { c0 = sq[l][0]; c1 = sq[l][1]; c2 = sq[l][2]; c3 = sq[l][3]; }
            if constexpr (CU) { f0 = sc[l][0]; f1 = sc[l][1]; }
        }
    };
    auto store = [&](int l) {
        if constexpr (CH == CH_XY) { sc[l][0] = c1; sc[l][1] = c2; }
        else if constexpr (CH == CH_MZC) { sq[l][0] = c0; sq[l][1] = c3; sq[l][2] = f0; sq[l][3] = f1; }
        else {
            if constexpr (QD) { sq[l][0] = c0; sq[l][1] = c1; sq[l][2] = c2; sq[l][3] = c3; }
            if constexpr (CU) { sc[l][0] = f0; sc[l][1] = f1; }
        }
    };
    load(at(0, 0, 0));
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ox = 0; ox < 4; ++ox) {
#pragma unroll
        for (int oy = 0; oy < 4; ++oy) {
            T P[NP][3], Q[NP][3], M[NP], CW[NP], CD[NP];
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                if constexpr (QD) {
                    const T WW = q[j].W[ox][0] * q[j].W[oy][1];
                    const T DW = q[j].D[ox][0] * q[j].W[oy][1];
                    const T WD = q[j].W[ox][0] * q[j].D[oy][1];
#pragma unroll
                    for (int c = K::mlo; c < K::mhi; ++c) {
                        P[j][c] = q[j].af[c][1] * WD + (q[j].af[c][0] * DW + WW * q[j].mv[c]);
                        Q[j][c] = q[j].af[c][2] * WW;
                    }
                    if constexpr (K::mass) M[j] = WW * q[j].mass;
                }
                if constexpr (CU) {
                    CW[j] = q[j].wc[ox][0] * q[j].wc[oy][1];
                    CD[j] = CW[j] * q[j].d;
                }
            }
#pragma unroll
            for (int oz = 0; oz < 4; ++oz) {
#pragma unroll
                for (int j = 0; j < NP; ++j) {
                    if constexpr (QD) {
                        const T Wk = q[j].W[oz][2], Dk = q[j].D[oz][2];
                        if constexpr (K::mass) c0 = M[j] * Wk + c0;
                        if constexpr (K::mlo <= 0 && K::mhi > 0) c1 = P[j][0] * Wk + (Q[j][0] * Dk + c1);
                        if constexpr (K::mlo <= 1 && K::mhi > 1) c2 = P[j][1] * Wk + (Q[j][1] * Dk + c2);
                        if constexpr (K::mlo <= 2 && K::mhi > 2) c3 = P[j][2] * Wk + (Q[j][2] * Dk + c3);
                    }
                    if constexpr (CU) {
                        const T Ck = q[j].wc[oz][2];
                        f0 = CD[j] * Ck + f0;
                        f1 = CW[j] * Ck + f1;
                    }
                }
                store(at(ox, oy, oz));
                __asm__ __volatile__("" ::: "memory");
                const int n = (ox * 4 + oy) * 4 + oz + 1;
                if (n < 64) load(at(n >> 4, (n >> 2) & 3, n & 3));
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
}

// A particle's frame in the arena: its cubic base (cell - 1 per axis)
// relative to the arena origin org sits at 0..3 (0..AZ-4 along z) for a
// particle inside the grid, so the frame's 4 nodes per axis follow from its
// first node (frame_origin; ax offsets x in stacked arenas).  A particle
// beyond the grid is binned to the edge cell while its base lies further
// out: its frame is clamped onto the arena's edge (frame_clamped true, node
// positions per axis from frame_nodes), and the grid pass folds the edge
// positions onto the edge nodes.
template <int AZ>
__device__ __forceinline__ bool frame_clamped(const int bcb[3], const int org[3]) {
    const int x = bcb[0] - org[0], y = bcb[1] - org[1], z = bcb[2] - org[2];
    return x < 0 || x > AR - 4 || y < 0 || y > AR - 4 || z < 0 || z > AZ - 4;
}

template <int AZ, int XS = AR * AZ>
__device__ __forceinline__ int frame_origin(const int bcb[3], const int org[3], int ax) {
    const int x = clampi(bcb[0] - org[0], 0, AR - 4) + ax, y = clampi(bcb[1] - org[1], 0, AR - 4),
              z = clampi(bcb[2] - org[2], 0, AZ - 4);
    return x * XS + y * AZ + z;
}

template <int AZ>
__device__ __forceinline__ void frame_nodes(const int bcb[3], const int org[3], int ax, int lp[3][4]) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 4; ++k) lp[a][k] = clampi(bcb[a] + k - org[a], 0, (a == 2 ? AZ : AR) - 1) + (a == 0 ? ax : 0);
}

// One pass of a lane over its cell: the sorted particle record i (and i + 1
// when NP is 2 and left > 1, left being the cell's records from i on)
// scattered into the arenas sq, sc (the channels of CH), whose node 0 is
// grid node org; ax offsets the arena position along x (arenas stacked
// along x in one array).
template <typename T, int NP, int AZ = AR, int CH = CH_ALL, int XS = AR * AZ>
__device__ __forceinline__ void p2g_pass(const StepArgs<T>& A, int i, int left, const int org[3], T (*sq)[4], T (*sc)[2],
                                         int ax = 0) {
    const int ng = A.ng;
    P2GPart<T> q[NP];
    int sh[3];
#pragma unroll
    for (int j = 0; j < NP; ++j) p2g_load<T, CH>(A, i + (j < left ? j : 0), ng, q[j], sh);
    // particles of one cell share the frame unless they lie beyond the grid
    // (cells clamped together); a particle beyond the grid (its frame
    // clamped) goes alone after the others, with its frame's node positions
    // clamped one by one
    const bool clamped = frame_clamped<AZ>(q[0].bcb, org);
    const int l0 = frame_origin<AZ, XS>(q[0].bcb, org, ax);
    bool alone[NP];
    alone[0] = clamped;
    if (clamped) p2g_mute<T, CH>(q[0]);
#pragma unroll
    for (int j = 1; j < NP; ++j) {
        const bool has = j < left;
        const bool same = has && q[j].bcb[0] == q[0].bcb[0] && q[j].bcb[1] == q[0].bcb[1] && q[j].bcb[2] == q[0].bcb[2];
        alone[j] = has && (!same || clamped);
        if (!same || clamped) p2g_mute<T, CH>(q[j]);
    }
    p2g_scatter<T, NP, AZ, CH, false, XS>(sq, sc, l0, nullptr, q);
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        if (!alone[j]) continue;
        P2GPart<T> q1[1];
        p2g_load<T, CH>(A, i + j, ng, q1[0], sh);
        int lp1[3][4];
        frame_nodes<AZ>(q1[0].bcb, org, ax, lp1);
        p2g_scatter<T, 1, AZ, CH, true, XS>(sq, sc, 0, lp1, q1);
    }
}

// One wavefront's share of a bucket's P2G: lane l owns cell l of bucket bk
// (its particles contiguous in the sorted records) and scatters the cell's
// particles first, first + step, ... (NP per arena update) into the
// wavefront's arenas sq, sc, whose node 0 is grid node org.  Waves with on =
// false only take part in the round count.
template <typename T, int NP>
__device__ __forceinline__ void p2g_cells(const StepArgs<T>& A, int bk, const int org[3], int first, int step, bool on,
                                          T (*sq)[4], T (*sc)[2]) {
    const int lane = threadIdx.x & 63;
    const long long cell = ((long long)bk << 6) + lane;
    const int c0 = A.start[cell], cnt = A.start[cell + 1] - c0;
    int rounds = cnt;
    for (int o = 32; o > 0; o >>= 1) rounds = max(rounds, __shfl_xor(rounds, o));
    for (int r = first; r < rounds; r += step) {
        if (!on || r >= cnt) continue;
        p2g_pass<T, NP>(A, c0 + r, cnt - r, org, sq, sc);
    }
}

// P2G (quadratic momentum, cubic damage) per bucket, from the records k_const
// wrote.  A work-group of P2G_WAVES (2 or 4) wavefronts takes one bucket; in every
// wavefront lane l owns cell l of the bucket and takes NP of the cell's
// particles per pass (wavefront w: particles NP w .. NP w + NP-1, then
// NP (w + P2G_WAVES), ...), summed per node in registers before the arena
// update.  Each wavefront scatters into its own LDS arena; the arenas are
// summed at the end and stored to the bucket's slot with plain coalesced
// stores.
constexpr int P2G_WG = 256;
template <typename T, int NP, int P2G_WAVES>
__global__ void __launch_bounds__(64 * P2G_WAVES) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1)))
k_p2g(StepArgs<T> A) {
    constexpr int NT = 64 * P2G_WAVES;
    // per wave: node-major (m, mvx, mvy, mvz) and (sum w d, sum w), so a node's
    // update is one 16-byte and one 8-byte LDS access (fp32)
    __shared__ __align__(16) T s_q[P2G_WAVES][AR3][4];
    __shared__ __align__(16) T s_c[P2G_WAVES][AR3][2];
    const int nact = A.misc[0];
    const int nb = A.nb;
    const int wv = threadIdx.x >> 6;
    for (BucketIter it(nact); it.i < it.end; it.i += it.step) {
        const int bk = A.alist[it.i];
        const int loc = (int)(bk % A.NBs);
        const int bx = loc / (nb * nb), by = (loc / nb) % nb, bz = loc % nb;
        const int org[3] = {BLK * bx - 1, BLK * by - 1, BLK * bz - 1};
        for (int l = threadIdx.x; l < P2G_WAVES * 4 * AR3; l += NT) (&s_q[0][0][0])[l] = T(0);
        for (int l = threadIdx.x; l < P2G_WAVES * 2 * AR3; l += NT) (&s_c[0][0][0])[l] = T(0);
        __syncthreads();
        p2g_cells<T, NP>(A, bk, org, NP * wv, NP * P2G_WAVES, wv < P2G_WAVES, s_q[wv < P2G_WAVES ? wv : 0],
                         s_c[wv < P2G_WAVES ? wv : 0]);
        __syncthreads();
        // sum the wavefront arenas in a fixed order and store this bucket's
        // slot with plain coalesced stores
        // slot layout: node-major [343][6] (m, mvx, mvy, mvz, sum w d, sum w)
        T* dst = A.arena + (size_t)it.i * (6 * AR3);
        for (int l = threadIdx.x; l < 6 * AR3; l += NT) {
            const int node = l / 6, c = l - node * 6;
            T v = c < 4 ? s_q[0][node][c] : s_c[0][node][c - 4];
#pragma unroll
            for (int w = 1; w < P2G_WAVES; ++w) v += c < 4 ? s_q[w][node][c] : s_c[w][node][c - 4];
            dst[l] = v;
        }
        __syncthreads();
    }
}

// P2G with the larger cell block: 4x4x8 cells (the two buckets adjacent
// along z) per wavefront, into one 7x7x11-node block arena stored at the
// block's dense index (scene, bx, by, bz / 2) -- the grid pass then folds
// 8-node-long block arenas along z instead of 4-node ones.  The block's 128
// (bucket, cell) items go to the 64 lanes packed by load: ranked by pair
// passes (ties by item index, so the order is fixed) and dealt in snake
// order (lane l takes ranks l and 127 - l), each lane running its items'
// passes back to back.  A thin or torn body leaves buckets part-empty and
// their cells unevenly filled, which one bucket per wavefront pays as
// max-cell-count passes with most lanes idle.  An item belongs to one lane,
// so the lanes of an instruction hold distinct cells of the block and, at a
// common frame offset, touch distinct nodes, as in k_p2g.  The work-group of
// a block is the one that meets its even bucket in the active list (or its
// odd bucket when the even one is empty); blocks are fixed by bucket
// position within a scene, so a scene's sums run in the same order batched
// or alone.
template <typename T, int NP>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1)))
k_p2g_cb8(StepArgs<T> A) {
    constexpr int NI = 128;
    __shared__ __align__(16) T s_q[AR3B][4];
    __shared__ __align__(16) T s_c[AR3B][2];
    // the item ranking runs before the arena is cleared, in the arena's memory
    unsigned* s_key = reinterpret_cast<unsigned*>(&s_q[0][0]);
    int* s_item = reinterpret_cast<int*>(s_key + NI);
    int* s_c0 = s_item + NI;
    int* s_n = s_c0 + NI;
    const int nact = A.misc[0];
    const int nb = A.nb, nbz = (nb + 1) >> 1;
    const int lane = threadIdx.x;
    for (BucketIter it(nact); it.i < it.end; it.i += it.step) {
        const int bk = A.alist[it.i];
        const int scene = (int)(bk / A.NBs), loc = (int)(bk - (long long)scene * A.NBs);
        const int bx = loc / (nb * nb), by = (loc / nb) % nb, bz = loc % nb;
        if ((bz & 1) && A.slotof[bk - 1] >= 0) continue;     // the even bucket's work-group has it
        const int bz0 = bz & ~1;
        const long long b0 = (long long)bk - bz + bz0;         // the block's even bucket
        // items j * 64 + cell: key = pair passes, then the reverse index
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            int c0 = 0, n = 0;
            if (bz0 + j < nb && A.slotof[b0 + j] >= 0) {
                const long long cell = ((b0 + j) << 6) + lane;
                c0 = A.start[cell];
                n = A.start[cell + 1] - c0;
            }
            const unsigned pp = (unsigned)((n + NP - 1) / NP);
            const int idx = j * 64 + lane;
            s_key[idx] = ((pp < 0xFFFFu ? pp : 0xFFFFu) << 16) | (unsigned)(0xFFFF - idx);
            s_c0[idx] = c0;
            s_n[idx] = n;
        }
        __syncthreads();
        {
            // rank = number of larger keys (all keys differ)
            const unsigned k0 = s_key[lane], k1 = s_key[64 + lane];
            int r0 = 0, r1 = 0;
            for (int m = 0; m < NI; m += 4) {
                const uint4 o = *reinterpret_cast<const uint4*>(&s_key[m]);
                r0 += (o.x > k0) + (o.y > k0) + (o.z > k0) + (o.w > k0);
                r1 += (o.x > k1) + (o.y > k1) + (o.z > k1) + (o.w > k1);
            }
            s_item[r0] = lane;
            s_item[r1] = 64 + lane;
        }
        __syncthreads();
        // this lane's two items (snake order: rank l, then 127 - l; the
        // empty items rank last, so an empty first item means no second)
        const int ia = s_item[lane], ib = s_item[127 - lane];
        int ca = s_c0[ia], na = s_n[ia];
        int cb = s_c0[ib], nbi = s_n[ib];
        const int load = (na + NP - 1) / NP + (nbi + NP - 1) / NP;
        int rounds = load;
        for (int o = 32; o > 0; o >>= 1) rounds = max(rounds, __shfl_xor(rounds, o));
        __syncthreads();                    // items read: the arena may be cleared
        for (int l = lane; l < 4 * AR3B; l += 64) (&s_q[0][0])[l] = T(0);
        for (int l = lane; l < 2 * AR3B; l += 64) (&s_c[0][0])[l] = T(0);
        __syncthreads();
        const int org[3] = {BLK * bx - 1, BLK * by - 1, BLK * bz0 - 1};
        int r = 0;
        for (int t = 0; t < rounds; ++t) {
            if (r < na) {
                p2g_pass<T, NP, ARZ8>(A, ca + r, na - r, org, s_q, s_c);
                r += NP;
                if (r >= na) {
                    ca = cb;
                    na = nbi;
                    nbi = 0;
                    r = 0;
                }
            }
        }
        __syncthreads();
        // the block arena, node-major [539][6] at node (lx * 7 + ly) * 11 + lz
        const long long blk = (long long)scene * A.NBs8 + (bx * nb + by) * nbz + (bz0 >> 1);
        T* dst = A.arena + blk * (6 * AR3B);
        for (int l = lane; l < 6 * AR3B; l += 64) {
            const int node = l / 6, c = l - node * 6;
            dst[l] = c < 4 ? s_q[node][c] : s_c[node][c - 4];
        }
        __syncthreads();
    }
}

// P2G split by channel, with cells split over two arenas: one work-group
// of two wavefronts per 4x4x4-cell bucket, wavefront 0 scattering the
// quadratic channels (m, mv) and wavefront 1 the cubic ones (sum w d,
// sum w) of all the bucket's particles, each into its own pair of arenas.
// With one arena per wavefront a bucket costs its fullest cell's passes
// however empty the other lanes are, and a torn or compressed body leaves
// cells unevenly filled (1 to 16 particles).  Here a cell of p >= 2 passes
// is split: a lane runs the first a = ceil(p / 2) passes of its own cell
// into arena 0 (one lane per cell, as in k_p2g),
// and the other p - a passes go, into arena 1, to a lane whose own load is
// small: the lanes ranked by a ascending take the cells ranked by p - a
// descending, rank for rank (ties by index, so the order is fixed).  A lane
// runs its two items back to back.  Every item belongs to one lane, so at a
// common frame offset the lanes of an instruction touch distinct entries:
// distinct cells of one arena hit distinct nodes, and a cell's two items lie
// in different arenas.  Splitting the channels between the wavefronts keeps
// the work-group's LDS at one six-channel arena pair (two arenas per
// wavefront at the size of one), and each wavefront's pass does its
// channels' share of the arithmetic, and the smaller per-wavefront state
// lets four wavefronts share a SIMD (128 VGPRs; k_p2g holds 168, three).
// The arenas are summed in a fixed order and stored to the bucket's slot as
// k_p2g stores it.  BAL splits the channels as (mv_x, mv_y) and (m, mv_z,
// sum w d, sum w) instead, which evens out the two wavefronts' arithmetic
// (14 FMAs per frame node and particle pair against 4 in the quadratic /
// cubic split; 8 against 10 balanced).  The arenas are padded (PAD_XS) and
// lane l owns cell pad_cell(l), free of bank conflicts among own cells.
constexpr int CHAN_E = 2 * AR * PAD_XS;   // entries of an arena pair

template <typename T, int NP, bool BAL>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1)))
k_p2g_chan(StepArgs<T> A) {
    // padded arena pairs stacked along x: arena w's node (lx, ly, lz) is
    // entry (w * 7 + lx) * PAD_XS + ly * 7 + lz
    __shared__ __align__(16) T s_q[CHAN_E][4];
    __shared__ __align__(16) T s_c[CHAN_E][2];
    __shared__ int s_item[2][128];          // per wavefront: second items by rank (first record, records)
    const int nact = A.misc[0];
    const int nb = A.nb;
    for (BucketIter it(nact); it.i < it.end; it.i += it.step) {
        // the lane's constants are recomputed per bucket (an opaque copy of
        // the thread index), so the compiler keeps them out of the
        // registers the passes need instead of spilling them
        int tid = (int)threadIdx.x;
        __asm__ __volatile__("" : "+v"(tid));
        const int lane = tid & 63, wv = tid >> 6;
        const unsigned long long below = (1ull << lane) - 1ull;
        const int bk = A.alist[it.i];
        const int loc = (int)(bk % A.NBs);
        const int bx = loc / (nb * nb), by = (loc / nb) % nb, bz = loc % nb;
        const long long cell = ((long long)bk << 6) + pad_cell(lane);
        const int c0 = A.start[cell], n = A.start[cell + 1] - c0;
        const int pp = (n + NP - 1) / NP;
        const int pa = pp >= 2 ? (pp + 1) >> 1 : pp, pb = pp - pa;
        const int na = min(n, pa * NP);
        int top = pa;
        for (int o = 32; o > 0; o >>= 1) top = max(top, __shfl_xor(top, o));
        // ranks: of pa ascending and of pb descending, ties by lane
        int ra = 0, rb = 0, acc_a = 0, acc_b = 0;
        for (int v = 0; v <= top; ++v) {
            const unsigned long long ma = __ballot(pa == v), mb = __ballot(pb == top - v);
            if (pa == v) ra = acc_a + __popcll(ma & below);
            if (pb == top - v) rb = acc_b + __popcll(mb & below);
            acc_a += __popcll(ma);
            acc_b += __popcll(mb);
        }
        s_item[wv][rb] = c0 + na;
        s_item[wv][64 + rb] = n - na;
        // this wavefront's arena pair cleared
        if (wv == 0) {
            for (int l = lane; l < 4 * CHAN_E; l += 64) (&s_q[0][0])[l] = T(0);
        } else {
            for (int l = lane; l < 2 * CHAN_E; l += 64) (&s_c[0][0])[l] = T(0);
        }
        __syncthreads();
        int ca = c0, nx = na, wa = 0;
        const int cb = s_item[wv][ra], nbi = s_item[wv][64 + ra];
        int rounds = pa + (nbi + NP - 1) / NP;
        for (int o = 32; o > 0; o >>= 1) rounds = max(rounds, __shfl_xor(rounds, o));
        const int org[3] = {BLK * bx - 1, BLK * by - 1, BLK * bz - 1};
        int r = 0, left = nbi;
        for (int t = 0; t < rounds; ++t) {
            if (r >= nx && left > 0) {
                ca = cb;
                nx = left;
                wa = AR;
                left = 0;
                r = 0;
            }
            if (r < nx) {
                if constexpr (BAL) {
                    if (wv == 0) p2g_pass<T, NP, AR, CH_XY, PAD_XS>(A, ca + r, nx - r, org, nullptr, s_c, wa);
                    else p2g_pass<T, NP, AR, CH_MZC, PAD_XS>(A, ca + r, nx - r, org, s_q, nullptr, wa);
                } else {
                    if (wv == 0) p2g_pass<T, NP, AR, CH_QUAD, PAD_XS>(A, ca + r, nx - r, org, s_q, nullptr, wa);
                    else p2g_pass<T, NP, AR, CH_CUBIC, PAD_XS>(A, ca + r, nx - r, org, nullptr, s_c, wa);
                }
                r += NP;
            }
        }
        __syncthreads();
        // arena 0 + arena 1, node-major [343][6] as k_p2g stores it
        T* dst = A.arena + (size_t)it.i * (6 * AR3);
        for (int l = threadIdx.x; l < 6 * AR3; l += 128) {
            const int node = l / 6, c = l - node * 6;
            const int e0 = pad_entry(node), e1 = e0 + PAD3;
            if constexpr (BAL) {
                // (m, mv_z, sum w d, sum w) in s_q, (mv_x, mv_y) in s_c
                const int e = c == 0 ? 0 : (c >= 3 ? c - 2 : c - 1);
                dst[l] = c == 1 || c == 2 ? s_c[e0][e] + s_c[e1][e] : s_q[e0][e] + s_q[e1][e];
            } else {
                dst[l] = c < 4 ? s_q[e0][c] + s_q[e1][c] : s_c[e0][c - 4] + s_c[e1][c - 4];
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ arena fold
// Arena slot of the arena unit k = lane of the 27 around node block
// (bx, by, bz), k = 9 dx + 3 dy + dz for unit offsets dx, dy, dz + 1 in
// {0,1,2}, on lanes 0..26 (-1: outside the grid, empty, or another lane).
// CB = 1: the units are the 4x4x4-cell buckets, the slot their compact arena
// slot; CB = 2: along z the units are 4x4x8-cell blocks (uz = bz / 2), the
// slot their dense block index, present when either bucket holds particles.
template <typename T, int CB>
__device__ __forceinline__ int unit_slot_lane(const StepArgs<T>& A, int scene, int bx, int by, int uz, int lane) {
    const int nb = A.nb;
    const int u = bx - 1 + lane / 9, v = by - 1 + (lane / 3) % 3, w = uz - 1 + lane % 3;
    if constexpr (CB == 1) {
        const bool in = lane < 27 && u >= 0 && u < nb && v >= 0 && v < nb && w >= 0 && w < nb;
        return in ? A.slotof[scene * A.NBs + (u * nb + v) * nb + w] : -1;
    } else {
        const int nbz = (nb + 1) >> 1;
        const bool in = lane < 27 && u >= 0 && u < nb && v >= 0 && v < nb && w >= 0 && w < nbz;
        if (!in) return -1;
        const long long b0 = scene * A.NBs + (u * nb + v) * nb + 2 * w;
        const bool act = A.slotof[b0] >= 0 || (2 * w + 1 < nb && A.slotof[b0 + 1] >= 0);
        return act ? (int)(scene * A.NBs8 + (u * nb + v) * nbz + w) : -1;
    }
}

// The first NCH channels of a node's 6 node-major values (8-byte pairs in
// fp32: a node's values start 24-byte aligned).
template <typename T, int NCH> __device__ __forceinline__ void load_node(const T* __restrict__ src, T out[NCH]) {
    if constexpr (sizeof(T) == 4 && NCH % 2 == 0) {
        const float2* s2 = reinterpret_cast<const float2*>(src);
#pragma unroll
        for (int c = 0; c < NCH / 2; ++c) {
            const float2 v = s2[c];
            out[2 * c] = v.x;
            out[2 * c + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int c = 0; c < NCH; ++c) out[c] = src[c];
    }
}

// Sum of the first NCH channels of node g over the arenas that hold it, in
// a fixed unit order (ascending unit offsets).  Per axis an arena unit of U
// nodes (4, or 8 along z for CB = 2) starting at node U u covers nodes
// [U u - 1, U u + U + 1]; node U u + i of unit u lies in unit u - 1 too for
// i = 0, 1 and in unit u + 1 too for i = U - 1: one or two units per axis,
// at most 8 in all.  The 8 candidates are loaded unconditionally -- a
// missing one reads slot 0 and is masked to zero -- so all loads are in
// flight together.  lslot: this wavefront's unit_slot_lane (every lane of
// the wavefront must be here).  Edge nodes 0 and ng-1 also collect the arena
// positions beyond the grid (the frames keep those positions unclamped; the
// reference clamps such nodes onto the edge node); off_grid: some
// particle's stencil reached past the grid this step -- otherwise those
// positions hold zeros.
template <typename T, int NCH, int CB>
__device__ __forceinline__ void fold_node(const StepArgs<T>& A, int scene, int lslot, const int g[3], int bx, int by, int bz,
                                          bool off_grid, T acc[NCH]) {
    constexpr int UZ = BLK * CB, ARZ = UZ + 3, SLOT = 6 * AR * AR * ARZ;
    const T* __restrict__ arena = A.arena;
    const int ng = A.ng;
    const int u[3] = {bx, by, CB == 1 ? bz : bz >> 1};
    const int U[3] = {BLK, BLK, UZ};
    int lo[3];
    bool two[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int i = g[a] - U[a] * u[a];
        lo[a] = i < 2 ? -1 : 0;
        two[a] = i < 2 || i == U[a] - 1;
    }
    T v[8][NCH];
    int cs[8];       // slot of candidate unit c, -1 when it does not hold the node
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int cx = c >> 2, cy = (c >> 1) & 1, cz = c & 1;
        const int ux = lo[0] + cx, uy = lo[1] + cy, uz = lo[2] + cz;
        const int s = __shfl(lslot, (ux + 1) * 9 + (uy + 1) * 3 + (uz + 1));
        const bool ok = (cx == 0 || two[0]) && (cy == 0 || two[1]) && (cz == 0 || two[2]) && s >= 0;
        cs[c] = ok ? s : -1;
        const int lx = g[0] - (BLK * (bx + ux) - 1), ly = g[1] - (BLK * (by + uy) - 1), lz = g[2] - (UZ * (u[2] + uz) - 1);
        const size_t off = ok ? (size_t)s * SLOT + 6 * ((lx * AR + ly) * ARZ + lz) : 0;
        load_node<T, NCH>(arena + off, v[c]);
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) v[c][ch] = ok ? v[c][ch] : T(0);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) acc[ch] += v[c][ch];
    const bool edge = g[0] == 0 || g[1] == 0 || g[2] == 0 || g[0] == ng - 1 || g[1] == ng - 1 || g[2] == ng - 1;
    if (!off_grid || !edge) return;
    // An arena position beyond the grid folds onto edge node g only from a
    // unit that holds g itself: per axis, position -1 lies in unit 0 (which
    // holds node 0), positions ng and ng+1 in the last unit (which holds
    // ng-1).  So the extra positions come from the candidates above, taken
    // in the same ascending unit order, each a box of at most 3 positions
    // per axis around the node's own one.
    const int ext[3] = {AR, AR, ARZ};
    for (int c = 0; c < 8; ++c) {
        const int sk = cs[c];
        if (sk < 0) continue;
        const int b3[3] = {u[0] + lo[0] + (c >> 2), u[1] + lo[1] + ((c >> 1) & 1), u[2] + lo[2] + (c & 1)};
        int lo3[3], hi3[3], own[3];
        bool any_extra = false;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int org = U[a] * b3[a] - 1;
            const int plo = g[a] == 0 ? org : g[a];
            const int phi = g[a] == ng - 1 ? org + ext[a] - 1 : g[a];
            lo3[a] = (plo > org ? plo : org) - org;
            hi3[a] = (phi < org + ext[a] - 1 ? phi : org + ext[a] - 1) - org;
            own[a] = g[a] - org;
            any_extra = any_extra || lo3[a] != hi3[a];
        }
        if (!any_extra) continue;
        const T* base = arena + (size_t)sk * SLOT;
        for (int lx = lo3[0]; lx <= hi3[0]; ++lx)
            for (int ly = lo3[1]; ly <= hi3[1]; ++ly)
                for (int lz = lo3[2]; lz <= hi3[2]; ++lz) {
                    // the node's own position was summed above
                    if (lx == own[0] && ly == own[1] && lz == own[2]) continue;
                    T e[NCH];
                    load_node<T, NCH>(base + 6 * ((lx * AR + ly) * ARZ + lz), e);
#pragma unroll
                    for (int ch = 0; ch < NCH; ++ch) acc[ch] += e[ch];
                }
    }
}

// One collider at a node position: its velocity (the forward difference of
// its placement, anisompm.py:837-839, :861-863), the outward normal there and
// whether the node is inside it (anisompm.py:828-867).
template <typename T>
__device__ __forceinline__ bool collider_setup(const pmpm_collider& cl, const T pos[3], T vc[3], T n[3]) {
    const T p0[3] = {T(cl.p0[0]), T(cl.p0[1]), T(cl.p0[2])};
    const T p1[3] = {T(cl.p1[0]), T(cl.p1[1]), T(cl.p1[2])};
#pragma unroll
    for (int a = 0; a < 3; ++a) vc[a] = (p1[a] - p0[a]) / T(1e-4);
    if (cl.kind == PMPM_COLLIDER_SPHERE) {
        const T r[3] = {pos[0] - p0[0], pos[1] - p0[1], pos[2] - p0[2]};
        const T dist = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const T den = dist + T(1e-9);
        n[0] = r[0] / den; n[1] = r[1] / den; n[2] = r[2] / den;
        return dist < T(cl.radius);
    }
    n[0] = T(cl.normal[0]); n[1] = T(cl.normal[1]); n[2] = T(cl.normal[2]);
    const T sd = (pos[0] - p0[0]) * n[0] + (pos[1] - p0[1]) * n[1] + (pos[2] - p0[2]) * n[2];
    return sd < T(0);
}

// Grid velocity update and colliders per marked 4x4x4 node block
// (anisompm.py:726-746), cubic normalisation (:777-779).  The nodes of block
// (bx,by,bz) lie in the arenas [4b-1, 4b+5] of buckets b = bx-1..bx+1 per
// axis; a node sums the arenas that cover it, in a fixed bucket order, so
// the loop over the 27 candidate buckets is uniform across the wavefront.
template <typename T, int CB>
__global__ void __launch_bounds__(WG) k_grid(StepArgs<T> A) {
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
    const bool off_grid = A.misc[3] != 0;
    const int nb = A.nb, ng = A.ng;
    unsigned long long nq = 0, nc = 0;
    for (int i = wave; i < nblocks; i += nwaves) {
        const int blk = A.blist[i];
        const int scene = (int)(blk / A.NBs), loc = (int)(blk - scene * A.NBs);
        const int bx = loc / (nb * nb), by = (loc / nb) % nb, bz = loc % nb;
        const int g[3] = {BLK * bx + (lane >> 4), BLK * by + ((lane >> 2) & 3), BLK * bz + (lane & 3)};
        T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        fold_node<T, 6, CB>(A, scene, unit_slot_lane<T, CB>(A, scene, bx, by, CB == 1 ? bz : bz >> 1, lane), g, bx, by,
                            bz, off_grid, acc);
        if (g[0] >= ng || g[1] >= ng || g[2] >= ng) continue;
        const int gx = g[0], gy = g[1], gz = g[2];
        const long long node = scene * A.Gs + ((long long)gx * ng + gy) * ng + gz;
        const T m = acc[0], mvx = acc[1], mvy = acc[2], mvz = acc[3], dsum = acc[4], wsum = acc[5];
        {
            const bool act = m > T(1e-12);
            T v[3] = {T(0), T(0), T(0)};
            if (act) {
                const T inv = T(1) / m;
                v[0] = mvx * inv + A.dtg[0];
                v[1] = mvy * inv + A.dtg[1];
                v[2] = mvz * inv + A.dtg[2];
                if (A.damp != T(1)) { v[0] = v[0] * A.damp; v[1] = v[1] * A.damp; v[2] = v[2] * A.damp; }
                const T pos[3] = {T(gx) * A.dx, T(gy) * A.dx, T(gz) * A.dx};
                for (int k = 0; k < A.n_col; ++k) {
                    T vc[3], n[3];
                    if (collider_setup<T>(s_col[k], pos, vc, n)) collider_resolve(v, vc, n, s_col[k].mode, T(s_col[k].friction));
                }
            }
            const bool cact = wsum > T(1e-12);
            T* g = A.gv + 4 * node;
            g[0] = v[0]; g[1] = v[1]; g[2] = v[2];
            g[3] = cact ? dsum / wsum : T(0);
            if (A.export_active) A.active[node] = act ? 1 : 0;
            if (A.gmw) {
                T* gw = A.gmw + GMW * node;
                gw[0] = m; gw[1] = wsum;
                gw[2] = act ? mvx * (T(1) / m) : T(0);
                gw[3] = act ? mvy * (T(1) / m) : T(0);
                gw[4] = act ? mvz * (T(1) / m) : T(0);
            }
            nq += act;
            nc += cact;
        }
    }
    // per-work-group totals of active nodes, stored (not accumulated: a
    // same-address HBM atomic from every wave serialises the whole pass)
    __shared__ unsigned long long s_n[2][WG / 64];
    for (int o = 32; o > 0; o >>= 1) {
        nq += __shfl_xor(nq, o);
        nc += __shfl_xor(nc, o);
    }
    if (lane == 0) { s_n[0][threadIdx.x >> 6] = nq; s_n[1][threadIdx.x >> 6] = nc; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long q = 0, c = 0;
        for (int w = 0; w < WG / 64; ++w) { q += s_n[0][w]; c += s_n[1][w]; }
        A.stats[2 * blockIdx.x] = q;
        A.stats[2 * blockIdx.x + 1] = c;
        // how many partials this pass wrote (the launch shape may change)
        if (blockIdx.x == 0) const_cast<int*>(A.misc)[2] = (int)gridDim.x;
    }
}

// G2P gathers of one particle from the bucket's staged nodes s_g (velocity
// in channels 0-2, normalised damage in 3), as separable sums -- over z per
// (x, y) row, then over y per x plane, then over x:
//   v = sum w g,  C[a][b] = sum w g_a dp_b  (anisompm.py:751-760, before the
//   4 inv_dx^2 factor),
//   lap = sum (w''x wy wz + wx w''y wz + wx wy w''z) gd  (:781-791, before
//   the inv_dx^2 factor).
// base: the stencil's first node per axis; org: the arena's first node.  A
// stencil node clamps to the grid as in the reference (anisompm.py:586-590).
// UNCLAMPED: no node of the stencil needs the clamp, so every node is the
// stencil base plus a compile-time offset.
template <int W, bool UNCLAMPED>
__device__ __forceinline__ void arena_pos(const int base[3], const int org[3], int ng, int& b0, int lp[3][W]) {
    b0 = ((base[0] - org[0]) * AR + (base[1] - org[1])) * AR + (base[2] - org[2]);
    if constexpr (!UNCLAMPED) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int o = 0; o < W; ++o) lp[a][o] = clampi(base[a] + o, 0, ng - 1) - org[a];
    }
}
template <typename T, bool UNCLAMPED>
__device__ __forceinline__ void g2p_gather_q(const T (*s_g)[AR3], const int bq[3], const int org[3], int ng, const T w[3][3],
                                             const T fx[3], T dx, T vn[3], M3<T>& Cn) {
    int q0, lq[3][3];
    arena_pos<3, UNCLAMPED>(bq, org, ng, q0, lq);
    T wd[3][3];                 // weight times node offset dp, per node and axis
#pragma unroll
    for (int o = 0; o < 3; ++o)
#pragma unroll
        for (int a = 0; a < 3; ++a) wd[o][a] = w[o][a] * ((T(o) - fx[a]) * dx);
#pragma unroll
    for (int ox = 0; ox < 3; ++ox) {
        T sy[3], syd[3], ty[3];
#pragma unroll
        for (int oy = 0; oy < 3; ++oy) {
            T sz[3], tz[3];
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
                const int l = UNCLAMPED ? q0 + (ox * AR + oy) * AR + oz : (lq[0][ox] * AR + lq[1][oy]) * AR + lq[2][oz];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const T g = s_g[a][l];
                    sz[a] = oz == 0 ? w[0][2] * g : sz[a] + w[oz][2] * g;
                    tz[a] = oz == 0 ? wd[0][2] * g : tz[a] + wd[oz][2] * g;
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                sy[a] = oy == 0 ? w[0][1] * sz[a] : sy[a] + w[oy][1] * sz[a];
                syd[a] = oy == 0 ? wd[0][1] * sz[a] : syd[a] + wd[oy][1] * sz[a];
                ty[a] = oy == 0 ? w[0][1] * tz[a] : ty[a] + w[oy][1] * tz[a];
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            vn[a] = ox == 0 ? w[0][0] * sy[a] : vn[a] + w[ox][0] * sy[a];
            Cn.m[a][0] = ox == 0 ? wd[0][0] * sy[a] : Cn.m[a][0] + wd[ox][0] * sy[a];
            Cn.m[a][1] = ox == 0 ? w[0][0] * syd[a] : Cn.m[a][1] + w[ox][0] * syd[a];
            Cn.m[a][2] = ox == 0 ? w[0][0] * ty[a] : Cn.m[a][2] + w[ox][0] * ty[a];
        }
    }
}
template <typename T, bool UNCLAMPED>
__device__ __forceinline__ void g2p_gather_c(const T (*s_g)[AR3], const int bc[3], const int org[3], int ng,
                                             const T wc[4][3], const T w2[4][3], T& lap) {
    int k0, lc[3][4];
    arena_pos<4, UNCLAMPED>(bc, org, ng, k0, lc);
#pragma unroll
    for (int ox = 0; ox < 4; ++ox) {
        T py, qy, ry;           // sum_y wy (sum_z wz gd), sum_y w''y (..), sum_y wy (sum_z w''z gd)
#pragma unroll
        for (int oy = 0; oy < 4; ++oy) {
            T az, bz;           // sum_z wz gd, sum_z w''z gd
#pragma unroll
            for (int oz = 0; oz < 4; ++oz) {
                const int l = UNCLAMPED ? k0 + (ox * AR + oy) * AR + oz : (lc[0][ox] * AR + lc[1][oy]) * AR + lc[2][oz];
                const T g = s_g[3][l];
                az = oz == 0 ? wc[0][2] * g : az + wc[oz][2] * g;
                bz = oz == 0 ? w2[0][2] * g : bz + w2[oz][2] * g;
            }
            py = oy == 0 ? wc[0][1] * az : py + wc[oy][1] * az;
            qy = oy == 0 ? w2[0][1] * az : qy + w2[oy][1] * az;
            ry = oy == 0 ? wc[0][1] * bz : ry + wc[oy][1] * bz;
        }
        const T px = w2[ox][0] * py + wc[ox][0] * (qy + ry);
        lap = ox == 0 ? px : lap + px;
    }
}

// G2P + cubic Laplacian + advection/F update/clamp per bucket (anisompm.py:748-803).
template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_g2p(StepArgs<T> A) {
    __shared__ T s_g[4][AR3];
    const int nact = A.misc[0];
    const int nb = A.nb, ng = A.ng;
    for (BucketIter it(nact); it.i < it.end; it.i += it.step) {
        const int bk = A.alist[it.i];
        const int scene = (int)(bk / A.NBs), loc = (int)(bk - scene * A.NBs);
        const int bx = loc / (nb * nb), by = (loc / nb) % nb, bz = loc % nb;
        const T* gvs = A.gv + 4 * scene * A.Gs;
        const int o0 = BLK * bx - 1, o1 = BLK * by - 1, o2 = BLK * bz - 1;
        const T* Fin = A.use_fclamp ? (const T*)A.oF : A.F;
        // this thread's first G2P_PREFETCH particles are fetched before the
        // node staging, so their dependent loads (perm, then x and F) overlap
        // the staging
        const int s0 = A.start[(long long)bk << 6], cnt = A.start[((long long)bk + 1) << 6] - s0;
        long long pf[G2P_PREFETCH];
        T xf[G2P_PREFETCH][3];
        M3<T> Ff[G2P_PREFETCH];
#pragma unroll
        for (int k = 0; k < G2P_PREFETCH; ++k) {
            const int i = (int)threadIdx.x + k * NT;
            pf[k] = 0;
            xf[k][0] = xf[k][1] = xf[k][2] = T(0);
            if (i < cnt) {
                pf[k] = A.perm[s0 + i];
                xf[k][0] = A.x[3 * pf[k]]; xf[k][1] = A.x[3 * pf[k] + 1]; xf[k][2] = A.x[3 * pf[k] + 2];
                Ff[k] = load_m3(Fin, pf[k]);
            }
        }
        for (int l = threadIdx.x; l < AR3; l += NT) {
            const int lx = l / (AR * AR), ly = (l / AR) % AR, lz = l % AR;
            const int gx = o0 + lx, gy = o1 + ly, gz = o2 + lz;
            if (gx < 0 || gy < 0 || gz < 0 || gx >= ng || gy >= ng || gz >= ng) {
                s_g[0][l] = T(0); s_g[1][l] = T(0); s_g[2][l] = T(0); s_g[3][l] = T(0);
            } else {
                const T* g = gvs + 4 * (((long long)gx * ng + gy) * ng + gz);
                s_g[0][l] = g[0]; s_g[1][l] = g[1]; s_g[2][l] = g[2]; s_g[3][l] = g[3];
            }
        }
        __syncthreads();
        // one particle: velocity, affine C and the damage Laplacian from the
        // staged nodes, then advection, the F update and the clamp; nk, off:
        // the cell key of the new position and its off-grid flag (bin_next)
        auto particle = [&](long long p, const T (&xp)[3], const M3<T>& F, int& nk, bool& off) {
            T Xp[3] = {scaled(xp[0], A.inv_dx), scaled(xp[1], A.inv_dx), scaled(xp[2], A.inv_dx)};
            int bq[3];
            T fx[3], w[3][3];
            quad_weights(Xp, ng, bq, fx, w);
            // a particle whose stencils need no grid clamp takes the fast
            // path (the cubic base is floor(Xp) - 1, as in cubic_weights)
            int bcb[3];
            const int org[3] = {o0, o1, o2};
            bool unclamped = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                bcb[a] = floor_idx(Xp[a], ng) - 1;
                unclamped = unclamped && bq[a] >= 0 && bq[a] + 2 < ng && bcb[a] >= 0 && bcb[a] + 3 < ng;
            }
            T vn[3], lap;
            M3<T> Cn;
            if (unclamped) g2p_gather_q<T, true>(s_g, bq, org, ng, w, fx, A.dx, vn, Cn);
            else g2p_gather_q<T, false>(s_g, bq, org, ng, w, fx, A.dx, vn, Cn);
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) Cn.m[a][b] = Cn.m[a][b] * A.c_scale;
            {
                T wc[4][3], w2[4][3];
                cubic_weights(Xp, ng, bcb, wc, w2);
                if (unclamped) g2p_gather_c<T, true>(s_g, bcb, org, ng, wc, w2, lap);
                else g2p_gather_c<T, false>(s_g, bcb, org, ng, wc, w2, lap);
            }
            lap = A.lap_scale * lap;
            M3<T> Ip;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) Ip.m[a][b] = (a == b ? T(1) : T(0)) + A.dt * Cn.m[a][b];
            store_m3(A.oF, p, mul(Ip, F));
            store_m3(A.oC, p, Cn);
            T xw[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                A.ov[3 * p + a] = vn[a];
                const T xn = xp[a] + A.dt * vn[a];
                xw[a] = minp(maxp(xn, A.pad_lo), A.pad_hi);
                A.ox[3 * p + a] = xw[a];
            }
            A.olap[p] = lap;
            if (A.bin_next) nk = cell_key(xw, A.inv_dx, ng, nb, off) + scene * (int)(A.NBs << 6);
        };
        // the next step's binning from the written positions, as k_bin_count
        // takes it from x: the lanes hold consecutive sorted particles, so
        // runs of equal keys are long (one count atomic per run)
        auto carry = [&](long long p, int nk, bool live, bool off) {
            const int lane = threadIdx.x & 63;
            if (__ballot(off) != 0 && lane == 0) A.cmisc[5] = 1;
            bool head;
            int first, len;
            wave_run(nk, live, lane, head, first, len);
            int base = 0;
            if (head) base = atomicAdd(&A.ccount[nk], len);
            base = __shfl(base, first);
            if (live) {
                A.ckey[p] = nk;
                A.cslot[p] = base + (lane - first);
            }
            const int kb = nk >> 6, kbprev = __shfl_up(kb, 1);
            if (live && (lane == 0 || kbprev != kb)) A.cbflag[kb] = 1;
        };
        // (the lanes of a wavefront run these loops together, live or not,
        // so carry's wavefront operations see every lane).  The prefetched
        // particles are all computed before their counts are taken, and
        // their count atomics are all issued before any result is used, so
        // one atomic round trip covers them.
        int nkk[G2P_PREFETCH];
        bool offk[G2P_PREFETCH], livek[G2P_PREFETCH];
#pragma unroll
        for (int k = 0; k < G2P_PREFETCH; ++k) {
            livek[k] = (int)threadIdx.x + k * NT < cnt;
            nkk[k] = -1;
            offk[k] = false;
            if (livek[k]) particle(pf[k], xf[k], Ff[k], nkk[k], offk[k]);
        }
        if (A.bin_next) {
            const int lane = threadIdx.x & 63;
            bool anyoff = false;
            int base[G2P_PREFETCH], first[G2P_PREFETCH];
#pragma unroll
            for (int k = 0; k < G2P_PREFETCH; ++k) {
                anyoff = anyoff || offk[k];
                bool head;
                int len;
                wave_run(nkk[k], livek[k], lane, head, first[k], len);
                base[k] = 0;
                if (head) base[k] = atomicAdd(&A.ccount[nkk[k]], len);
            }
            if (__ballot(anyoff) != 0 && lane == 0) A.cmisc[5] = 1;
#pragma unroll
            for (int k = 0; k < G2P_PREFETCH; ++k) {
                const int b = __shfl(base[k], first[k]);
                if (livek[k]) {
                    A.ckey[pf[k]] = nkk[k];
                    A.cslot[pf[k]] = b + (lane - first[k]);
                }
                const int kb = nkk[k] >> 6, kbprev = __shfl_up(kb, 1);
                if (livek[k] && (lane == 0 || kbprev != kb)) A.cbflag[kb] = 1;
            }
        }
        for (int i0 = (int)(threadIdx.x & ~63u) + G2P_PREFETCH * NT; i0 < cnt; i0 += NT) {
            const int i = i0 + (int)(threadIdx.x & 63);
            const bool live = i < cnt;
            long long p = 0;
            int nk = -1;
            bool off = false;
            if (live) {
                p = A.perm[s0 + i];
                const T xp[3] = {A.x[3 * p], A.x[3 * p + 1], A.x[3 * p + 2]};
                particle(p, xp, load_m3(Fin, p), nk, off);
            }
            if (A.bin_next) carry(p, nk, live, off);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ implicit damage (PCG)
// One backward-Euler step of  eta d' = (1-d) D - d + l0^2 lap d  on the cubic
// grid of the reference's Laplacian pass (anisompm.py:762-791), with the
// particles as quadrature points and the nodal mass lumped:
//   (diag(m) + K) d_n = b,  m_i = sum_p N_ip (1 + c_p (1 + D_p)),  c_p = dt_damage / eta_p,
//   K_ij = sum_p c_p l0_p^2 grad N_ip . grad N_jp,  b_i = sum_p N_ip (d_p + c_p D_p),
// over the nodes with sum_p N_ip > 1e-12.  The reference has no implicit
// damage solve; this is the build's extension.  Jacobi-preconditioned CG,
// matrix-free: K p goes through the particles (gather grad p, scatter
// kappa grad N . grad p) with the P2G arena machinery.  Vectors live per node
// in gv as (x, r, p, Ap) during the solve (gv is free until the grid pass)
// and in pdiag as (Jacobi diagonal, lumped mass).  Every kernel of an
// iteration first reads the done flag, so a converged solve costs one load
// per remaining launch and the host never waits.

// Cubic B-spline slopes dN/dX (grid units) per axis at the nodes base+0..3.
template <typename T> __device__ __forceinline__ void cubic_slopes(const T Xp[3], T ds[4][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T f = sub_exact(Xp[a], floor(Xp[a]));
        const T o = T(1) - f;
        ds[0][a] = T(-0.5) * (o * o);
        ds[1][a] = (T(1.5) * f - T(2)) * f;
        ds[2][a] = (T(2) - T(1.5) * o) * o;
        ds[3][a] = T(0.5) * (f * f);
    }
}

// Work-group sum of NCH doubles, stored as this group's partials
// (channel-major, GRID_WG_COUNT per channel), in a fixed order within the
// group; which node blocks a group holds follows the marked-block list,
// whose order the atomic appends of its compaction set.
template <int NCH> __device__ __forceinline__ void store_partials(double v[NCH], double* __restrict__ part) {
    __shared__ double s[NCH][WG / 64];
#pragma unroll
    for (int c = 0; c < NCH; ++c)
        for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < NCH; ++c) s[c][threadIdx.x >> 6] = v[c];
    __syncthreads();
    if (threadIdx.x < NCH) {
        double t = 0;
        for (int w = 0; w < WG / 64; ++w) t += s[threadIdx.x][w];
        part[threadIdx.x * GRID_WG_COUNT + blockIdx.x] = t;
    }
}

// Drive per particle, in particle order (the particle arrays are read
// coalesced): the constitutive update up to the D_hist history max (phi,
// phi_shear, D_hist written), then the particle's solve record at its sorted
// position -- x, kappa = c l0^2 inv_dx^2, the mass coefficient
// 1 + c (1 + D_hist) and the right-hand side d + c D_hist (c = dt_dmg / eta)
// -- and, for the interpolation after the solve, the old d and the allow
// flag, in the P2G record buffer, which the step fills only after the solve.
constexpr int PCG_REC = 8;              // values per solve record: x(3), kappa, mcoef, bcoef, d, allow
template <typename T, bool EXACT_SIGMA>
__global__ void __launch_bounds__(WG) k_pcg_const(StepArgs<T> A) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n) return;
    M3<T> F = load_m3(A.F, p);
    ConstIn<T> q;
    load_const_in<T, true>(A, p, q);
    const T xp[3] = {A.x[3 * p], A.x[3 * p + 1], A.x[3 * p + 2]};
    const T l0 = q.l0;
    const long long i = A.inv[p];
    T d = q.d, J, Dh;
    constitutive<T, true, EXACT_SIGMA, true>(A, p, q, F, d, J, true, &Dh);
    const T c = A.dt_dmg / q.eta;
    T* r = A.rec + (size_t)PCG_REC * i;
    r[0] = xp[0]; r[1] = xp[1]; r[2] = xp[2];
    r[3] = (c * (l0 * l0)) * (A.inv_dx * A.inv_dx);
    r[4] = T(1) + c * (T(1) + Dh);
    r[5] = d + c * Dh;
    r[6] = q.d;
    r[7] = q.allow ? T(1) : T(0);
}

// One particle's solve record unpacked for the cubic stencil: weights and
// slopes per axis, kappa, the two coefficients and the frame's base node.
template <typename T> struct PcgPart {
    T wc[4][3], ds[4][3], kg, mco, bco;
    int bcb[3];
};

template <typename T, bool COEF>
__device__ __forceinline__ void pcg_load(const StepArgs<T>& A, long long i, int ng, PcgPart<T>& q) {
    const T* rc = (const T*)__builtin_assume_aligned(A.rec + (size_t)PCG_REC * i, 4 * sizeof(T));
    const T Xp[3] = {scaled(rc[0], A.inv_dx), scaled(rc[1], A.inv_dx), scaled(rc[2], A.inv_dx)};
    q.kg = rc[3];
    if (COEF) { q.mco = rc[4]; q.bco = rc[5]; }
    T w2[4][3];
    cubic_weights(Xp, ng, q.bcb, q.wc, w2);
    cubic_slopes(Xp, q.ds);
}

// A second particle that does not share the first one's frame contributes
// nothing to a pair's pass (it is scattered alone afterwards).
template <typename T> __device__ __forceinline__ void pcg_mute(PcgPart<T>& q) {
    q.kg = T(0); q.mco = T(0); q.bco = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) { q.wc[k][a] = T(0); q.ds[k][a] = T(0); }
}

// Scatter NP particles that share one frame into one wavefront's arena of
// (N mcoef, N bcoef, kappa |grad N|^2, N) per node: the NP contributions
// are summed in registers, then one read-modify-write per node, in program
// order, software-pipelined as the P2G scatter (the next node's entry is
// read right after this node's write).
template <typename T, int NP, bool CLAMP = false, int XS = AR * AR>
__device__ __forceinline__ void pcg_drive_scatter(T (*sq)[4], int l0, const int (*lp)[4], const PcgPart<T> (&q)[NP]) {
    auto at = [&](int ox, int oy, int oz) {
        if constexpr (CLAMP) return lp[0][ox] * XS + lp[1][oy] * AR + lp[2][oz];
        else return l0 + ox * XS + oy * AR + oz;
    };
    T c0, c1, c2, c3;
    {
        const int l = at(0, 0, 0);
        c0 = sq[l][0]; c1 = sq[l][1]; c2 = sq[l][2]; c3 = sq[l][3];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ox = 0; ox < 4; ++ox)
#pragma unroll
        for (int oy = 0; oy < 4; ++oy) {
            T wxy[NP], gxy[NP], gyx[NP];
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                wxy[j] = q[j].wc[ox][0] * q[j].wc[oy][1];
                gxy[j] = q[j].ds[ox][0] * q[j].wc[oy][1];
                gyx[j] = q[j].wc[ox][0] * q[j].ds[oy][1];
            }
#pragma unroll
            for (int oz = 0; oz < 4; ++oz) {
                T a0, a1, a2, a3;
#pragma unroll
                for (int j = 0; j < NP; ++j) {
                    const T nw = wxy[j] * q[j].wc[oz][2];
                    const T gx = gxy[j] * q[j].wc[oz][2];
                    const T gy = gyx[j] * q[j].wc[oz][2];
                    const T gz = wxy[j] * q[j].ds[oz][2];
                    const T t0 = nw * q[j].mco, t1 = nw * q[j].bco, t2 = q[j].kg * (gx * gx + gy * gy + gz * gz);
                    if (j == 0) { a0 = t0; a1 = t1; a2 = t2; a3 = nw; }
                    else { a0 += t0; a1 += t1; a2 += t2; a3 += nw; }
                }
                const int l = at(ox, oy, oz);
                c0 += a0; c1 += a1; c2 += a2; c3 += a3;
                sq[l][0] = c0; sq[l][1] = c1; sq[l][2] = c2; sq[l][3] = c3;
                __asm__ __volatile__("" ::: "memory");
                const int n = (ox * 4 + oy) * 4 + oz + 1;
                if (n < 64) {
                    const int ln = at(n >> 4, (n >> 2) & 3, n & 3);
                    c0 = sq[ln][0]; c1 = sq[ln][1]; c2 = sq[ln][2]; c3 = sq[ln][3];
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
}

// K p for NP particles that share one frame: gather g_j = kappa_j sum_i
// grad N_ij p_i from the staged search direction (one LDS read per node for
// the NP particles), then scatter sum_j grad N_ij . g_j into one wavefront's
// arena, software-pipelined as above.  Both halves are separable: per
// (ox, oy) the gather keeps sum_z N_z p and sum_z N'_z p, and the scatter
// forms N_z (N'_x N_y g_x + N_x N'_y g_y) + N'_z (N_x N_y g_z).
template <typename T, int NP, bool CLAMP = false, int XS = AR * AR>
__device__ __forceinline__ void pcg_apply_scatter(const T* sp, T* sa, int l0, const int (*lp)[4], const PcgPart<T> (&q)[NP]) {
    auto at = [&](int ox, int oy, int oz) {
        if constexpr (CLAMP) return lp[0][ox] * XS + lp[1][oy] * AR + lp[2][oz];
        else return l0 + ox * XS + oy * AR + oz;
    };
    T g[NP][3];
#pragma unroll
    for (int j = 0; j < NP; ++j) g[j][0] = g[j][1] = g[j][2] = T(0);
#pragma unroll
    for (int ox = 0; ox < 4; ++ox)
#pragma unroll
        for (int oy = 0; oy < 4; ++oy) {
            T pz[4];
#pragma unroll
            for (int oz = 0; oz < 4; ++oz) pz[oz] = sp[at(ox, oy, oz)];
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                T s = q[j].wc[0][2] * pz[0], t = q[j].ds[0][2] * pz[0];
#pragma unroll
                for (int oz = 1; oz < 4; ++oz) { s += q[j].wc[oz][2] * pz[oz]; t += q[j].ds[oz][2] * pz[oz]; }
                g[j][0] += (q[j].ds[ox][0] * q[j].wc[oy][1]) * s;
                g[j][1] += (q[j].wc[ox][0] * q[j].ds[oy][1]) * s;
                g[j][2] += (q[j].wc[ox][0] * q[j].wc[oy][1]) * t;
            }
            // at most one row of reads in flight (bounds the live registers)
            if (oy == 3) __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
    for (int j = 0; j < NP; ++j) { g[j][0] *= q[j].kg; g[j][1] *= q[j].kg; g[j][2] *= q[j].kg; }
    T c = sa[at(0, 0, 0)];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ox = 0; ox < 4; ++ox)
#pragma unroll
        for (int oy = 0; oy < 4; ++oy) {
            T u[NP], v[NP];
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                // associated unlike the gather's products, which therefore
                // need not stay live across the two halves
                u[j] = q[j].ds[ox][0] * (q[j].wc[oy][1] * g[j][0]) + q[j].wc[ox][0] * (q[j].ds[oy][1] * g[j][1]);
                v[j] = q[j].wc[ox][0] * (q[j].wc[oy][1] * g[j][2]);
            }
#pragma unroll
            for (int oz = 0; oz < 4; ++oz) {
                T a = q[0].wc[oz][2] * u[0] + q[0].ds[oz][2] * v[0];
#pragma unroll
                for (int j = 1; j < NP; ++j) a += q[j].wc[oz][2] * u[j] + q[j].ds[oz][2] * v[j];
                c += a;
                sa[at(ox, oy, oz)] = c;
                __asm__ __volatile__("" ::: "memory");
                const int n = (ox * 4 + oy) * 4 + oz + 1;
                if (n < 64) c = sa[at(n >> 4, (n >> 2) & 3, n & 3)];
                __builtin_amdgcn_sched_barrier(0);
            }
        }
}

// One wavefront's pass over a cell's particles r, r + 1 (pairs that share a
// frame).  A particle beyond the grid (its frame clamped onto the arena's
// edge, which a second particle of the cell need not share) is scattered
// alone afterwards, with its frame's node positions clamped one by one.
template <typename T, bool APPLY, int XS = AR * AR>
__device__ __forceinline__ void pcg_pair(const StepArgs<T>& A, long long i0, bool two, const int org[3], const T* sp, T* arena) {
    PcgPart<T> q[2];
    pcg_load<T, !APPLY>(A, i0, A.ng, q[0]);
    pcg_load<T, !APPLY>(A, i0 + (two ? 1 : 0), A.ng, q[1]);
    const bool clamped = frame_clamped<AR>(q[0].bcb, org);
    const bool same = two && q[1].bcb[0] == q[0].bcb[0] && q[1].bcb[1] == q[0].bcb[1] && q[1].bcb[2] == q[0].bcb[2];
    const bool alone[2] = {clamped, two && (!same || clamped)};
    if (clamped) pcg_mute(q[0]);
    if (!same || clamped) pcg_mute(q[1]);
    const int l0 = frame_origin<AR, XS>(q[0].bcb, org, 0);
    if (APPLY) pcg_apply_scatter<T, 2, false, XS>(sp, arena, l0, nullptr, q);
    else pcg_drive_scatter<T, 2, false, XS>((T(*)[4])arena, l0, nullptr, q);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (!alone[j]) continue;
        PcgPart<T> q1[1];
        pcg_load<T, !APPLY>(A, i0 + j, A.ng, q1[0]);
        int lp[3][4];
        frame_nodes<AR>(q1[0].bcb, org, 0, lp);
        if (APPLY) pcg_apply_scatter<T, 1, true, XS>(sp, arena, 0, lp, q1);
        else pcg_drive_scatter<T, 1, true, XS>((T(*)[4])arena, 0, lp, q1);
    }
}

// Assembly per bucket from the solve records: the cubic scatter of
// (N mcoef, N bcoef, kappa |grad N|^2, N) into the bucket's arena slot, lane
// per cell as in P2G, two particles of a cell per pass (wavefront w:
// particles 2w, 2w + 1, then 2(w + W), ...).
template <typename T>
__global__ void __launch_bounds__(P2G_WG) k_pcg_drive(StepArgs<T> A) {
    constexpr int W = sizeof(T) == 4 ? 4 : 2;
    __shared__ __align__(16) T s_q[W][PAD3][4];   // padded arenas, lane l on cell pad_cell(l)
    const int nact = A.misc[0];
    const int nb = A.nb;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (BucketIter it(nact); it.i < it.end; it.i += it.step) {
        const int bk = A.alist[it.i];
        const int bx = bk / (nb * nb), by = (bk / nb) % nb, bz = bk % nb;
        const int org[3] = {BLK * bx - 1, BLK * by - 1, BLK * bz - 1};
        for (int l = threadIdx.x; l < W * 4 * PAD3; l += P2G_WG) (&s_q[0][0][0])[l] = T(0);
        __syncthreads();
        const long long cell = ((long long)bk << 6) + pad_cell(lane);
        const int c0 = A.start[cell], cnt = A.start[cell + 1] - c0;
        int rounds = cnt;
        for (int o = 32; o > 0; o >>= 1) rounds = max(rounds, __shfl_xor(rounds, o));
        for (int r = 2 * wv; r < rounds; r += 2 * W) {
            if (wv >= W || r >= cnt) continue;
            pcg_pair<T, false, PAD_XS>(A, c0 + r, r + 1 < cnt, org, nullptr, &s_q[wv][0][0]);
        }
        __syncthreads();
        T* dst = A.arena + (size_t)it.i * (6 * AR3);
        for (int l = threadIdx.x; l < 4 * AR3; l += P2G_WG) {
            const int node = l / 4, c = l - node * 4, e = pad_entry(node);
            T v = s_q[0][e][c];
#pragma unroll
            for (int w = 1; w < W; ++w) v += s_q[w][e][c];
            dst[6 * node + c] = v;
        }
        __syncthreads();
    }
}

// K p through the particles, per bucket: stage the bucket's 343 nodes of the
// search direction (p_new = r / diag + beta p, or p itself when DIRECT) in
// LDS, then per pair of a cell's particles g = kappa sum_i grad N_i p_i and
// the scatter of grad N_i . g into the arena slot (channel 0).
template <typename T, bool DIRECT>
__global__ void __launch_bounds__(P2G_WG) k_pcg_apply(StepArgs<T> A) {
    if (A.pscal[PCG_DONE] != 0.0) return;
    constexpr int W = sizeof(T) == 4 ? 4 : 2;
    __shared__ T s_p[PAD3];                // padded, lane l on cell pad_cell(l)
    __shared__ T s_a[W][PAD3];
    const int nact = A.misc[0];
    const int nb = A.nb, ng = A.ng;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const T beta = T(A.pscal[PCG_BETA]);
    for (BucketIter it(nact); it.i < it.end; it.i += it.step) {
        const int bk = A.alist[it.i];
        const int bx = bk / (nb * nb), by = (bk / nb) % nb, bz = bk % nb;
        const int org[3] = {BLK * bx - 1, BLK * by - 1, BLK * bz - 1};
        for (int l = threadIdx.x; l < AR3; l += P2G_WG) {
            const int lx = l / (AR * AR), ly = (l / AR) % AR, lz = l % AR;
            // arena positions beyond the grid read the edge node (the
            // reference's clamped node index)
            const long long node = node_lin(org[0] + lx, org[1] + ly, org[2] + lz, ng);
            const T* v = A.gv + 4 * node;
            s_p[pad_entry(l)] = DIRECT ? v[2] : v[1] / A.pdiag[2 * node] + beta * v[2];
        }
        for (int l = threadIdx.x; l < W * PAD3; l += P2G_WG) (&s_a[0][0])[l] = T(0);
        __syncthreads();
        const long long cell = ((long long)bk << 6) + pad_cell(lane);
        const int c0 = A.start[cell], cnt = A.start[cell + 1] - c0;
        int rounds = cnt;
        for (int o = 32; o > 0; o >>= 1) rounds = max(rounds, __shfl_xor(rounds, o));
        for (int r = 2 * wv; r < rounds; r += 2 * W) {
            if (wv >= W || r >= cnt) continue;
            pcg_pair<T, true, PAD_XS>(A, c0 + r, r + 1 < cnt, org, s_p, s_a[wv]);
        }
        __syncthreads();
        T* dst = A.arena + (size_t)it.i * (6 * AR3);
        for (int l = threadIdx.x; l < AR3; l += P2G_WG) {
            const int e = pad_entry(l);
            T v = s_a[0][e];
#pragma unroll
            for (int w = 1; w < W; ++w) v += s_a[w][e];
            dst[6 * l] = v;
        }
        __syncthreads();
    }
}

// Per marked block, one wavefront (lane = node).  ASSEMBLE: fold the drive's
// four channels into the Jacobi diagonal, lumped mass, b and x0 = b / m.
// APPLY / APPLY_DIRECT: fold K p, form Ap = m p + K p on the active nodes,
// store the search direction and Ap, and the partial dots p.Ap.
enum { PCG_ASSEMBLE = 0, PCG_APPLY = 1, PCG_APPLY_DIRECT = 2 };
template <typename T, int MODE>
__global__ void __launch_bounds__(WG) k_pcg_fold(StepArgs<T> A) {
    if (MODE != PCG_ASSEMBLE && A.pscal[PCG_DONE] != 0.0) return;
    constexpr int NCH = MODE == PCG_ASSEMBLE ? 4 : 1;
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * (unsigned)WG + threadIdx.x) >> 6);
    const int nwaves = gridDim.x * (WG / 64);
    const int nblocks = A.misc[1];
    const bool off_grid = A.misc[3] != 0;
    const int nb = A.nb, ng = A.ng;
    const T beta = T(A.pscal[PCG_BETA]);
    double part[1] = {0.0};
    for (int i = wave; i < nblocks; i += nwaves) {
        const int blk = A.blist[i];
        const int bx = blk / (nb * nb), by = (blk / nb) % nb, bz = blk % nb;
        const int g[3] = {BLK * bx + (lane >> 4), BLK * by + ((lane >> 2) & 3), BLK * bz + (lane & 3)};
        T acc[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) acc[c] = T(0);
        fold_node<T, NCH, 1>(A, 0, unit_slot_lane<T, 1>(A, 0, bx, by, bz, lane), g, bx, by, bz, off_grid, acc);
        if (g[0] >= ng || g[1] >= ng || g[2] >= ng) continue;
        const long long node = ((long long)g[0] * ng + g[1]) * ng + g[2];
        T* v = A.gv + 4 * node;
        T* dg = A.pdiag + 2 * node;
        if constexpr (MODE == PCG_ASSEMBLE) {
            if (i == 0 && lane == 0) {
                // a new solve: clear the previous one's flag and scalars (no
                // kernel of this launch reads them)
                A.pscal[PCG_DONE] = 0.0; A.pscal[PCG_BETA] = 0.0; A.pscal[PCG_ITER] = 0.0; A.pscal[PCG_REL] = 0.0;
            }
            const bool act = acc[3] > T(1e-12);
            const T m = acc[0], b = acc[1];
            dg[0] = act ? m + acc[2] : T(1);
            dg[1] = act ? m : T(0);
            const T x0 = act ? b / m : T(0);
            v[0] = x0; v[1] = act ? b : T(0); v[2] = x0; v[3] = T(0);
        } else {
            const T m = dg[1];
            const T pn = MODE == PCG_APPLY_DIRECT ? v[2] : v[1] / dg[0] + beta * v[2];
            const T ap = m > T(0) ? m * pn + acc[0] : T(0);
            v[2] = pn;
            v[3] = ap;
            part[0] += (double)pn * (double)ap;
        }
    }
    if constexpr (MODE == PCG_APPLY) store_partials<1>(part, A.ppart);
}

// r = b - A x0 (r holds b, Ap holds A x0), p = 0; partial r.z, r.r, b.D^-1 b.
template <typename T>
__global__ void __launch_bounds__(WG) k_pcg_init(StepArgs<T> A) {
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * (unsigned)WG + threadIdx.x) >> 6);
    const int nwaves = gridDim.x * (WG / 64);
    const int nblocks = A.misc[1];
    const int nb = A.nb, ng = A.ng;
    double part[3] = {0.0, 0.0, 0.0};
    for (int i = wave; i < nblocks; i += nwaves) {
        const int blk = A.blist[i];
        const int bx = blk / (nb * nb), by = (blk / nb) % nb, bz = blk % nb;
        const int g[3] = {BLK * bx + (lane >> 4), BLK * by + ((lane >> 2) & 3), BLK * bz + (lane & 3)};
        if (g[0] >= ng || g[1] >= ng || g[2] >= ng) continue;
        const long long node = ((long long)g[0] * ng + g[1]) * ng + g[2];
        T* v = A.gv + 4 * node;
        const T b = v[1];
        const T r = b - v[3];
        const T dinv = T(1) / A.pdiag[2 * node];
        v[1] = r;
        v[2] = T(0);
        part[0] += (double)r * (double)(r * dinv);
        part[1] += (double)r * (double)r;
        part[2] += (double)b * (double)(b * dinv);
    }
    store_partials<3>(part, A.ppart);
}

// x += alpha p, r -= alpha Ap; partial r.z, r.r.
template <typename T>
__global__ void __launch_bounds__(WG) k_pcg_update(StepArgs<T> A) {
    if (A.pscal[PCG_DONE] != 0.0) return;
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * (unsigned)WG + threadIdx.x) >> 6);
    const int nwaves = gridDim.x * (WG / 64);
    const int nblocks = A.misc[1];
    const int nb = A.nb, ng = A.ng;
    const T alpha = T(A.pscal[PCG_ALPHA]);
    double part[2] = {0.0, 0.0};
    for (int i = wave; i < nblocks; i += nwaves) {
        const int blk = A.blist[i];
        const int bx = blk / (nb * nb), by = (blk / nb) % nb, bz = blk % nb;
        const int g[3] = {BLK * bx + (lane >> 4), BLK * by + ((lane >> 2) & 3), BLK * bz + (lane & 3)};
        if (g[0] >= ng || g[1] >= ng || g[2] >= ng) continue;
        const long long node = ((long long)g[0] * ng + g[1]) * ng + g[2];
        T* v = A.gv + 4 * node;
        v[0] = v[0] + alpha * v[2];
        const T r = v[1] - alpha * v[3];
        v[1] = r;
        const T z = r / A.pdiag[2 * node];
        part[0] += (double)r * (double)z;
        part[1] += (double)r * (double)r;
    }
    store_partials<2>(part, A.ppart);
}

// One work-group: the partial dots in a fixed order, then the CG scalars.
enum { PCG_RED_INIT = 0, PCG_RED_ALPHA = 1, PCG_RED_BETA = 2 };
// hflag (host-mapped, NULL under stream capture): thread 0 publishes
// (seq << 32) | iterations done, with bit 31 set once the solve has stopped,
// as one 8-byte system-scope store, for the host's wait between iteration
// chunks.
template <int MODE>
__global__ void __launch_bounds__(WG) k_pcg_reduce(const double* __restrict__ part, int nparts, double* __restrict__ sc,
                                                   double tol, int max_iter, unsigned long long* hflag, unsigned seq) {
    if (MODE != PCG_RED_INIT && sc[PCG_DONE] != 0.0) return;
    constexpr int NCH = MODE == PCG_RED_INIT ? 3 : (MODE == PCG_RED_ALPHA ? 1 : 2);
    __shared__ double s[NCH][WG];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        double t = 0;
        for (int i = threadIdx.x; i < nparts; i += WG) t += part[c * GRID_WG_COUNT + i];
        s[c][threadIdx.x] = t;
    }
    __syncthreads();
    for (int h = WG / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int c = 0; c < NCH; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    // convergence is measured on the preconditioned residual,
    // sqrt(r^T D^-1 r / b^T D^-1 b): it weights every node by its own
    // diagonal, so nodes with a thin particle support (small lumped mass)
    // count, where the plain |r| / |b| would let their error through
    if constexpr (MODE == PCG_RED_INIT) {
        const double rz = s[0][0], bb = s[2][0];
        const double rel = bb > 0 ? sqrt(fmax(rz, 0.0) / bb) : 0.0;
        sc[PCG_RZ] = rz; sc[PCG_BB] = bb; sc[PCG_BETA] = 0.0; sc[PCG_ITER] = 0.0; sc[PCG_REL] = rel;
        sc[PCG_DONE] = (rel <= tol || max_iter <= 0) ? 1.0 : 0.0;
    } else if constexpr (MODE == PCG_RED_ALPHA) {
        const double pap = s[0][0];
        if (pap > 0) {
            sc[PCG_ALPHA] = sc[PCG_RZ] / pap;
        } else {
            sc[PCG_ALPHA] = 0.0;
            sc[PCG_DONE] = 1.0;           // breakdown: keep the current iterate
        }
    } else {
        const double rz = s[0][0], bb = sc[PCG_BB];
        const double it = sc[PCG_ITER] + 1.0;
        const double rel = bb > 0 ? sqrt(fmax(rz, 0.0) / bb) : 0.0;
        sc[PCG_BETA] = sc[PCG_RZ] > 0 ? rz / sc[PCG_RZ] : 0.0;
        sc[PCG_RZ] = rz;
        sc[PCG_ITER] = it;
        sc[PCG_REL] = rel;
        sc[PCG_DONE] = (rel <= tol || it >= (double)max_iter) ? 1.0 : 0.0;
    }
    if (hflag) {
        // low word: the iterations done, with the top bit set once the solve has stopped
        const unsigned lo = (unsigned)sc[PCG_ITER] | (sc[PCG_DONE] != 0.0 ? 0x80000000u : 0u);
        __hip_atomic_store(hflag, ((unsigned long long)seq << 32) | lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Per host thread and device, 64 bytes of pinned host memory the solve's
// reductions publish their state to (allocated once, held until the process
// ends), and the sequence number of the solve using it.
struct PcgFlag {
    int dev;
    unsigned long long *host, *dptr;
    unsigned seq;
    int hint;       // iterations the last solve polled to its end took: the next solve's first chunk
};
thread_local std::vector<PcgFlag> t_pcg_flags;

PcgFlag* pcg_flag() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    for (auto& f : t_pcg_flags)
        if (f.dev == dev) return &f;
    void* h = nullptr;
    void* d = nullptr;
    if (hipHostMalloc(&h, 64, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) return nullptr;
    if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) return nullptr;
    memset(h, 0, 64);
    t_pcg_flags.push_back({dev, (unsigned long long*)h, (unsigned long long*)d, 0, 1});
    return &t_pcg_flags.back();
}

// Wait for the published state of solve seq to show `issued` iterations
// done or the solve stopped: 1 stopped (ran = the iterations it took), 0 not
// yet, -1 not seen within a second (the caller then reads the device flag
// the synchronous way).
int pcg_wait(const PcgFlag& f, int issued, int& ran) {
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spin = 1;; ++spin) {
        const unsigned long long v = __atomic_load_n(f.host, __ATOMIC_ACQUIRE);
        if ((unsigned)(v >> 32) == f.seq) {
            const unsigned lo = (unsigned)v;
            if (lo & 0x80000000u) {
                ran = (int)(lo & 0x7FFFFFFFu);
                return 1;
            }
            if (lo >= (unsigned)issued) return 0;
        }
        if ((spin & 255) == 0) {
            const auto dt = std::chrono::steady_clock::now() - t0;
            if (dt > std::chrono::seconds(1)) return -1;
            if (dt > std::chrono::microseconds(200)) std::this_thread::yield();
        }
    }
}

// New particle damage: the interpolant of the nodal solution, clamped to
// [0, 1] and made irreversible where allowed, as the explicit update's value
// is (anisompm.py:684-685).  Per bucket: the bucket's 343 nodes of the
// solution staged in LDS (positions beyond the grid read the edge node, the
// reference's clamped node index), then lane per cell over the solve
// records, which also carry the old d and the allow flag.
template <typename T>
__global__ void __launch_bounds__(P2G_WG) k_pcg_interp(StepArgs<T> A) {
    constexpr int W = P2G_WG / 64;
    __shared__ T s_x[AR3];
    const int nact = A.misc[0];
    const int nb = A.nb, ng = A.ng;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (BucketIter it(nact); it.i < it.end; it.i += it.step) {
        const int bk = A.alist[it.i];
        const int bx = bk / (nb * nb), by = (bk / nb) % nb, bz = bk % nb;
        const int org[3] = {BLK * bx - 1, BLK * by - 1, BLK * bz - 1};
        for (int l = threadIdx.x; l < AR3; l += P2G_WG) {
            const int lx = l / (AR * AR), ly = (l / AR) % AR, lz = l % AR;
            s_x[l] = A.gv[4 * node_lin(org[0] + lx, org[1] + ly, org[2] + lz, ng)];
        }
        __syncthreads();
        const long long cell = ((long long)bk << 6) + lane;
        const int c0 = A.start[cell], cnt = A.start[cell + 1] - c0;
        for (int r = wv; r < cnt; r += W) {
            const T* rc = (const T*)__builtin_assume_aligned(A.rec + (size_t)PCG_REC * (c0 + r), 4 * sizeof(T));
            const T Xp[3] = {scaled(rc[0], A.inv_dx), scaled(rc[1], A.inv_dx), scaled(rc[2], A.inv_dx)};
            const T dold = rc[6];
            const bool allow = rc[7] != T(0);
            const long long p = A.perm[c0 + r];
            int bcb[3];
            T wc[4][3], w2[4][3];
            cubic_weights(Xp, ng, bcb, wc, w2);
            int lp[3][4];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int k = 0; k < 4; ++k) lp[a][k] = clampi(bcb[a] + k - org[a], 0, AR - 1);
            T di = T(0);
#pragma unroll
            for (int ox = 0; ox < 4; ++ox)
#pragma unroll
                for (int oy = 0; oy < 4; ++oy)
#pragma unroll
                    for (int oz = 0; oz < 4; ++oz)
                        di += (wc[ox][0] * wc[oy][1] * wc[oz][2]) * s_x[(lp[0][ox] * AR + lp[1][oy]) * AR + lp[2][oz]];
            const T dn = minp(maxp(di, T(0)), T(1));
            A.od[p] = allow ? maxp(dn, dold) : dold;
        }
        __syncthreads();
    }
}

// Zero n bytes at a 16-byte aligned p (a fill kernel rather than
// hipMemsetAsync, so that a captured step holds no memset node).
__global__ void __launch_bounds__(WG) k_zero_bytes(unsigned char* __restrict__ p, long long n) {
    const long long n16 = n >> 4;
    const long long stride = (long long)gridDim.x * WG;
    for (long long i = (long long)blockIdx.x * WG + threadIdx.x; i < n16; i += stride)
        reinterpret_cast<uint4*>(p)[i] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && (long long)threadIdx.x < (n & 15)) p[(n16 << 4) + threadIdx.x] = 0;
}

#include "pmpm_adjoint.h"
#include "pmpm_carry.h"
#include "pmpm_observe.h"
#include "pmpm_scene.h"
#include "pmpm_variational.h"

// Option B on its own (the interface branch of the step's stress): tau =
// the projection split of tau_c by m = F n0 / |F n0| with g(d) = (1 - d)^2
// (1 - k_r) + k_r for both modes, through the step's own degrade().
template <typename T>
__global__ void __launch_bounds__(WG) k_dir_kirchhoff(long long n, const T* __restrict__ tau_c, const T* __restrict__ Fp,
                                                      const T* __restrict__ dp, const T* __restrict__ n0p,
                                                      const T* __restrict__ res_a, T res_s, int gate, T* __restrict__ out) {
    const long long p = (long long)blockIdx.x * WG + threadIdx.x;
    if (p >= n) return;
    const T n0[3] = {n0p[3 * p], n0p[3 * p + 1], n0p[3 * p + 2]};
    const M3<T> tc = load_m3(tau_c, p);
    store_m3(out, p, degrade<T>(res_a ? res_a[p] : res_s, load_m3(Fp, p), dp[p], T(1), tc, T(0), true, n0, gate != 0));
}

// ------------------------------------------------------------------ test kernels
template <typename T>
__global__ void __launch_bounds__(WG) k_sym_eig(const T* __restrict__ A, long long n, T* __restrict__ ev, T* __restrict__ V) {
    long long p = (long long)blockIdx.x * WG + threadIdx.x;
    if (p >= n) return;
    M3<T> a = load_m3(A, p), Q;
    T e[3];
    sym_eig3(a, e, Q);
    ev[3 * p] = e[0]; ev[3 * p + 1] = e[1]; ev[3 * p + 2] = e[2];
    store_m3(V, p, Q);
}

template <typename T>
__global__ void __launch_bounds__(WG) k_stencil(const T* __restrict__ x, long long n, int ng, T inv_dx, int* bq_o, int* i27, int* bc_o,
                                                int* i64, T* w_o, T* wc_o, T* w2_o) {
    long long p = (long long)blockIdx.x * WG + threadIdx.x;
    if (p >= n) return;
    T Xp[3] = {scaled(x[3 * p], inv_dx), scaled(x[3 * p + 1], inv_dx), scaled(x[3 * p + 2], inv_dx)};
    int bq[3], bcb[3];
    T fx[3], w[3][3], wc[4][3], w2[4][3];
    quad_weights(Xp, ng, bq, fx, w);
    cubic_weights(Xp, ng, bcb, wc, w2);
    for (int a = 0; a < 3; ++a) {
        // report the reference's unclamped base cells (the kernels clamp internally)
        const T lim = T(1 << 30);
        if (bq_o) bq_o[3 * p + a] = (int)minp(maxp(floor(sub_exact(Xp[a], T(0.5))), -lim), lim);
        if (bc_o) bc_o[3 * p + a] = (int)minp(maxp(floor(Xp[a]), -lim), lim) - 1;
        for (int k = 0; k < 3; ++k)
            if (w_o) w_o[9 * p + 3 * k + a] = w[k][a];
        for (int k = 0; k < 4; ++k) {
            if (wc_o) wc_o[12 * p + 3 * k + a] = wc[k][a];
            if (w2_o) w2_o[12 * p + 3 * k + a] = w2[k][a];
        }
    }
    if (i27)
        for (int k = 0; k < 27; ++k) i27[27 * p + k] = node_lin(bq[0] + k / 9, bq[1] + (k / 3) % 3, bq[2] + k % 3, ng);
    if (i64)
        for (int k = 0; k < 64; ++k) i64[64 * p + k] = node_lin(bcb[0] + k / 16, bcb[1] + (k / 4) % 4, bcb[2] + k % 4, ng);
}

// ------------------------------------------------------------------ host side
inline unsigned nblocks(long long n) { return (unsigned)((n + WG - 1) / WG); }

// The second stream of a device (launch shape const_stream 2): the
// constitutive kernel runs on it beside the binning.  fork is recorded on the
// launch stream and waited for by the second stream; join the other way.
// Default priority, non-blocking (no implicit ordering with the null stream:
// the two events order it).  Made once per device and kept until the process
// ends.  g_side_mu also serialises a step's fork ... join enqueueing, so the
// two events are never re-recorded by another host thread in between.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
constexpr int SIDE_MAX_DEVICES = 64;
std::mutex g_side_mu;
SideStream g_side[SIDE_MAX_DEVICES];

int side_stream_locked(SideStream** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= SIDE_MAX_DEVICES) return fail(PMPM_EARG, "%sdevice ordinal out of range: %lld", "", dev);
    SideStream& ss = g_side[dev];
    if (!ss.s) {
        HIPCHK(hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ss.join, hipEventDisableTiming));
        HIPCHK(hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking));
    }
    *out = &ss;
    return PMPM_OK;
}

int side_stream(SideStream** out) {
    std::lock_guard<std::mutex> lk(g_side_mu);
    return side_stream_locked(out);
}

// The grid pass over the P2G arenas of the launch's cell block (A.cb).
template <typename T> void launch_grid(const StepArgs<T>& A, int wgs, hipStream_t st) {
    if (A.cb == 8) hipLaunchKernelGGL((k_grid<T, 2>), dim3(wgs), dim3(WG), 0, st, A);
    else hipLaunchKernelGGL((k_grid<T, 1>), dim3(wgs), dim3(WG), 0, st, A);
}

// Constitutive update (particle order).
template <typename T> void launch_const(const StepArgs<T>& A, bool damage, bool exact_sigma, hipStream_t st) {
    const int ct = g_launch.const_threads;
    const dim3 PB((unsigned)((A.n + ct - 1) / ct)), B(ct);
    if (damage && exact_sigma) hipLaunchKernelGGL((k_const<T, true, true>), PB, B, 0, st, A);
    else if (damage) hipLaunchKernelGGL((k_const<T, true, false>), PB, B, 0, st, A);
    else hipLaunchKernelGGL((k_const<T, false, false>), PB, B, 0, st, A);
}

// The PCG mode's constitutive update up to the damage drive and its solve
// records (particle order).
template <typename T> void launch_pcg_const(const StepArgs<T>& A, bool exact_sigma, hipStream_t st) {
    const int ct = g_launch.const_threads;
    const dim3 CG((unsigned)((A.n + ct - 1) / ct)), CB(ct);
    if (exact_sigma) hipLaunchKernelGGL((k_pcg_const<T, true>), CG, CB, 0, st, A);
    else hipLaunchKernelGGL((k_pcg_const<T, false>), CG, CB, 0, st, A);
}

// P2G (bucket order) from the constitutive kernel's records.
template <typename T> void launch_p2g(const StepArgs<T>& A, hipStream_t st) {
    const dim3 G(g_launch.bucket_wgs);
    if (A.cb == 8) {
        if (g_launch.p2g_pair == 1) hipLaunchKernelGGL((k_p2g_cb8<T, 1>), G, dim3(64), 0, st, A);
        else hipLaunchKernelGGL((k_p2g_cb8<T, 2>), G, dim3(64), 0, st, A);
        return;
    }
    if (g_launch.p2g_split == 2) {
        if (g_launch.p2g_pair == 1) hipLaunchKernelGGL((k_p2g_chan<T, 1, false>), G, dim3(128), 0, st, A);
        else hipLaunchKernelGGL((k_p2g_chan<T, 2, false>), G, dim3(128), 0, st, A);
        return;
    }
    if (g_launch.p2g_split == 3) {
        if (g_launch.p2g_pair == 1) hipLaunchKernelGGL((k_p2g_chan<T, 1, true>), G, dim3(128), 0, st, A);
        else hipLaunchKernelGGL((k_p2g_chan<T, 2, true>), G, dim3(128), 0, st, A);
        return;
    }
    // four fp64 arenas would exceed 64 KiB of LDS per work-group: fp64 always takes two
    if constexpr (sizeof(T) == 4) {
        if (g_launch.p2g_waves == 4) {
            if (g_launch.p2g_pair == 1) hipLaunchKernelGGL((k_p2g<T, 1, 4>), G, dim3(256), 0, st, A);
            else hipLaunchKernelGGL((k_p2g<T, 2, 4>), G, dim3(256), 0, st, A);
            return;
        }
    }
    if (g_launch.p2g_waves == 1) {
        if (g_launch.p2g_pair == 1) hipLaunchKernelGGL((k_p2g<T, 1, 1>), G, dim3(64), 0, st, A);
        else hipLaunchKernelGGL((k_p2g<T, 2, 1>), G, dim3(64), 0, st, A);
        return;
    }
    {
        if (g_launch.p2g_pair == 1) hipLaunchKernelGGL((k_p2g<T, 1, 2>), G, dim3(128), 0, st, A);
        else hipLaunchKernelGGL((k_p2g<T, 2, 2>), G, dim3(128), 0, st, A);
    }
}

// Constitutive update (particle order) then P2G (bucket order).
template <typename T> void launch_transfer(const StepArgs<T>& A, bool damage, bool exact_sigma, hipStream_t st) {
    launch_const<T>(A, damage, exact_sigma, st);
    launch_p2g<T>(A, st);
}

// The binning.  pre: the carried count of the previous step's G2P (count2,
// btot2, tiles2, and the keys and slots it wrote) is this x's, so the count pass is
// skipped; otherwise k_bin_count counts afresh and the scan also clears the
// carried set.
template <typename T>
int run_bin(const T* x, long long n, int ng, double inv_dx, int stable, const long long* scene_off, int ns, const Layout& L,
            char* ws, hipStream_t st, bool pre = false) {
    const int nb = nblk(ng);
    int* count = (int*)(ws + (pre ? L.count2 : L.count));
    int* start = (int*)(ws + L.start);
    int* btot = (int*)(ws + L.btot);
    int* bstart = (int*)(ws + L.bstart);
    int* alist = (int*)(ws + L.alist);
    int* slotof = (int*)(ws + L.slotof);
    int* blist = (int*)(ws + L.blist);
    int* key = (int*)(ws + L.key);
    int* slot = (int*)(ws + L.slot);
    int* perm = (int*)(ws + L.perm);
    int* inv = (int*)(ws + L.inv);
    int* misc = (int*)(ws + L.misc);
    int* bflag = (int*)(ws + (pre ? L.bflag2 : L.bflag));
    int* tiles = (int*)(ws + L.tiles);
    int* mark = (int*)(ws + L.mark);
    TimedScope ts(0, st);
    if (n > 0 && !pre)
        hipLaunchKernelGGL(k_bin_count<T>, dim3(nblocks(n)), dim3(WG), 0, st, x, n, T(inv_dx), ng, nb, scene_off, ns, count, bflag,
                           key, slot, misc);
    const dim3 TG((unsigned)L.ntiles);
    hipLaunchKernelGGL(k_bucket_tiles, TG, dim3(WG), 0, st, bflag, count, L.NB, btot, tiles);
    hipLaunchKernelGGL(k_bucket_apply, TG, dim3(WG), 0, st, bflag, count, btot, L.NB, nb, tiles, bstart, alist, slotof, mark,
                       start, misc, pre ? nullptr : (int*)(ws + L.bflag2), pre ? nullptr : (int*)(ws + L.count2));
    const long long nmark = nblocks(L.NB);
    hipLaunchKernelGGL(k_bin_scatter, dim3((unsigned)(nmark + (n > 0 ? nblocks(n) : 0))), dim3(WG), 0, st, n, key, slot, start,
                       perm, inv, nmark, mark, L.NB, blist, misc);
    if (stable && n > 0)
        hipLaunchKernelGGL(k_bin_stabilize, dim3(STAB_WGS), dim3(64), 0, st, start, alist, misc, perm, inv);
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}

template <typename T> StepArgs<T> make_args(const pmpm_step_params* p, const pmpm_particles_in* in, const pmpm_particles_out* out,
                                           const Layout& L, char* ws) {
    StepArgs<T> A;
    memset(&A, 0, sizeof A);
    A.x = (const T*)in->x; A.v = (const T*)in->v; A.F = (const T*)in->F; A.C = (const T*)in->C;
    A.d = (const T*)in->d; A.lap = (const T*)in->lap; A.D_hist = (const T*)in->D_hist;
    A.vol0 = (const T*)in->vol0; A.mass = (const T*)in->mass; A.mu = (const T*)in->mu; A.lam = (const T*)in->lam;
    A.sigma_c = (const T*)in->sigma_c; A.eta = (const T*)in->eta; A.zeta = (const T*)in->zeta; A.l0 = (const T*)in->l0;
    A.residual = (const T*)in->residual; A.A0 = (const T*)in->A0; A.n_if = (const T*)in->n_interface;
    A.umask = in->uniform_mask & ((1u << 10) - 1u);
    for (int k = 0; k < 9; ++k) A.uval[k] = T(in->uniform_value[k]);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) A.uA0.m[a][b] = T(in->uniform_A0[3 * a + b]);
    A.allow = in->allow; A.imask = in->interface_mask; A.bc_id = in->bc_id;
    A.ox = (T*)out->x; A.ov = (T*)out->v; A.oF = (T*)out->F; A.oC = (T*)out->C; A.olap = (T*)out->lap;
    A.od = (T*)out->d; A.oD = (T*)out->D_hist; A.ophi = (T*)out->phi; A.ophs = (T*)out->phi_shear;
    A.otau = (T*)out->tau_last; A.oJ = (T*)out->J_last;
    A.arena = (T*)(ws + L.arena); A.gv = (T*)(ws + L.gv);
    A.active = (uint8_t*)(ws + L.active);
    A.start = (const int*)(ws + L.start); A.alist = (const int*)(ws + L.alist); A.perm = (const int*)(ws + L.perm);
    A.key = (const int*)(ws + L.key); A.inv = (const int*)(ws + L.inv);
    A.rec = (T*)(ws + L.rec); A.dsrt = (T*)(ws + L.dsrt);
    A.slotof = (const int*)(ws + L.slotof); A.blist = (const int*)(ws + L.blist);
    A.misc = (const int*)(ws + L.misc);
    A.ckey = (int*)(ws + L.key); A.cslot = (int*)(ws + L.slot);
    A.ccount = (int*)(ws + L.count2); A.cbflag = (int*)(ws + L.bflag2); A.cmisc = (int*)(ws + L.misc);
    A.bin_next = 0;
    A.stats = (unsigned long long*)(ws + L.stats);
    A.gmw = p->keep_for_adjoint ? (T*)(ws + L.gmw) : nullptr;
    A.pdiag = (T*)(ws + L.pcgd);
    A.ppart = (double*)(ws + L.pcgr);
    A.pscal = A.ppart + 3 * GRID_WG_COUNT;
    A.pcg_tol = p->pcg_tol;
    A.pcg_max = p->pcg_max_iter;
    A.n = in->n;
    A.ng = p->n_grid; A.nb = nblk(p->n_grid);
    A.NBs = L.NBs; A.Gs = L.Gs;
    A.NBs8 = (long long)A.nb * A.nb * ((A.nb + 1) / 2);
    A.cb = g_launch.cell_block;
    A.nscenes = p->n_scenes > 1 ? p->n_scenes : 1;
    A.scene_dmg = p->scene_damage;
    A.dx = T(p->dx); A.inv_dx = T(p->inv_dx); A.dt = T(p->dt); A.lim = T(p->grid_lim);
    A.pad_lo = T(2.0 * p->dx); A.pad_hi = T(p->grid_lim - 2.0 * p->dx);
    A.fc_lo = T(p->fclamp_lo); A.fc_hi = T(p->fclamp_hi); A.use_fclamp = p->use_fclamp;
    A.dt_dmg = T(p->dt_damage); A.damp = T(p->grid_damp);
    A.c_scale = T(4.0 * p->inv_dx * p->inv_dx);
    A.lap_scale = T(p->inv_dx * p->inv_dx);
    for (int a = 0; a < 3; ++a) A.dtg[a] = T(p->dt) * T(p->gravity[a]);
    A.export_active = p->export_active;
    A.n_col = p->n_colliders;
    for (int k = 0; k < p->n_colliders; ++k) A.col[k] = p->colliders[k];
    A.btab = p->boundary_table;
    A.n_bc = p->n_vel_bc;
    for (int k = 0; k < p->n_vel_bc; ++k)
        for (int a = 0; a < 3; ++a) A.bcv[k][a] = T(p->vel_bc[k][a]);
    return A;
}

// Arguments of a step call.  ws: the caller's workspace, checked against
// the shape pmpm_workspace_init set it up for (null: the library's own
// scratch, laid out per call).
int check_step(const pmpm_step_params* p, const pmpm_particles_in* in, const pmpm_particles_out* out, const void* ws,
               size_t ws_bytes, int64_t& max_n_out) {
    if (!p || !in || !out) return fail(PMPM_EARG, "%snull argument", "");
    if (p->dtype != PMPM_F32 && p->dtype != PMPM_F64) return fail(PMPM_EARG, "%sunknown dtype %lld", "", p->dtype);
    if (p->n_grid < 4 || p->n_grid > 1024) return fail(PMPM_EARG, "%sn_grid out of range: %lld", "", p->n_grid);
    if (in->n < 0 || in->n > 0x7fffffffLL) return fail(PMPM_EARG, "%sparticle count out of range: %lld", "", in->n);
    if (p->n_colliders < 0 || p->n_colliders > PMPM_MAX_COLLIDERS) return fail(PMPM_EARG, "%stoo many colliders: %lld", "", p->n_colliders);
    if (p->n_vel_bc < 0 || p->n_vel_bc > PMPM_MAX_VEL_BC) return fail(PMPM_EARG, "%stoo many velocity BCs: %lld", "", p->n_vel_bc);
    if (p->n_scenes > 1 && !p->scene_offsets) return fail(PMPM_EARG, "%sbatched step needs scene_offsets", "");
    if (p->damage_solver != PMPM_DAMAGE_EXPLICIT && p->damage_solver != PMPM_DAMAGE_PCG)
        return fail(PMPM_EARG, "%sunknown damage_solver %lld", "", p->damage_solver);
    if (p->damage_solver == PMPM_DAMAGE_PCG) {
        if (p->n_scenes > 1) return fail(PMPM_EARG, "%sthe PCG damage solve takes one scene", "");
        if (p->pcg_max_iter < 0 || p->pcg_max_iter > 10000)
            return fail(PMPM_EARG, "%spcg_max_iter out of range: %lld", "", p->pcg_max_iter);
        if (!(p->pcg_tol >= 0.0)) return fail(PMPM_EARG, "%spcg_tol must be >= 0", "");
    }
    if (p->bin_next && p->keep_for_adjoint) return fail(PMPM_EARG, "%sbin_next does not combine with keep_for_adjoint", "");
    Layout L(in->n, p->n_grid, p->dtype, p->n_scenes);
    if (ws_bytes < L.total) return fail(PMPM_ESIZE, "%sworkspace too small, need %lld bytes", "", (long long)L.total);
    if (ws) {
        const int rc = ws_check(ws, in->n, p->n_grid, p->n_scenes, p->dtype);
        if (rc) return rc;
    }
    if (in->n > 0) {
        // a material field given as a uniform value needs no array
        auto fieldp = [&](const void* a, int k) { return ((in->uniform_mask >> k) & 1u) ? (const void*)in : a; };
        const void* need[] = {in->x, in->v, in->F, in->C, in->d, in->lap, fieldp(in->vol0, PMPM_U_VOL0),
                              fieldp(in->mass, PMPM_U_MASS), fieldp(in->mu, PMPM_U_MU), fieldp(in->lam, PMPM_U_LAM),
                              fieldp(in->residual, PMPM_U_RESIDUAL), in->interface_mask, out->x, out->v, out->F, out->C,
                              out->lap};
        for (const void* q : need)
            if (!q) return fail(PMPM_EARG, "%smissing particle array", "");
        if (in->n_interface == nullptr) return fail(PMPM_EARG, "%smissing n_interface", "");
        if (!out->tau_last != !out->J_last) return fail(PMPM_EARG, "%stau_last and J_last are skipped together", "");
        if (p->do_damage) {
            const void* dm[] = {in->D_hist, fieldp(in->sigma_c, PMPM_U_SIGMA_C), fieldp(in->eta, PMPM_U_ETA),
                                fieldp(in->zeta, PMPM_U_ZETA), fieldp(in->l0, PMPM_U_L0), fieldp(in->A0, PMPM_U_A0),
                                in->allow, out->d, out->D_hist};
            for (const void* q : dm)
                if (!q) return fail(PMPM_EARG, "%smissing damage array", "");
            if (!out->phi != !out->phi_shear) return fail(PMPM_EARG, "%sphi and phi_shear are skipped together", "");
        }
        if (out->x == in->x || out->F == in->F || out->v == in->v || out->C == in->C)
            return fail(PMPM_EARG, "%soutputs must not alias inputs", "");
    }
    max_n_out = in->n;
    return PMPM_OK;
}

// carry: the workspace may carry binning between steps (not the
// no-allocation scratch, which is cleared every call).
// side_ok: the step may put its constitutive kernel on the device's second
// stream (launch shape const_stream 2); not for the no-allocation scratch.
template <typename T>
int step_begin(const pmpm_step_params* p, const pmpm_particles_in* in, const pmpm_particles_out* out, char* ws, hipStream_t st,
               int exact_sigma, bool host_poll = true, bool carry = true, bool side_ok = true) {
    Layout L(in->n, p->n_grid, p->dtype, p->n_scenes);
    const bool pre = carry && carry_take(ws, p->prebinned != 0, in->x, in->n, st);
    StepArgs<T> A = make_args<T>(p, in, out, L, ws);
    const bool pcg = p->do_damage && p->damage_solver == PMPM_DAMAGE_PCG && in->n > 0;
    // The constitutive kernel beside the binning: it reads the particle
    // arrays and writes its outputs and its records in particle order, none
    // of which the binning touches; it reads the cell keys only for a
    // batched damage gate, and only a fresh count (k_bin_count, in run_bin)
    // writes them.  P2G, after the join, reads the records through perm.
    const bool side = side_ok && g_launch.const_stream == 2 && in->n > 0 && !pcg && !p->keep_for_adjoint &&
                      (pre || A.scene_dmg == nullptr);
    if (side) {
        std::lock_guard<std::mutex> lk(g_side_mu);
        SideStream* ss = nullptr;
        if (int rc = side_stream_locked(&ss)) return rc;
        A.rec_pp = 1;
        HIPCHK(hipEventRecord(ss->fork, st));
        HIPCHK(hipStreamWaitEvent(ss->s, ss->fork, 0));
        {
            TimedScope ts(6, ss->s);
            launch_const<T>(A, p->do_damage != 0, exact_sigma != 0, ss->s);
        }
        HIPCHK(hipEventRecord(ss->join, ss->s));
        int rc = run_bin<T>((const T*)in->x, in->n, p->n_grid, p->inv_dx, p->stable_sort, (const long long*)p->scene_offsets,
                            p->n_scenes, L, ws, st, pre);
        if (rc) {
            // the launch stream still joins the second stream before returning
            (void)hipStreamWaitEvent(st, ss->join, 0);
            return rc;
        }
        HIPCHK(hipStreamWaitEvent(st, ss->join, 0));
    } else {
        int rc = run_bin<T>((const T*)in->x, in->n, p->n_grid, p->inv_dx, p->stable_sort, (const long long*)p->scene_offsets,
                            p->n_scenes, L, ws, st, pre);
        if (rc) return rc;
    }
    if (p->export_active)
        hipLaunchKernelGGL(k_zero_bytes, dim3(nblocks(((long long)L.G + 15) / 16 < 1 ? 1 : ((long long)L.G + 15) / 16)),
                           dim3(WG), 0, st, (unsigned char*)(ws + L.active), (long long)L.G);
    if (pcg) {
        TimedScope ts(4, st);
        const dim3 PG(g_launch.bucket_wgs), PB(P2G_WG), GG(g_launch.block_wgs), GB(WG);
        const int nparts = g_launch.block_wgs;
        launch_pcg_const<T>(A, exact_sigma != 0, st);
        hipLaunchKernelGGL(k_pcg_drive<T>, PG, PB, 0, st, A);
        hipLaunchKernelGGL((k_pcg_fold<T, PCG_ASSEMBLE>), GG, GB, 0, st, A);
        hipLaunchKernelGGL((k_pcg_apply<T, true>), PG, PB, 0, st, A);            // A x0
        hipLaunchKernelGGL((k_pcg_fold<T, PCG_APPLY_DIRECT>), GG, GB, 0, st, A);
        hipLaunchKernelGGL(k_pcg_init<T>, GG, GB, 0, st, A);
        // Iterations go out in chunks with a read of the done flag between
        // chunks, so a solve that converges in a few iterations does not pay
        // for pcg_max_iter launches: first as many as the last solve polled
        // to its end took (none when it converged at the start: the flag is
        // read right after the start's reduction), then 1, 2, 4, ...
        // Iterations issued past convergence exit at their first load but
        // still cost their launches.  The reductions publish the solve's
        // state to pinned host memory, which the host polls: no copy and no
        // stream synchronisation between chunks.  Under stream capture no
        // host read is possible: all pcg_max_iter iterations are enqueued and
        // the ones after convergence exit at their first load.
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(st, &cap));
        // enqueue every iteration under capture, and when the caller allows no
        // host-side set-up (the pinned flag is allocated on first use)
        const bool capturing = cap != hipStreamCaptureStatusNone || !host_poll;
        PcgFlag* pf = capturing ? nullptr : pcg_flag();
        unsigned long long* hflag = nullptr;
        unsigned seq = 0;
        if (pf) {
            seq = ++pf->seq;
            hflag = pf->dptr;
        }
        hipLaunchKernelGGL(k_pcg_reduce<PCG_RED_INIT>, dim3(1), GB, 0, st, (const double*)A.ppart, nparts, A.pscal,
                           A.pcg_tol, A.pcg_max, hflag, seq);
        int issued = 0, chunk = pf ? std::max(pf->hint, 0) : 1, next = 1;
        while (issued < p->pcg_max_iter) {
            const int k_end = capturing ? p->pcg_max_iter : std::min(p->pcg_max_iter, issued + chunk);
            for (; issued < k_end; ++issued) {
                TimedScope ti(5, st);
                hipLaunchKernelGGL((k_pcg_apply<T, false>), PG, PB, 0, st, A);
                hipLaunchKernelGGL((k_pcg_fold<T, PCG_APPLY>), GG, GB, 0, st, A);
                hipLaunchKernelGGL(k_pcg_reduce<PCG_RED_ALPHA>, dim3(1), GB, 0, st, (const double*)A.ppart, nparts,
                                   A.pscal, A.pcg_tol, A.pcg_max, hflag, seq);
                hipLaunchKernelGGL(k_pcg_update<T>, GG, GB, 0, st, A);
                hipLaunchKernelGGL(k_pcg_reduce<PCG_RED_BETA>, dim3(1), GB, 0, st, (const double*)A.ppart, nparts,
                                   A.pscal, A.pcg_tol, A.pcg_max, hflag, seq);
            }
            if (capturing || issued >= p->pcg_max_iter) break;
            HIPCHK(hipGetLastError());
            int ran = -1;
            int done = pf ? pcg_wait(*pf, issued, ran) : -1;
            if (done < 0) {
                double flag = 0.0;
                HIPCHK(hipMemcpyAsync(&flag, A.pscal + PCG_DONE, sizeof flag, hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                done = flag != 0.0;
            }
            if (done) {
                if (ran >= 0) pf->hint = ran;
                break;
            }
            chunk = next;
            next *= 2;
        }
        hipLaunchKernelGGL(k_pcg_interp<T>, PG, PB, 0, st, A);
    }
    if (in->n > 0) {
        TimedScope ts(1, st);
        if (pcg) {
            // the solve wrote d: the transfer uses it (no second update)
            StepArgs<T> A2 = A;
            A2.d = A.od;
            launch_transfer<T>(A2, false, false, st);
        } else if (side) {
            launch_p2g<T>(A, st);      // the constitutive kernel ran beside the binning
        } else {
            launch_transfer<T>(A, p->do_damage != 0, exact_sigma != 0, st);
        }
    }
    {
        TimedScope ts(2, st);
        launch_grid<T>(A, g_launch.block_wgs, st);
    }
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}

template <typename T>
int step_end(const pmpm_step_params* p, const pmpm_particles_in* in, const pmpm_particles_out* out, char* ws, hipStream_t st,
             bool carry = true) {
    Layout L(in->n, p->n_grid, p->dtype, p->n_scenes);
    StepArgs<T> A = make_args<T>(p, in, out, L, ws);
    if (carry && p->bin_next && in->n > 0) {
        A.bin_next = 1;
        carry_set(ws, out->x, in->n, st);
    }
    if (in->n > 0) {
        TimedScope ts(3, st);
        const dim3 G(g_launch.bucket_wgs);
        if (g_launch.g2p_threads == 128) hipLaunchKernelGGL((k_g2p<T, 128>), G, dim3(128), 0, st, A);
        else if (g_launch.g2p_threads == 512) hipLaunchKernelGGL((k_g2p<T, 512>), G, dim3(512), 0, st, A);
        else hipLaunchKernelGGL((k_g2p<T, 256>), G, dim3(256), 0, st, A);
    }
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}

template <typename T>
int step_adjoint(const pmpm_step_params* p, const pmpm_particles_in* in, const pmpm_particles_out* fwd,
                 const pmpm_adjoint_in* gin, const pmpm_adjoint_out* gout, char* ws, char* aws, hipStream_t st) {
    Layout L(in->n, p->n_grid, p->dtype, 1);
    StepArgs<T> A = make_args<T>(p, in, fwd, L, ws);
    A.gmw = (T*)(ws + L.gmw);
    if (!p->keep_for_adjoint) {
        // recompute the step's binning, P2G and grid pass, keeping (m, sum w)
        int rc = run_bin<T>((const T*)in->x, in->n, p->n_grid, p->inv_dx, 1, nullptr, 1, L, ws, st);
        if (rc) return rc;
        if (in->n > 0) launch_transfer<T>(A, p->do_damage != 0, false, st);
        launch_grid<T>(A, GRID_WG_COUNT, st);
    }
    AdjArgs<T> B;
    B.gx = (const T*)gin->x; B.gv = (const T*)gin->v; B.gF = (const T*)gin->F; B.gC = (const T*)gin->C;
    B.gd = (const T*)gin->d; B.glap = (const T*)gin->lap; B.gD = (const T*)gin->D_hist;
    B.nv = (const T*)fwd->v; B.nC = (const T*)fwd->C;
    B.ox = (T*)gout->x; B.ov = (T*)gout->v; B.oF = (T*)gout->F; B.oC = (T*)gout->C; B.od = (T*)gout->d;
    B.olap = (T*)gout->lap; B.oD = (T*)gout->D_hist; B.omu = (T*)gout->mu; B.olam = (T*)gout->lam;
    B.oA0 = (T*)gout->A0; B.osc = (T*)gout->sigma_c;
    B.gmw = A.gmw;
    B.gadj = (T*)aws;
    B.pbar = (T*)aws + 6 * L.G;
    B.dmg_on = p->do_damage;
    if (in->n > 0) {
        hipLaunchKernelGGL(k_adj_scatter<T>, dim3(P2G_GRID), dim3(P2G_WG), 0, st, A, B);
        hipLaunchKernelGGL(k_adj_grid<T>, dim3(GRID_WG_COUNT), dim3(WG), 0, st, A, B);
        hipLaunchKernelGGL(k_adj_gather<T>, dim3(nblocks(in->n)), dim3(WG), 0, st, A, B);
        hipLaunchKernelGGL(k_adj_const<T>, dim3(nblocks(in->n)), dim3(WG), 0, st, A, B);
    }
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* pmpm_last_error(void) { return g_err; }
const char* pmpm_version(void) { return "pmpm 0.1 gfx950"; }

int pmpm_abi_sizes(size_t* sizes, int32_t n) {
    const size_t all[8] = {sizeof(pmpm_collider), sizeof(pmpm_step_params), sizeof(pmpm_particles_in),
                           sizeof(pmpm_particles_out), sizeof(pmpm_frame_in), sizeof(pmpm_frame_obs), sizeof(pmpm_pcg64),
                           sizeof(pmpm_boundary_table)};
    int k = 0;
    for (; k < n && k < 8; ++k) sizes[k] = all[k];
    return k;
}

size_t pmpm_workspace_bytes(int64_t max_n, int32_t n_grid, int32_t n_scenes, int32_t dtype) {
    if (n_grid < 1 || max_n < 0) return 0;
    return Layout(max_n, n_grid, dtype, n_scenes).total;
}

int pmpm_workspace_init(void* ws, size_t ws_bytes, int64_t max_n, int32_t n_grid, int32_t n_scenes, int32_t dtype,
                        void* stream) {
    if (!ws) return fail(PMPM_EARG, "%snull workspace", "");
    if (n_grid < 4 || n_grid > 1024 || max_n < 0 || max_n > 0x7fffffffLL || n_scenes < 0)
        return fail(PMPM_EARG, "%sworkspace shape out of range (max_n %lld)", "", (long long)max_n);
    if (dtype != PMPM_F32 && dtype != PMPM_F64) return fail(PMPM_EARG, "%sunknown dtype %lld", "", dtype);
    Layout L(max_n, n_grid, dtype, n_scenes);
    if (ws_bytes < L.total) return fail(PMPM_ESIZE, "%sworkspace too small, need %lld bytes", "", (long long)L.total);
    HIPCHK(hipMemsetAsync(ws, 0, L.total, (hipStream_t)stream));
    ws_register(ws, WsShape{max_n, n_grid, n_scenes < 1 ? 1 : n_scenes, dtype, ws_bytes});
    return PMPM_OK;
}

int pmpm_grid_views(int64_t max_n, int32_t n_grid, int32_t n_scenes, int32_t dtype, size_t* gv_offset,
                    size_t* active_offset) {
    Layout L(max_n, n_grid, dtype, n_scenes);
    if (gv_offset) *gv_offset = L.gv;
    if (active_offset) *active_offset = L.active;
    return PMPM_OK;
}

int pmpm_last_step_stats(const void* ws, int64_t max_n, int32_t n_grid, int32_t n_scenes, int32_t dtype, int64_t* n_active_q,
                         int64_t* n_active_c, void* stream) {
    if (int rc = ws_check(ws, 0, n_grid, n_scenes, dtype)) return rc;
    Layout L(max_n, n_grid, dtype, n_scenes);
    std::vector<unsigned long long> h(2 * GRID_WG_COUNT);
    int misc[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(h.data(), (const char*)ws + L.stats, h.size() * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIPCHK(hipMemcpyAsync(misc, (const char*)ws + L.misc, sizeof misc, hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    const int used = misc[2] > 0 && misc[2] <= GRID_WG_COUNT ? misc[2] : GRID_WG_COUNT;
    for (int i = 1; i < used; ++i) { h[0] += h[2 * i]; h[1] += h[2 * i + 1]; }
    if (n_active_q) *n_active_q = (int64_t)h[0];
    if (n_active_c) *n_active_c = (int64_t)h[1];
    return PMPM_OK;
}

int pmpm_set_launch_config(const pmpm_launch_config* cfg) {
    pmpm_launch_config c = {PERSIST_WG, GRID_WG_COUNT, 2, 256, 2, 256, 4, 1, 1};
    if (cfg) {
        if (cfg->const_stream) c.const_stream = cfg->const_stream;
        if (cfg->cell_block) c.cell_block = cfg->cell_block;
        if (cfg->p2g_split) c.p2g_split = cfg->p2g_split;
        if (cfg->bucket_wgs) c.bucket_wgs = cfg->bucket_wgs;
        if (cfg->block_wgs) c.block_wgs = cfg->block_wgs;
        if (cfg->p2g_pair) c.p2g_pair = cfg->p2g_pair;
        if (cfg->const_threads) c.const_threads = cfg->const_threads;
        if (cfg->p2g_waves) c.p2g_waves = cfg->p2g_waves;
        if (cfg->g2p_threads) c.g2p_threads = cfg->g2p_threads;
    }
    if (c.bucket_wgs < 8 || c.bucket_wgs > 16384 || c.bucket_wgs % 8 != 0)
        return fail(PMPM_EARG, "%sbucket_wgs must be a multiple of 8 in [8, 16384]: %lld", "", c.bucket_wgs);
    if (c.block_wgs < 1 || c.block_wgs > GRID_WG_COUNT)
        return fail(PMPM_EARG, "%sblock_wgs must be in [1, 2048]: %lld", "", c.block_wgs);
    if (c.p2g_pair != 1 && c.p2g_pair != 2) return fail(PMPM_EARG, "%sp2g_pair must be 1 or 2: %lld", "", c.p2g_pair);
    if (c.const_threads != 64 && c.const_threads != 128 && c.const_threads != 256)
        return fail(PMPM_EARG, "%sconst_threads must be 64, 128 or 256: %lld", "", c.const_threads);
    if (c.p2g_waves != 1 && c.p2g_waves != 2 && c.p2g_waves != 4)
        return fail(PMPM_EARG, "%sp2g_waves must be 1, 2 or 4: %lld", "", c.p2g_waves);
    if (c.g2p_threads != 128 && c.g2p_threads != 256 && c.g2p_threads != 512)
        return fail(PMPM_EARG, "%sg2p_threads must be 128, 256 or 512: %lld", "", c.g2p_threads);
    if (c.cell_block != 4 && c.cell_block != 8)
        return fail(PMPM_EARG, "%scell_block must be 4 or 8: %lld", "", c.cell_block);
    if (c.p2g_split < 1 || c.p2g_split > 3)
        return fail(PMPM_EARG, "%sp2g_split must be 1, 2 or 3: %lld", "", c.p2g_split);
    if (c.const_stream != 1 && c.const_stream != 2)
        return fail(PMPM_EARG, "%sconst_stream must be 1 or 2: %lld", "", c.const_stream);
    if (c.const_stream == 2) {
        // the second stream of the current device, made here, outside any
        // stream capture a later step may run under
        SideStream* ss = nullptr;
        if (int rc = side_stream(&ss)) return rc;
    }
    g_launch = c;
    return PMPM_OK;
}

int pmpm_get_launch_config(pmpm_launch_config* cfg) {
    if (!cfg) return fail(PMPM_EARG, "%snull argument", "");
    *cfg = g_launch;
    return PMPM_OK;
}

int pmpm_damage_solve_info(const void* ws, int64_t max_n, int32_t n_grid, int32_t n_scenes, int32_t dtype,
                           int32_t* iterations, double* rel_residual, void* stream) {
    if (!ws) return fail(PMPM_EARG, "%snull workspace", "");
    if (int rc = ws_check(ws, 0, n_grid, n_scenes, dtype)) return rc;
    Layout L(max_n, n_grid, dtype, n_scenes);
    double h[PCG_NSCAL];
    HIPCHK(hipMemcpyAsync(h, (const char*)ws + L.pcgr + (size_t)3 * GRID_WG_COUNT * 8, sizeof h, hipMemcpyDeviceToHost,
                          (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    if (iterations) *iterations = (int32_t)h[PCG_ITER];
    if (rel_residual) *rel_residual = h[PCG_REL];
    return PMPM_OK;
}

int pmpm_substep_begin(const pmpm_step_params* p, const pmpm_particles_in* in, const pmpm_particles_out* out, void* ws,
                       size_t ws_bytes, void* stream) {
    int64_t n;
    int rc = check_step(p, in, out, ws, ws_bytes, n);
    if (rc) return rc;
    if (!ws) return fail(PMPM_EARG, "%snull workspace", "");
    return p->dtype == PMPM_F64 ? step_begin<double>(p, in, out, (char*)ws, (hipStream_t)stream, p->exact_sigma_eig)
                                : step_begin<float>(p, in, out, (char*)ws, (hipStream_t)stream, p->exact_sigma_eig);
}

int pmpm_substep_end(const pmpm_step_params* p, const pmpm_particles_in* in, const pmpm_particles_out* out, void* ws,
                     size_t ws_bytes, void* stream) {
    int64_t n;
    int rc = check_step(p, in, out, ws, ws_bytes, n);
    if (rc) return rc;
    if (!ws) return fail(PMPM_EARG, "%snull workspace", "");
    return p->dtype == PMPM_F64 ? step_end<double>(p, in, out, (char*)ws, (hipStream_t)stream)
                                : step_end<float>(p, in, out, (char*)ws, (hipStream_t)stream);
}

int pmpm_substep(const pmpm_step_params* p, const pmpm_particles_in* in, const pmpm_particles_out* out, void* ws,
                 size_t ws_bytes, void* stream) {
    int rc = pmpm_substep_begin(p, in, out, ws, ws_bytes, stream);
    if (rc) return rc;
    return pmpm_substep_end(p, in, out, ws, ws_bytes, stream);
}

int pmpm_carried_steps(const void* ws, int64_t* steps) {
    if (!ws || !steps) return fail(PMPM_EARG, "%snull argument", "");
    *steps = (int64_t)carry_taken(ws);
    return PMPM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- no-allocation step
// The scratch of pmpm_substep_noalloc: the step's workspace, in device
// memory the library itself holds (every GPU the library runs on holds its
// own copy until the process ends).  Host code never passes this name to a
// kernel or a copy: it stands in for the array on the host.  The device
// address comes from hipGetSymbolAddress, once per device, below.
static __device__ __align__(256) unsigned char scratch[PMPM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of the current device's scratch, looked up on the first
// call on that device (with the device current) and kept.
int scratch_address(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(PMPM_EARG, "%sdevice ordinal out of range: %lld", "", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(PMPM_EHIP, "%sno device address for the scratch array", "");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return PMPM_OK;
}
}  // namespace

extern "C" {

int pmpm_substep_noalloc(const pmpm_step_params* p, const pmpm_particles_in* in, const pmpm_particles_out* out,
                         void* stream) {
    if (!p || !in || !out) return fail(PMPM_EARG, "%snull argument", "");
    if (p->dtype != PMPM_F32 && p->dtype != PMPM_F64) return fail(PMPM_EARG, "%sunknown dtype %lld", "", p->dtype);
    if (p->n_grid < 4 || p->n_grid > 1024) return fail(PMPM_EARG, "%sn_grid out of range: %lld", "", p->n_grid);
    if (in->n < 0 || in->n > 0x7fffffffLL) return fail(PMPM_EARG, "%sparticle count out of range: %lld", "", in->n);
    if (p->keep_for_adjoint) return fail(PMPM_EARG, "%skeep_for_adjoint needs a caller-owned workspace", "");
    const Layout L(in->n, p->n_grid, p->dtype, p->n_scenes);
    if (L.total > (size_t)PMPM_NOALLOC_SCRATCH_BYTES)
        return fail(PMPM_ESIZE, "%sthis step needs %lld bytes of scratch; pmpm_substep_noalloc holds "
                    "PMPM_NOALLOC_SCRATCH_BYTES (use pmpm_substep with a workspace)", "", (long long)L.total);
    int64_t n;
    if (int rc = check_step(p, in, out, nullptr, L.total, n)) return rc;
    void* ws = nullptr;
    if (int rc = scratch_address(&ws)) return rc;
    hipStream_t st = (hipStream_t)stream;
    // the previous call may have laid the scratch out for another shape:
    // clear the zero-state sections with a fill kernel
    const long long nz = (long long)L.zero_end();
    hipLaunchKernelGGL(k_zero_bytes, dim3(nblocks((nz / 16) < 1 ? 1 : nz / 16)), dim3(WG), 0, st, (unsigned char*)ws, nz);
    HIPCHK(hipGetLastError());
    t_untimed = true;
    int rc = p->dtype == PMPM_F64 ? step_begin<double>(p, in, out, (char*)ws, st, p->exact_sigma_eig, false, false, false)
                                  : step_begin<float>(p, in, out, (char*)ws, st, p->exact_sigma_eig, false, false, false);
    if (!rc)
        rc = p->dtype == PMPM_F64 ? step_end<double>(p, in, out, (char*)ws, st, false)
                                  : step_end<float>(p, in, out, (char*)ws, st, false);
    t_untimed = false;
    return rc;
}

int pmpm_timing(int32_t enable) {
    g_timing.reset();
    g_timing.on = enable != 0;
    return PMPM_OK;
}

int pmpm_timing_read(double* total_ms, int64_t* launches, int32_t n) {
    for (int g = 0; g < NTIMED && g < n; ++g) {
        double tot = 0;
        auto& v = g_timing.ev[g];
        for (size_t i = 0; i + 1 < v.size(); i += 2) {
            HIPCHK(hipEventSynchronize(v[i + 1]));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, v[i], v[i + 1]));
            tot += ms;
        }
        if (total_ms) total_ms[g] = tot;
        if (launches) launches[g] = (int64_t)(v.size() / 2);
    }
    return PMPM_OK;
}

size_t pmpm_adjoint_workspace_bytes(int64_t max_n, int32_t n_grid, int32_t dtype) {
    if (n_grid < 1 || max_n < 0) return 0;
    const size_t ts = dtype == PMPM_F64 ? 8 : 4;
    return ((size_t)n_grid * n_grid * n_grid * 6 + (size_t)max_n * 10) * ts;
}

int pmpm_substep_adjoint(const pmpm_step_params* p, const pmpm_particles_in* in, const pmpm_particles_out* fwd_out,
                         const pmpm_adjoint_in* gin, const pmpm_adjoint_out* gout, void* ws, size_t ws_bytes, void* aws,
                         size_t aws_bytes, void* stream) {
    int64_t n;
    if (!ws || !aws || !gin || !gout) return fail(PMPM_EARG, "%snull argument", "");
    int rc = check_step(p, in, fwd_out, ws, ws_bytes, n);
    if (rc) return rc;
    if (aws_bytes < pmpm_adjoint_workspace_bytes(in->n, p->n_grid, p->dtype))
        return fail(PMPM_ESIZE, "%sadjoint workspace too small, need %lld bytes", "",
                    (long long)pmpm_adjoint_workspace_bytes(in->n, p->n_grid, p->dtype));
    if (p->n_scenes > 1 || p->damage_solver != PMPM_DAMAGE_EXPLICIT || p->export_active)
        return fail(PMPM_EARG, "%sthe adjoint supports one scene, device colliders and the explicit damage update", "");
    if (in->uniform_mask) return fail(PMPM_EARG, "%sthe adjoint takes per-particle material arrays (uniform_mask 0)", "");
    if (in->n > 0) {
        const void* need[] = {gin->x, gin->v, gin->F, gin->C, gin->d, gin->lap, gin->D_hist, gout->x, gout->v, gout->F,
                              gout->C, gout->d, gout->lap, gout->D_hist, gout->mu, gout->lam, gout->A0, gout->sigma_c, fwd_out->d,
                              fwd_out->D_hist, fwd_out->phi, fwd_out->phi_shear};
        for (const void* q : need)
            if (!q) return fail(PMPM_EARG, "%smissing adjoint array", "");
    }
    carry_take(ws, false, nullptr, 0, (hipStream_t)stream);
    return p->dtype == PMPM_F64
               ? step_adjoint<double>(p, in, fwd_out, gin, gout, (char*)ws, (char*)aws, (hipStream_t)stream)
               : step_adjoint<float>(p, in, fwd_out, gin, gout, (char*)ws, (char*)aws, (hipStream_t)stream);
}

int pmpm_bin(const void* x, int32_t dtype, int64_t n, int32_t n_grid, double inv_dx, int32_t stable, int32_t* key, int32_t* perm,
             int32_t* cell_start, void* ws, size_t ws_bytes, void* stream) {
    if (!ws || (n > 0 && (!x || !perm))) return fail(PMPM_EARG, "%snull argument", "");
    if (dtype != PMPM_F32 && dtype != PMPM_F64) return fail(PMPM_EARG, "%sunknown dtype %lld", "", dtype);
    if (n_grid < 4 || n_grid > 1024 || n < 0 || n > 0x7fffffffLL) return fail(PMPM_EARG, "%sbad size %lld", "", n);
    Layout L(n, n_grid, dtype);
    if (ws_bytes < L.total) return fail(PMPM_ESIZE, "%sworkspace too small, need %lld bytes", "", (long long)L.total);
    if (int rc = ws_check(ws, n, n_grid, 1, dtype)) return rc;
    hipStream_t st = (hipStream_t)stream;
    carry_take(ws, false, nullptr, 0, st);
    int rc = dtype == PMPM_F64 ? run_bin<double>((const double*)x, n, n_grid, inv_dx, stable, nullptr, 1, L, (char*)ws, st)
                               : run_bin<float>((const float*)x, n, n_grid, inv_dx, stable, nullptr, 1, L, (char*)ws, st);
    if (rc) return rc;
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(perm, (char*)ws + L.perm, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
        if (key) HIPCHK(hipMemcpyAsync(key, (char*)ws + L.key, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    }
    if (cell_start) {
        char* w = (char*)ws;
        hipLaunchKernelGGL(k_dense_starts, dim3(nblocks(L.NC)), dim3(WG), 0, st, (const int*)(w + L.bstart),
                           (const int*)(w + L.slotof), L.NC, (int*)(w + L.start));
        HIPCHK(hipMemcpyAsync(cell_start, (char*)ws + L.start, (size_t)(L.NC + 1) * 4, hipMemcpyDeviceToDevice, st));
    }
    return PMPM_OK;
}

int pmpm_stencil(const void* x, int32_t dtype, int64_t n, int32_t n_grid, double inv_dx, int32_t* base_q, int32_t* idx27,
                 int32_t* base_c, int32_t* idx64, void* w, void* wc, void* wc2, void* stream) {
    if (n < 0 || (n > 0 && !x)) return fail(PMPM_EARG, "%snull argument", "");
    if (n == 0) return PMPM_OK;
    hipStream_t st = (hipStream_t)stream;
    if (dtype == PMPM_F64)
        hipLaunchKernelGGL(k_stencil<double>, dim3(nblocks(n)), dim3(WG), 0, st, (const double*)x, (long long)n, n_grid, inv_dx,
                           base_q, idx27, base_c, idx64, (double*)w, (double*)wc, (double*)wc2);
    else
        hipLaunchKernelGGL(k_stencil<float>, dim3(nblocks(n)), dim3(WG), 0, st, (const float*)x, (long long)n, n_grid,
                           (float)inv_dx, base_q, idx27, base_c, idx64, (float*)w, (float*)wc, (float*)wc2);
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}

int pmpm_sym_eig(const void* A, int32_t dtype, int64_t n, void* evals, void* V, void* stream) {
    if (n < 0 || (n > 0 && (!A || !evals || !V))) return fail(PMPM_EARG, "%snull argument", "");
    if (n == 0) return PMPM_OK;
    hipStream_t st = (hipStream_t)stream;
    if (dtype == PMPM_F64)
        hipLaunchKernelGGL(k_sym_eig<double>, dim3(nblocks(n)), dim3(WG), 0, st, (const double*)A, (long long)n, (double*)evals,
                           (double*)V);
    else
        hipLaunchKernelGGL(k_sym_eig<float>, dim3(nblocks(n)), dim3(WG), 0, st, (const float*)A, (long long)n, (float*)evals,
                           (float*)V);
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}

int pmpm_gaussian_carry(int32_t dtype, int64_t n_gauss, int64_t n_particles, const int64_t* bind, const void* x,
                        const void* F, const void* d, const void* off0, const void* cov0, const void* color,
                        double grid_lim, double scale_to_sim, const double* center, double crack_tint, void* pos,
                        void* cov, void* color_out, void* stream) {
    if (dtype != PMPM_F32 && dtype != PMPM_F64) return fail(PMPM_EARG, "%sunknown dtype %lld", "", dtype);
    if (n_gauss < 0 || n_particles < 0) return fail(PMPM_EARG, "%sbad size %lld", "", n_gauss < 0 ? n_gauss : n_particles);
    if (n_gauss == 0) return PMPM_OK;
    if (!bind || !x || !F || !off0 || !cov0 || !center || !pos || !cov || (color_out && !color) ||
        (color_out && crack_tint > 0.0 && !d))
        return fail(PMPM_EARG, "%snull argument", "");
    if (n_particles == 0) return fail(PMPM_EARG, "%ssplats bound to an empty particle set", "");
    if (!(grid_lim > 0.0) || !(scale_to_sim > 0.0))
        return fail(PMPM_EARG, "%sgrid_lim and scale_to_sim must be positive", "");
    hipStream_t st = (hipStream_t)stream;
    return dtype == PMPM_F64
               ? launch_gaussian_carry<double>(n_gauss, (const long long*)bind, x, F, d, off0, cov0, color, grid_lim,
                                               scale_to_sim, center, crack_tint, pos, cov, color_out, st)
               : launch_gaussian_carry<float>(n_gauss, (const long long*)bind, x, F, d, off0, cov0, color, grid_lim,
                                              scale_to_sim, center, crack_tint, pos, cov, color_out, st);
}

size_t pmpm_observe_workspace_bytes(void) { return (size_t)OBS_WG_COUNT * OBS_NSLOT * sizeof(double); }

int pmpm_observe_frame(int32_t dtype, int64_t n, int32_t frame, const pmpm_frame_in* in, const pmpm_frame_obs* obs, void* ws,
                       size_t ws_bytes, void* stream) {
    if (dtype != PMPM_F32 && dtype != PMPM_F64) return fail(PMPM_EARG, "%sunknown dtype %lld", "", dtype);
    if (!in || !obs || !ws) return fail(PMPM_EARG, "%snull argument", "");
    if (n <= 0) return fail(PMPM_EARG, "%sno particles: %lld", "", n);
    if (frame < 0) return fail(PMPM_EARG, "%snegative frame %lld", "", frame);
    if (!in->d || !in->x || !in->region || !in->cap || !obs->peaks || !obs->onset || !obs->log || !obs->d_e ||
        !obs->x_e || !obs->phipp_e || !obs->pshpp_e || !obs->phi_pp || !obs->psh_pp)
        return fail(PMPM_EARG, "%snull argument", "");
    if ((obs->tau_e && !in->tau_last) || (obs->J_e && !in->J_last)) return fail(PMPM_EARG, "%snull argument", "");
    if (ws_bytes < pmpm_observe_workspace_bytes())
        return fail(PMPM_ESIZE, "%sobserve workspace too small, need %lld bytes", "", (long long)pmpm_observe_workspace_bytes());
    hipStream_t st = (hipStream_t)stream;
    return dtype == PMPM_F64 ? launch_observe<double>(n, frame, in, obs, (double*)ws, st)
                             : launch_observe<float>(n, frame, in, obs, (double*)ws, st);
}

int pmpm_directional_pk1(int32_t dtype, int64_t n, const void* F, const void* n0, const void* d, const void* mu_arr,
                         double mu, const void* lam_arr, double lam, const void* gamma_I_arr, double gamma_I,
                         const void* gamma_II_arr, double gamma_II, void* P, void* psi, void* stream) {
    if (dtype != PMPM_F32 && dtype != PMPM_F64) return fail(PMPM_EARG, "%sunknown dtype %lld", "", dtype);
    if (n < 0) return fail(PMPM_EARG, "%sbad size %lld", "", n);
    if (n == 0) return PMPM_OK;
    if (!F || !n0 || !d || (!P && !psi)) return fail(PMPM_EARG, "%snull argument", "");
    hipStream_t st = (hipStream_t)stream;
    return dtype == PMPM_F64 ? launch_option_a<double>(n, F, n0, d, mu_arr, mu, lam_arr, lam, gamma_I_arr, gamma_I,
                                                       gamma_II_arr, gamma_II, P, psi, st)
                             : launch_option_a<float>(n, F, n0, d, mu_arr, mu, lam_arr, lam, gamma_I_arr, gamma_I,
                                                      gamma_II_arr, gamma_II, P, psi, st);
}

int pmpm_directional_kirchhoff(int32_t dtype, int64_t n, const void* tau_c, const void* F, const void* d, const void* n0,
                               const void* res_arr, double res, int32_t tension_gate, void* tau, void* stream) {
    if (dtype != PMPM_F32 && dtype != PMPM_F64) return fail(PMPM_EARG, "%sunknown dtype %lld", "", dtype);
    if (n < 0) return fail(PMPM_EARG, "%sbad size %lld", "", n);
    if (n == 0) return PMPM_OK;
    if (!tau_c || !F || !d || !n0 || !tau) return fail(PMPM_EARG, "%snull argument", "");
    hipStream_t st = (hipStream_t)stream;
    const dim3 G((unsigned)((n + WG - 1) / WG)), B(WG);
    if (dtype == PMPM_F64)
        hipLaunchKernelGGL(k_dir_kirchhoff<double>, G, B, 0, st, n, (const double*)tau_c, (const double*)F, (const double*)d,
                           (const double*)n0, (const double*)res_arr, res, tension_gate, (double*)tau);
    else
        hipLaunchKernelGGL(k_dir_kirchhoff<float>, G, B, 0, st, n, (const float*)tau_c, (const float*)F, (const float*)d,
                           (const float*)n0, (const float*)res_arr, (float)res, tension_gate, (float*)tau);
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}

size_t pmpm_compact_workspace_bytes(int64_t n_candidates) { return n_candidates < 0 ? 0 : sc_bytes(n_candidates); }

int pmpm_ball_particles(const double* g, int64_t n, double R, double js, const double* center, const pmpm_pcg64* rng,
                        float* out, int64_t cap, int64_t* count, void* ws, size_t ws_bytes, void* stream) {
    if (!g || !center || !rng || !out || !count || !ws) return fail(PMPM_EARG, "%snull argument", "");
    // the compaction's offsets are 32-bit: fewer than 2^31 candidates
    if (n < 1 || n > 1290) return fail(PMPM_EARG, "%slattice size out of range: %lld", "", n);
    if (ws_bytes < sc_bytes(n * n * n))
        return fail(PMPM_ESIZE, "%scompaction workspace too small, need %lld bytes", "", (long long)sc_bytes(n * n * n));
    return launch_ball(g, n, R, js, center, *rng, out, cap, (long long*)count, (char*)ws, (hipStream_t)stream);
}

int pmpm_grid_block(const double* xs, const double* ys, const double* zs, int64_t nx, int64_t ny, int64_t nz, double* out,
                    void* stream) {
    if (!xs || !ys || !zs || !out) return fail(PMPM_EARG, "%snull argument", "");
    if (nx < 0 || ny < 0 || nz < 0 || (nx > 0 && ny > 0 && nz > (1LL << 40) / nx / ny))
        return fail(PMPM_EARG, "%sbad block size", "");
    const long long N = nx * ny * nz;
    if (N == 0) return PMPM_OK;
    hipLaunchKernelGGL(k_grid_block, dim3((unsigned)((N + WG - 1) / WG)), dim3(WG), 0, (hipStream_t)stream, xs, ys, zs, nx,
                       ny, nz, out);
    HIPCHK(hipGetLastError());
    return PMPM_OK;
}

size_t pmpm_fill_workspace_bytes(const int64_t* dims) {
    const size_t bounds = al(3 * FILL_PARTS * sizeof(double)) + al(3 * FILL_PARTS * sizeof(long long));
    if (!dims) return bounds;
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1) return 0;
    return std::max(bounds, fill_bytes(dims[0] * dims[1] * dims[2]));
}

static int fill_check(int32_t w_dtype, int32_t ct_dtype, const void* w, int64_t n, const double* center, double scale,
                      double dx) {
    if (w_dtype != PMPM_F32 && w_dtype != PMPM_F64) return fail(PMPM_EARG, "%sunknown point dtype %lld", "", w_dtype);
    if (ct_dtype != PMPM_F32 && ct_dtype != PMPM_F64) return fail(PMPM_EARG, "%sunknown compute dtype %lld", "", ct_dtype);
    if (w_dtype == PMPM_F64 && ct_dtype == PMPM_F32) return fail(PMPM_EARG, "%sfp64 points need fp64 arithmetic", "");
    if (!w || !center) return fail(PMPM_EARG, "%snull argument", "");
    if (n < 1) return fail(PMPM_EARG, "%sno points: %lld", "", n);
    if (!(dx > 0.0) || !(scale == scale)) return fail(PMPM_EARG, "%svoxel size must be positive", "");
    return PMPM_OK;
}

int pmpm_fill_bounds(int32_t w_dtype, int32_t ct_dtype, const void* w, int64_t n, const double* center, double scale,
                     double dx, double* lo, int64_t* dims, void* ws, size_t ws_bytes, void* stream) {
    int rc = fill_check(w_dtype, ct_dtype, w, n, center, scale, dx);
    if (rc) return rc;
    if (!lo || !dims || !ws) return fail(PMPM_EARG, "%snull argument", "");
    if (ws_bytes < pmpm_fill_workspace_bytes(nullptr))
        return fail(PMPM_ESIZE, "%sfill workspace too small, need %lld bytes", "", (long long)pmpm_fill_workspace_bytes(nullptr));
    hipStream_t st = (hipStream_t)stream;
    long long* D = (long long*)dims;
    if (w_dtype == PMPM_F64) return launch_fill_bounds<double, double>((const double*)w, n, center, scale, dx, lo, D, (char*)ws, st);
    if (ct_dtype == PMPM_F64) return launch_fill_bounds<float, double>((const float*)w, n, center, scale, dx, lo, D, (char*)ws, st);
    return launch_fill_bounds<float, float>((const float*)w, n, center, scale, dx, lo, D, (char*)ws, st);
}

int pmpm_fill_interior(int32_t w_dtype, int32_t ct_dtype, const void* w, int64_t n, const double* center, double scale,
                       double dx, const double* lo, const int64_t* dims, double js, const pmpm_pcg64* rng, float* out,
                       int64_t cap, int64_t* count, void* ws, size_t ws_bytes, void* stream) {
    int rc = fill_check(w_dtype, ct_dtype, w, n, center, scale, dx);
    if (rc) return rc;
    if (!lo || !dims || !out || !count || !ws) return fail(PMPM_EARG, "%snull argument", "");
    // the compaction's offsets are 32-bit: fewer than 2^31 voxels
    if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || dims[0] > (1LL << 20) || dims[1] > (1LL << 20) || dims[2] > (1LL << 20) ||
        dims[0] * dims[1] >= (1LL << 31) / dims[2])
        return fail(PMPM_EARG, "%svoxel grid out of range", "");
    if (ws_bytes < pmpm_fill_workspace_bytes(dims))
        return fail(PMPM_ESIZE, "%sfill workspace too small, need %lld bytes", "", (long long)pmpm_fill_workspace_bytes(dims));
    hipStream_t st = (hipStream_t)stream;
    const long long* D = (const long long*)dims;
    long long* cnt = (long long*)count;
    if (w_dtype == PMPM_F64)
        return launch_fill_interior<double, double>((const double*)w, n, center, scale, dx, lo, D, js, rng, out, cap, cnt,
                                                    (char*)ws, st);
    if (ct_dtype == PMPM_F64)
        return launch_fill_interior<float, double>((const float*)w, n, center, scale, dx, lo, D, js, rng, out, cap, cnt,
                                                   (char*)ws, st);
    return launch_fill_interior<float, float>((const float*)w, n, center, scale, dx, lo, D, js, rng, out, cap, cnt, (char*)ws,
                                              st);
}

}  // extern "C"
