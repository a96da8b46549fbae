// One very large world on the whole device: the same step as the per-world
// kernel (Step(AabbTreeWorld&, const StepConf&), AabbTreeWorld.cpp:2144-2219,
// doToi == false) but every phase spread over the GPU, with two changes of
// method:
//
// * Broadphase: a linear BVH rebuilt from the proxies' fat AABBs whenever
//   pairs are looked for (Morton codes of the AABB centres, a workgroup-wide
//   LSD radix sort, Karras' radix-tree construction, bottom-up bounds refit,
//   escape indices for a stackless traversal).  It answers the same queries as
//   the reference's DynamicTree (FindContacts, AabbTreeWorld.cpp:709-751):
//   for every moved proxy, all proxies of other bodies whose fat AABBs
//   overlap it.  The pair sets are identical; new contacts are appended in
//   proxy order instead of key order.
// * Solver: graph-coloured Gauss-Seidel.  Touching contacts are coloured so no
//   two contacts of a colour share a dynamic body; each colour is solved in
//   parallel, colours in sequence, with the reference's per-constraint
//   arithmetic (warm start, SolveVelocityConstraint with the block solver,
//   SolvePositionConstraint).  All awake dynamic bodies form one system (no
//   island split, no sleeping).  Results are not the reference's bits (the
//   sweep order differs); they are judged by penetration, momentum and
//   stacking behaviour.
//
// Order-sensitive bookkeeping (stable contact compaction, scans, the sort and
// the colouring) runs in single-workgroup kernels so every run is
// deterministic.
#include <mutex>

#include "../../include/playrho_amd.h"
#include "big_world.hpp"
#include "step_device.hpp"

namespace prxd {

using prx::BigState;

constexpr int kBlock = prx::kBigBlock;  // threads of the single-workgroup kernels
constexpr int kGrid = 256;              // threads per workgroup of the grid-wide kernels
constexpr int kSweepBlocks = 64;        // workgroups of a colour sweep launch (16,384 threads)
constexpr std::uint64_t kEmptyKey = ~0ull;
constexpr int kBigToiPasses = 2;  // time-of-impact passes per step against static shapes
// Bound on tree depth for the parent walks (a Karras tree over 30-bit codes
// with index tie-breaks is at most 30 + 32 deep).
constexpr int kMaxDepth = 96;

__device__ __forceinline__ WorldCtx big_ctx(const DeviceState& d) {
    const WorldHeader& H = d.hdr[0];
    WorldCtx x;
    x.w = 0;
    x.bo = 0;
    x.nb = H.nbodies;
    x.po = 0;
    x.np = H.nproxies;
    x.co = 0;
    x.cap = H.contact_cap;
    x.jo = 0;
    x.nj = 0;
    x.eo = 0;
    x.ecap = 0;
    x.lane = int(threadIdx.x) & (kWave - 1);
    return x;
}

// Exclusive prefix sum over the workgroup (kBlock threads); lds holds
// kBlock / 64 + 1 ints.  Returns the thread's prefix; total receives the sum.
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int& total) {
    const int lane = int(threadIdx.x) & (kWave - 1), wave = int(threadIdx.x) >> 6;
    const int incl = wave_incl_scan(v, lane);
    if (lane == kWave - 1) lds[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < kBlock / kWave; ++w) {
            const int t = lds[w];
            lds[w] = run;
            run += t;
        }
        lds[kBlock / kWave] = run;
    }
    __syncthreads();
    const int r = lds[wave] + incl - v;
    total = lds[kBlock / kWave];
    __syncthreads();
    return r;
}

__device__ __forceinline__ float block_reduce_max(float v, float* lds) {
    for (int o = 32; o > 0; o >>= 1) v = fmax_std(v, __shfl_xor(v, o));
    const int lane = int(threadIdx.x) & (kWave - 1), wave = int(threadIdx.x) >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    float r = lds[0];
    for (int w = 1; w < kBlock / kWave; ++w) r = fmax_std(r, lds[w]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ float block_reduce_min(float v, float* lds) {
    for (int o = 32; o > 0; o >>= 1) v = fmin_std(v, __shfl_xor(v, o));
    const int lane = int(threadIdx.x) & (kWave - 1), wave = int(threadIdx.x) >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    float r = lds[0];
    for (int w = 1; w < kBlock / kWave; ++w) r = fmin_std(r, lds[w]);
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Pre-phase.

// CreateProxies (AabbTreeWorld.cpp:557-579); also clears the step counters.
__global__ void k_big_pending(DeviceState d, StepParams c) {
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p == 0) {
        for (int i = 0; i < prx::kStats; ++i) d.hdr[0].stats[i] = 0u;
        d.hdr[0].stats[prx::kStRegMinSeparation] = 0x7f800000u;
        d.hdr[0].stats[prx::kStToiMinSeparation] = 0x7f800000u;
    }
    if (p >= d.hdr[0].nproxies) return;
    if (d.pflags[p] & prx::kProxyPending) {
        const int gb = d.pbody[p];
        d.paabb[p] = proxy_aabb(shape_proxy(d, d.pshape[p]), body_xf0(d, gb), body_xf1(d, gb), c, true);
        d.pflags[p] = prx::kProxyMoved;
    }
}

// DestroyContacts (AabbTreeWorld.cpp:2290-2331): contacts whose fat AABBs no
// longer overlap are removed by a stable in-place compaction, tile by tile
// (every record of a tile is read before any of the tile is written, and
// records only move down).
__global__ __launch_bounds__(kBlock) void k_big_destroy(DeviceState d) {
    __shared__ int s_scan[kBlock / kWave + 1];
    WorldHeader& H = d.hdr[0];
    const int nc = H.ncontacts;
    int out = 0;
    std::uint32_t destroyed = 0;
    for (int base = 0; base < nc; base += kBlock) {
        const int i = base + int(threadIdx.x);
        const bool valid = i < nc;
        bool keep = false;
        if (valid) {
            const int2 pr = d.c_prox[i];
            keep = overlap(d.paabb[pr.x], d.paabb[pr.y]);
            if (!keep) {
                const std::uint32_t mh = d.m_hdr[i].x;
                if (((mh >> 8) & 0xffu) > 0 && !(d.c_flags[i] & prx::kSensor)) {
                    const int2 bd = d.c_body[i];
                    body_set_awake(d, bd.x);
                    body_set_awake(d, bd.y);
                }
            }
        }
        int tot = 0;
        const int dst = out + block_exclusive_scan((valid && keep) ? 1 : 0, s_scan, tot);
        const bool move = valid && keep && dst != i;
        int2 pr{0, 0}, bd{0, 0}, sh{0, 0};
        std::uint32_t id = 0, fl = 0;
        float4 mt{0, 0, 0, 0}, mg{0, 0, 0, 0}, p0{0, 0, 0, 0}, p1{0, 0, 0, 0};
        uint4 mh{0, 0, 0, 0};
        if (move) {
            pr = d.c_prox[i];
            bd = d.c_body[i];
            sh = d.c_shape[i];
            id = d.c_id[i];
            fl = d.c_flags[i];
            mt = d.c_mat[i];
            mh = d.m_hdr[i];
            mg = d.m_geo[i];
            p0 = d.m_p0[i];
            p1 = d.m_p1[i];
        }
        __syncthreads();
        if (move) {
            d.c_prox[dst] = pr;
            d.c_body[dst] = bd;
            d.c_shape[dst] = sh;
            d.c_id[dst] = id;
            d.c_flags[dst] = fl;
            d.c_mat[dst] = mt;
            d.m_hdr[dst] = mh;
            d.m_geo[dst] = mg;
            d.m_p0[dst] = p0;
            d.m_p1[dst] = p1;
        }
        destroyed += std::uint32_t(min(kBlock, nc - base) - tot);
        out += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        H.ncontacts = out;
        H.stats[prx::kStPreDestroyed] += destroyed;
    }
}

// UpdateContacts (AabbTreeWorld.cpp:2333-2421): narrowphase of flagged contacts.
__global__ void k_big_update(DeviceState d, StepParams c, int statSlot) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= d.hdr[0].ncontacts) return;
    if (d.c_flags[i] & prx::kNeedsUpdating) {
        update_contact(d, 0, i, c);
        if (statSlot >= 0) atomicAdd(&d.hdr[0].stats[statSlot], 1u);
    }
}

// ---------------------------------------------------------------------------
// Linear BVH over the proxies' fat AABBs.

// Bounds of the AABB centres (one workgroup).
__global__ __launch_bounds__(kBlock) void k_lbvh_bounds(DeviceState d, BigState g) {
    __shared__ float s_red[kBlock / kWave];
    const int np = d.hdr[0].nproxies;
    float lox = __builtin_inff(), loy = __builtin_inff(), hix = -__builtin_inff(), hiy = -__builtin_inff();
    for (int p = int(threadIdx.x); p < np; p += kBlock) {
        const float4 a = d.paabb[p];
        const float cx = (a.x + a.z) * 0.5f, cy = (a.y + a.w) * 0.5f;
        lox = fmin_std(lox, cx);
        loy = fmin_std(loy, cy);
        hix = fmax_std(hix, cx);
        hiy = fmax_std(hiy, cy);
    }
    lox = block_reduce_min(lox, s_red);
    loy = block_reduce_min(loy, s_red);
    hix = block_reduce_max(hix, s_red);
    hiy = block_reduce_max(hiy, s_red);
    if (threadIdx.x == 0) g.bounds[0] = float4{lox, loy, hix, hiy};
}

__device__ __forceinline__ std::uint32_t spread15(std::uint32_t v) {
    v &= 0x7fffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// 30-bit Morton code of each proxy's AABB centre.
__global__ void k_lbvh_morton(DeviceState d, BigState g) {
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= d.hdr[0].nproxies) return;
    const float4 b = g.bounds[0];
    const float4 a = d.paabb[p];
    const float cx = (a.x + a.z) * 0.5f, cy = (a.y + a.w) * 0.5f;
    const float sx = (b.z > b.x) ? 32767.0f / (b.z - b.x) : 0.0f;
    const float sy = (b.w > b.y) ? 32767.0f / (b.w - b.y) : 0.0f;
    const float fx = fmin_std(fmax_std((cx - b.x) * sx, 0.0f), 32767.0f);
    const float fy = fmin_std(fmax_std((cy - b.y) * sy, 0.0f), 32767.0f);
    g.morton[p] = (spread15(std::uint32_t(fx)) << 1) | spread15(std::uint32_t(fy));
    g.order[p] = p;
}

// Stable LSD radix sort of (key, value) pairs by the low `bits` bits of the
// key, one workgroup, 8-bit digits, ping-ponging with the second buffers;
// an even number of passes leaves the result in the first buffers.  Ranks
// within a tile come from per-wave digit matching (8 ballots) plus per-wave
// digit counts in LDS.
__device__ __forceinline__ void block_radix_sort(std::uint32_t* k0, int* v0, std::uint32_t* k1, int* v1, int n,
                                                 int passes) {
    __shared__ int s_base[256];
    __shared__ int s_tile[256];
    __shared__ int s_wave[kBlock / kWave][256];
    const int lane = int(threadIdx.x) & (kWave - 1), wave = int(threadIdx.x) >> 6;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = 8 * pass;
        std::uint32_t* ks = (pass & 1) ? k1 : k0;
        int* vs = (pass & 1) ? v1 : v0;
        std::uint32_t* kd = (pass & 1) ? k0 : k1;
        int* vd = (pass & 1) ? v0 : v1;
        for (int t = int(threadIdx.x); t < 256; t += kBlock) s_base[t] = 0;
        for (int t = int(threadIdx.x); t < 256 * (kBlock / kWave); t += kBlock) (&s_wave[0][0])[t] = 0;
        __syncthreads();
        for (int i = int(threadIdx.x); i < n; i += kBlock) atomicAdd(&s_base[(ks[i] >> shift) & 0xffu], 1);
        __syncthreads();
        if (threadIdx.x == 0) {
            int run = 0;
            for (int t = 0; t < 256; ++t) {
                const int c = s_base[t];
                s_base[t] = run;
                run += c;
            }
        }
        __syncthreads();
        for (int base = 0; base < n; base += kBlock) {
            const int i = base + int(threadIdx.x);
            const bool valid = i < n;
            std::uint32_t key = 0;
            int val = 0;
            int dig = 0;
            if (valid) {
                key = ks[i];
                val = vs[i];
                dig = int((key >> shift) & 0xffu);
            }
            unsigned long long peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const bool bit = (dig >> b) & 1;
                const unsigned long long m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            const int rank = __popcll(peers & lanes_below(lane));
            if (valid && rank == 0) s_wave[wave][dig] = __popcll(peers);
            __syncthreads();
            for (int t = int(threadIdx.x); t < 256; t += kBlock) {
                int run = 0;
                for (int w = 0; w < kBlock / kWave; ++w) {
                    const int c = s_wave[w][t];
                    s_wave[w][t] = run;
                    run += c;
                }
                s_tile[t] = run;
            }
            __syncthreads();
            if (valid) {
                const int dst = s_base[dig] + s_wave[wave][dig] + rank;
                kd[dst] = key;
                vd[dst] = val;
            }
            __syncthreads();
            for (int t = int(threadIdx.x); t < 256; t += kBlock) {
                s_base[t] += s_tile[t];
                for (int w = 0; w < kBlock / kWave; ++w) s_wave[w][t] = 0;
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_lbvh_sort(DeviceState d, BigState g) {
    block_radix_sort(g.morton, g.order, g.morton2, g.order2, d.hdr[0].nproxies, 4);
}

__device__ __forceinline__ int lbvh_delta(const std::uint32_t* code, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const std::uint32_t a = code[i], b = code[j];
    if (a == b) return 32 + __clz(std::uint32_t(i) ^ std::uint32_t(j));
    return __clz(a ^ b);
}

// Karras' radix-tree construction: internal node i (0 is the root) covers a
// key range with one end at i; children are internal nodes or leaves.  Node
// ids: internal 0..n-2, leaf k at n-1+k.
__global__ void k_lbvh_build(DeviceState d, BigState g) {
    const int n = d.hdr[0].nproxies;
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i == 0) g.parent[0] = -1;
    if (i >= n - 1) return;
    const std::uint32_t* code = g.morton;
    const int dr = lbvh_delta(code, n, i, i + 1), dl = lbvh_delta(code, n, i, i - 1);
    const int dir = (dr - dl) > 0 ? 1 : -1;
    const int dmin = lbvh_delta(code, n, i, i - dir);
    int lmax = 2;
    while (lbvh_delta(code, n, i, i + lmax * dir) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1) {
        if (lbvh_delta(code, n, i, i + (l + t) * dir) > dmin) l += t;
    }
    const int j = i + l * dir;
    const int dnode = lbvh_delta(code, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (lbvh_delta(code, n, i, i + (s + t) * dir) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * dir + (dir < 0 ? -1 : 0);
    const int lo = min(i, j), hi = max(i, j);
    const int left = (lo == gamma) ? (n - 1 + gamma) : gamma;
    const int right = (hi == gamma + 1) ? (n - 1 + gamma + 1) : (gamma + 1);
    g.left[i] = left;
    g.right[i] = right;
    g.parent[left] = i;
    g.parent[right] = i;
    g.visit[i] = 0;
}

// Bottom-up bounds: each leaf climbs; the second child to arrive at a node
// writes the node's union and continues.
__global__ void k_lbvh_refit(DeviceState d, BigState g) {
    const int n = d.hdr[0].nproxies;
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n) return;
    const int leaf = n - 1 + k;
    g.naabb[leaf] = d.paabb[g.order[k]];
    if (n == 1) return;
    __threadfence();
    int cur = g.parent[leaf];
    for (int depth = 0; cur >= 0; ++depth) {
        if (depth > kMaxDepth) {
            d.hdr[0].status |= prx::kStatusScratchOverflow;
            return;
        }
        if (atomicAdd(&g.visit[cur], 1) == 0) return;
        __threadfence();
        const volatile float4* nb = g.naabb;
        const int l = g.left[cur], r = g.right[cur];
        const float4 a = float4{nb[l].x, nb[l].y, nb[l].z, nb[l].w};
        const float4 b = float4{nb[r].x, nb[r].y, nb[r].z, nb[r].w};
        g.naabb[cur] = float4{fmin_std(a.x, b.x), fmin_std(a.y, b.y), fmax_std(a.z, b.z), fmax_std(a.w, b.w)};
        __threadfence();
        cur = g.parent[cur];
    }
}

// Escape index of every node: the next node of a depth-first pre-order walk
// once the node's subtree is done (-1 past the end).
__global__ void k_lbvh_escape(DeviceState d, BigState g) {
    const int n = d.hdr[0].nproxies;
    const int node = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= 2 * n - 1) return;
    int cur = node, esc = -1;
    for (int depth = 0; cur != 0 && depth <= kMaxDepth; ++depth) {
        const int p = g.parent[cur];
        if (g.left[p] == cur) {
            esc = g.right[p];
            break;
        }
        cur = p;
    }
    g.escape[node] = esc;
}

// FindContacts for one moved proxy p (AabbTreeWorld.cpp:709-751): proxies of
// other bodies whose fat AABB overlaps p's.  A pair of two moved proxies is
// emitted by the lower index only, so every key appears once.  Pass 0
// counts, pass 1 writes at the proxy's offset.
template <int kPass>
__global__ void k_lbvh_pairs(DeviceState d, BigState g) {
    const int n = d.hdr[0].nproxies;
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= n) return;
    int count = 0;
    if (d.pflags[p] & prx::kProxyMoved) {
        const float4 q = d.paabb[p];
        const int bodyP = d.pbody[p];
        int at = 0;
        if constexpr (kPass == 1) at = g.pcount[p];
        const int end = (kPass == 1) ? min(g.pcount[p + 1], g.pairCap) : 0;
        int node = 0;
        for (int visits = 0; node >= 0 && visits < 4 * n; ++visits) {
            if (overlap(g.naabb[node], q)) {
                if (node >= n - 1) {
                    const int r = g.order[node - (n - 1)];
                    if (r != p && d.pbody[r] != bodyP && (!(d.pflags[r] & prx::kProxyMoved) || p < r)) {
                        if constexpr (kPass == 0) {
                            ++count;
                        } else {
                            if (at < end) g.pairs[at] = int2{min(p, r), max(p, r)};
                            ++at;
                        }
                    }
                    node = g.escape[node];
                } else {
                    node = g.left[node];
                }
            } else {
                node = g.escape[node];
            }
        }
    }
    if constexpr (kPass == 0) g.pcount[p] = count;
}

// Exclusive scan of the per-proxy pair counts (one workgroup); pcount[n] is
// the total (clamped to the pair capacity, with the overflow status set).
__global__ __launch_bounds__(kBlock) void k_lbvh_scan(DeviceState d, BigState g) {
    __shared__ int s_scan[kBlock / kWave + 1];
    const int n = d.hdr[0].nproxies;
    int carry = 0;
    for (int base = 0; base < n; base += kBlock) {
        const int i = base + int(threadIdx.x);
        const int v = (i < n) ? g.pcount[i] : 0;
        int tot = 0;
        const int r = block_exclusive_scan(v, s_scan, tot);
        if (i < n) g.pcount[i] = carry + r;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        g.pcount[n] = carry;
        g.counters[prx::kBigPairs] = min(carry, g.pairCap);
        if (carry > g.pairCap) d.hdr[0].status |= prx::kStatusScratchOverflow;
    }
}

__device__ __forceinline__ std::uint32_t hash_slot(std::uint64_t key, int mask) {
    const std::uint64_t h = key * 0x9e3779b97f4a7c15ull;
    return std::uint32_t(h >> 32) & std::uint32_t(mask);
}

__global__ void k_hash_clear(BigState g) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i <= g.hashMask) g.hash[i] = kEmptyKey;
}

// Keys of the existing contacts into an open-addressing set.
__global__ void k_hash_insert(DeviceState d, BigState g) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= d.hdr[0].ncontacts) return;
    const int2 pr = d.c_prox[i];
    const std::uint64_t key = (std::uint64_t(std::uint32_t(pr.x)) << 32) | std::uint32_t(pr.y);
    std::uint32_t h = hash_slot(key, g.hashMask);
    for (int probe = 0; probe <= g.hashMask; ++probe) {
        const unsigned long long prev =
            atomicCAS(reinterpret_cast<unsigned long long*>(g.hash + h), kEmptyKey, (unsigned long long)key);
        if (prev == kEmptyKey || prev == key) return;
        h = (h + 1u) & std::uint32_t(g.hashMask);
    }
}

// AddContacts filters (AabbTreeWorld.cpp:2424-2470): existing key, neither
// body accelerable, shape filters.
__global__ void k_pairs_filter(DeviceState d, BigState g) {
    const int j = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= g.counters[prx::kBigPairs]) return;
    const int2 pq = g.pairs[j];
    const std::uint64_t key = (std::uint64_t(std::uint32_t(pq.x)) << 32) | std::uint32_t(pq.y);
    std::uint32_t h = hash_slot(key, g.hashMask);
    bool exists = false;
    for (int probe = 0; probe <= g.hashMask; ++probe) {
        const std::uint64_t v = g.hash[h];
        if (v == key) {
            exists = true;
            break;
        }
        if (v == kEmptyKey) break;
        h = (h + 1u) & std::uint32_t(g.hashMask);
    }
    bool acc = false;
    if (!exists) {
        const int ba = d.pbody[pq.x], bb = d.pbody[pq.y];
        const std::uint32_t fa = d.bflags[ba], fb = d.bflags[bb];
        acc = ((fa | fb) & prx::kAccelerable) != 0 && should_collide(d, d.pshape[pq.x], d.pshape[pq.y]);
    }
    g.paccept[j] = acc ? 1 : 0;
}

// Contact creation (AabbTreeWorld.cpp:2470-2547) for the accepted pairs, in
// pair order, appended after the existing contacts; bodies woken.
__global__ __launch_bounds__(kBlock) void k_pairs_append(DeviceState d, BigState g, int statSlot) {
    __shared__ int s_scan[kBlock / kWave + 1];
    WorldHeader& H = d.hdr[0];
    const int total = g.counters[prx::kBigPairs];
    const int nc0 = H.ncontacts;
    int added = 0;
    for (int base = 0; base < total; base += kBlock) {
        const int j = base + int(threadIdx.x);
        const bool acc = j < total && g.paccept[j] != 0;
        int tot = 0;
        const int r = block_exclusive_scan(acc ? 1 : 0, s_scan, tot);
        const int idx = nc0 + added + r;
        if (acc && idx < H.contact_cap) {
            const int2 pq = g.pairs[j];
            const int pa = pq.x, pb = pq.y;
            const int ba = d.pbody[pa], bb = d.pbody[pb];
            const int sa = d.pshape[pa], sb = d.pshape[pb];
            const std::uint32_t fa = d.bflags[ba], fb = d.bflags[bb];
            const int4 sia = d.shape_i[sa], sib = d.shape_i[sb];
            const float4 sfa = d.shape_f[sa], sfb = d.shape_f[sb];
            std::uint32_t flags = prx::kCEnabled | prx::kNeedsUpdating;
            if ((fa | fb) & prx::kImpenetrable) flags |= prx::kCImpenetrable;
            const bool sensor = sia.z != 0 || sib.z != 0;
            if (sensor) flags |= prx::kSensor;
            d.c_prox[idx] = int2{pa, pb};
            d.c_body[idx] = int2{ba, bb};
            d.c_shape[idx] = int2{sa, sb};
            d.c_id[idx] = std::uint32_t(idx);
            d.c_flags[idx] = flags;
            d.c_mat[idx] = float4{__builtin_sqrtf(sfa.y * sfb.y), (sfa.z > sfb.z) ? sfa.z : sfb.z, 0.0f, 0.0f};
            d.m_hdr[idx] = uint4{0u, kDefaultCf, kDefaultCf, 0u};
            d.m_geo[idx] = float4{0, 0, 0, 0};
            d.m_p0[idx] = float4{0, 0, 0, 0};
            d.m_p1[idx] = float4{0, 0, 0, 0};
            if (!sensor) {
                if (fa & prx::kSpeedable) atomicOr(&d.bflags[ba], prx::kAwake);
                if (fb & prx::kSpeedable) atomicOr(&d.bflags[bb], prx::kAwake);
            }
        }
        added += tot;
    }
    if (threadIdx.x == 0) {
        int nadd = added;
        if (nc0 + nadd > H.contact_cap) {
            H.status |= prx::kStatusContactOverflow;
            nadd = H.contact_cap - nc0;
        }
        H.ncontacts = nc0 + nadd;
        if (statSlot >= 0) H.stats[statSlot] += std::uint32_t(nadd);
    }
}

__global__ void k_clear_moved(DeviceState d) {
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p < d.hdr[0].nproxies) d.pflags[p] &= ~prx::kProxyMoved;
}

// ---------------------------------------------------------------------------
// Solve.

// Body constraints (GetBodyConstraints, AabbTreeWorld.cpp:276-287): every
// dynamic body takes part and is kept awake; others keep their pose.
__global__ void k_big_bodies(DeviceState d, StepParams c) {
    const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= d.hdr[0].nbodies) return;
    const std::uint32_t f = d.bflags[b];
    const float4 p1 = d.pos1[b];
    if ((f & prx::kSpeedable) && (f & prx::kEnabled)) {
        d.bflags[b] = f | prx::kAwake;
        d.pos0[b] = p1;
        d.s_bc_pos[b] = p1;
        d.s_bc_vel[b] = integrated_velocity(d, b, c);
    } else {
        d.s_bc_pos[b] = p1;
        d.s_bc_vel[b] = float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
}

__device__ __forceinline__ bool moves(const DeviceState& d, int b) {
    return (d.bflags[b] & prx::kAccelerable) != 0;
}

// Touching, enabled, non-sensor contacts in contact order (one workgroup),
// and the colouring's initial state: no contact coloured, no colour used at
// any body, a new claim epoch.
__global__ __launch_bounds__(kBlock) void k_big_touch_list(DeviceState d, BigState g) {
    __shared__ int s_scan[kBlock / kWave + 1];
    const int nc = d.hdr[0].ncontacts;
    const int nb = d.hdr[0].nbodies;
    int m = 0;
    for (int base = 0; base < nc; base += kBlock) {
        const int i = base + int(threadIdx.x);
        bool t = false;
        if (i < nc) {
            const std::uint32_t f = d.c_flags[i];
            t = (f & prx::kCEnabled) && (f & prx::kTouching) && !(f & prx::kSensor);
        }
        int tot = 0;
        const int r = block_exclusive_scan(t ? 1 : 0, s_scan, tot);
        if (t) g.list[m + r] = i;
        m += tot;
    }
    for (int e = int(threadIdx.x); e < m; e += kBlock) g.color[e] = -1;
    for (int b = int(threadIdx.x); b < nb; b += kBlock) g.used[b] = 0ull;
    if (threadIdx.x == 0) {
        g.counters[prx::kBigTouching] = m;
        g.counters[prx::kBigColors] = 0;
        g.counters[prx::kBigRounds] = 0;
        g.counters[prx::kBigLeft + 1] = m;  // read by round 0 as "left after the previous round"
        g.counters[prx::kBigEpoch] += 1;
    }
}

// Graph colouring by rounds of speculative first-fit (two grid-wide launches
// per round).  A colour is a set of touching contacts in which no dynamic
// body appears twice: every uncoloured contact proposes the lowest colour
// free at both its dynamic bodies; at each (body, colour) the lowest list
// position wins; contacts that won all their claims take the colour.  Static
// and kinematic bodies are never changed by a contact solve, so they do not
// constrain the colouring.  A round starts only if the previous one left
// contacts uncoloured.
__device__ __forceinline__ unsigned long long claim_key(const BigState& g, int round, int e) {
    const unsigned epoch = unsigned(g.counters[prx::kBigEpoch]) * 64u + unsigned(round);
    return (static_cast<unsigned long long>(~epoch) << 32) | static_cast<unsigned>(e);
}

__global__ __launch_bounds__(kGrid) void k_col_propose(DeviceState d, BigState g, int round) {
    // The count this round's commit accumulates starts at zero, also when the
    // colouring is already complete, so that "done" carries to later rounds.
    const bool done = g.counters[prx::kBigLeft + ((round + 1) & 1)] == 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        g.counters[prx::kBigLeft + (round & 1)] = 0;
        if (!done) g.counters[prx::kBigRounds] = round + 1;
    }
    if (done) return;
    const int m = g.counters[prx::kBigTouching];
    for (int e = int(blockIdx.x * blockDim.x + threadIdx.x); e < m; e += int(gridDim.x * blockDim.x)) {
        if (g.color[e] >= 0) continue;
        const int2 bd = d.c_body[g.list[e]];
        const bool ma = moves(d, bd.x), mb = moves(d, bd.y);
        std::uint64_t mask = 0ull;
        if (ma) mask |= g.used[bd.x];
        if (mb) mask |= g.used[bd.y];
        const unsigned long long freeCols = ~mask & ((1ull << prx::kBigMaxColors) - 1ull);
        const int col = freeCols ? (__ffsll(freeCols) - 1) : -1;
        g.tent[e] = col;
        if (col < 0) continue;
        const unsigned long long key = claim_key(g, round, e);
        if (ma) atomicMin(&g.claim[bd.x * prx::kBigMaxColors + col], key);
        if (mb) atomicMin(&g.claim[bd.y * prx::kBigMaxColors + col], key);
    }
}

__global__ __launch_bounds__(kGrid) void k_col_commit(DeviceState d, BigState g, int round) {
    if (g.counters[prx::kBigLeft + ((round + 1) & 1)] == 0) return;
    const int m = g.counters[prx::kBigTouching];
    int left = 0, ncol = 0;
    for (int e = int(blockIdx.x * blockDim.x + threadIdx.x); e < m; e += int(gridDim.x * blockDim.x)) {
        if (g.color[e] >= 0) continue;
        const int col = g.tent[e];
        if (col < 0) {
            ++left;
            continue;
        }
        const int2 bd = d.c_body[g.list[e]];
        const bool ma = moves(d, bd.x), mb = moves(d, bd.y);
        const unsigned long long key = claim_key(g, round, e);
        const bool wa = !ma || g.claim[bd.x * prx::kBigMaxColors + col] == key;
        const bool wb = !mb || g.claim[bd.y * prx::kBigMaxColors + col] == key;
        if (wa && wb) {
            g.color[e] = col;
            const unsigned long long bit = 1ull << col;
            if (ma) atomicOr(reinterpret_cast<unsigned long long*>(g.used + bd.x), bit);
            if (mb) atomicOr(reinterpret_cast<unsigned long long*>(g.used + bd.y), bit);
            ncol = (col + 1 > ncol) ? col + 1 : ncol;
        } else {
            ++left;
        }
    }
    // One atomic per wavefront.
    for (int off = kWave / 2; off > 0; off >>= 1) {
        left += __shfl_xor(left, off);
        const int o = __shfl_xor(ncol, off);
        ncol = (o > ncol) ? o : ncol;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (left) atomicAdd(&g.counters[prx::kBigLeft + (round & 1)], left);
        if (ncol) atomicMax(&g.counters[prx::kBigColors], ncol);
    }
}


// Colour counts and the stable colour-bucket order of the touching list (one
// workgroup), after the colouring; contacts left uncoloured go to a last
// bucket solved by one thread in list order.  Also clears the per-iteration
// early-exit slots of the sweeps.
__global__ __launch_bounds__(kBlock) void k_big_buckets(DeviceState d, BigState g) {
    __shared__ int s_cnt[prx::kBigMaxColors + 2];
    const int m = g.counters[prx::kBigTouching];
    const int serial = prx::kBigMaxColors;
    for (int e = int(threadIdx.x); e < m; e += kBlock)
        if (g.color[e] < 0) g.color[e] = serial;
    for (int k = int(threadIdx.x); k < prx::kBigMaxColors + 2; k += kBlock) s_cnt[k] = 0;
    for (int k = int(threadIdx.x); k < prx::kBigSweepSlots; k += kBlock)
        g.sweep[k] = (k < prx::kBigPosSlot) ? 0 : 0x7f800000;  // 0.0f; key of +inf
    __syncthreads();
    for (int e = int(threadIdx.x); e < m; e += kBlock) atomicAdd(&s_cnt[g.color[e]], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int k = 0; k <= prx::kBigMaxColors; ++k) {
            const int cnt = s_cnt[k];
            g.colorStart[k] = run;
            run += cnt;
        }
        g.colorStart[prx::kBigMaxColors + 1] = run;
    }
    __syncthreads();
    // Stable bucket order by colour (list order kept within a colour): the
    // list is walked in tiles; each tile's rank comes from the radix-sort
    // style per-wave matching on the colour (< 256).
    {
        __shared__ int s_base[prx::kBigMaxColors + 1];
        __shared__ int s_wave[kBlock / kWave][prx::kBigMaxColors + 1];
        __shared__ int s_tile[prx::kBigMaxColors + 1];
        const int lane = int(threadIdx.x) & (kWave - 1), wave = int(threadIdx.x) >> 6;
        for (int k = int(threadIdx.x); k <= prx::kBigMaxColors; k += kBlock) s_base[k] = g.colorStart[k];
        for (int t = int(threadIdx.x); t < (kBlock / kWave) * (prx::kBigMaxColors + 1); t += kBlock)
            (&s_wave[0][0])[t] = 0;
        __syncthreads();
        for (int base = 0; base < m; base += kBlock) {
            const int e = base + int(threadIdx.x);
            const bool valid = e < m;
            const int col = valid ? g.color[e] : 0;
            unsigned long long peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 7; ++b) {
                const bool bit = (col >> b) & 1;
                const unsigned long long mm = __ballot(bit);
                peers &= bit ? mm : ~mm;
            }
            const int rank = __popcll(peers & lanes_below(lane));
            if (valid && rank == 0) s_wave[wave][col] = __popcll(peers);
            __syncthreads();
            for (int t = int(threadIdx.x); t <= prx::kBigMaxColors; t += kBlock) {
                int run = 0;
                for (int w = 0; w < kBlock / kWave; ++w) {
                    const int cnt = s_wave[w][t];
                    s_wave[w][t] = run;
                    run += cnt;
                }
                s_tile[t] = run;
            }
            __syncthreads();
            if (valid) g.list2[s_base[col] + s_wave[wave][col] + rank] = g.list[e];
            __syncthreads();
            for (int t = int(threadIdx.x); t <= prx::kBigMaxColors; t += kBlock) {
                s_base[t] += s_tile[t];
                for (int w = 0; w < kBlock / kWave; ++w) s_wave[w][t] = 0;
            }
            __syncthreads();
        }
    }
}

// Constraint records in colour order (GetPositionConstraints /
// GetVelocityConstraints, AabbTreeWorld.cpp:289-357), one thread per contact.
__global__ __launch_bounds__(kGrid) void k_big_records(DeviceState d, BigState g, StepParams c) {
    const int m = g.counters[prx::kBigTouching];
    const Bodies<false> bs{d.s_bc_pos, d.s_bc_vel, d.hdr[0].nbodies};
    for (int e = int(blockIdx.x * blockDim.x + threadIdx.x); e < m; e += int(gridDim.x * blockDim.x)) {
        const int gi = g.list2[e];
        const int2 bd = d.c_body[gi];
        const int2 ss = d.c_shape[gi];
        const float rA = d.shape_f[ss.x].x, rB = d.shape_f[ss.y].x;
        const BodyC A = load_bodyc(d, bd.x), B = load_bodyc(d, bd.y);
        const float4 pA = bs.pos(bd.x), pB = bs.pos(bd.y);
        const Xf xfA = get_xf(pA.x, pA.y, pA.z, A.lc);
        const Xf xfB = get_xf(pB.x, pB.y, pB.z, B.lc);
        const Manifold mf = load_manifold(d, gi);
        const WM wm = world_manifold(mf, xfA, rA, xfB, rB);
        const float4 mt = d.c_mat[gi];
        const Vc vc = make_vc(wm, mt.x, mt.y, mt.z, bd.x, bd.y, A, pA, bs.vel(bd.x), B, pB, bs.vel(bd.y), c);
        store_vc(d.s_vc + std::size_t(e) * prx::kVcStride, vc, A, B);
        store_pc(d.s_pc + std::size_t(e) * prx::kPcStride, mf, bd.x, bd.y, rA + rB, A, B);
    }
}

// Contacts [e0, e1) of colour k for this launch: the whole grid strides over a
// colour (no dynamic body appears twice in it); the last bucket is walked in
// order by one thread.
struct ColourRange {
    int e, e1, step;
    __device__ __forceinline__ ColourRange(const BigState& g, int k) {
        const int a = g.colorStart[k], b = g.colorStart[k + 1];
        if (k == prx::kBigMaxColors) {
            const bool first = blockIdx.x == 0 && threadIdx.x == 0;
            e = a;
            e1 = first ? b : a;
            step = 1;
        } else {
            e = a + int(blockIdx.x * blockDim.x + threadIdx.x);
            e1 = b;
            step = int(gridDim.x * blockDim.x);
        }
    }
};

// Velocity iteration `it` stops the sweeps if an earlier iteration's largest
// impulse change was at most regMinMomentum (the early exit of
// SolveRegIslandViaGS, AabbTreeWorld.cpp:1610-1616).
__device__ __forceinline__ bool velocity_done(const BigState& g, const StepParams& c, int it) {
    for (int j = 0; j < it; ++j)
        if (__int_as_float(g.sweep[prx::kBigVelSlot + j]) <= c.regMinMomentum) return true;
    return false;
}

// Position iteration `it` stops if an earlier one reached regMinSeparation
// (AabbTreeWorld.cpp:1627-1642).  Slots hold order-preserving keys of floats.
__device__ __forceinline__ int sep_key(float f) {
    const int bits = __float_as_int(f);
    return bits >= 0 ? bits : (bits ^ 0x7fffffff);
}
__device__ __forceinline__ float sep_of_key(int key) { return __int_as_float(key >= 0 ? key : (key ^ 0x7fffffff)); }
__device__ __forceinline__ bool position_done(const BigState& g, const StepParams& c, int it) {
    for (int j = 0; j < it; ++j)
        if (sep_of_key(g.sweep[prx::kBigPosSlot + j]) >= c.regMinSeparation) return true;
    return false;
}

// Warm start of colour k (WarmStartVelocities, AabbTreeWorld.cpp:264-274).
__global__ __launch_bounds__(kGrid) void k_big_warm(DeviceState d, BigState g, int k) {
    const Bodies<false> bs{d.s_bc_pos, d.s_bc_vel, d.hdr[0].nbodies};
    for (ColourRange r(g, k); r.e < r.e1; r.e += r.step) {
        BodyC A, B;
        const Vc vc = load_vc(d.s_vc + std::size_t(r.e) * prx::kVcStride, A, B);
        warm_start_one(vc, A, B, bs);
    }
}

// Velocity iteration `it` over colour k (SolveVelocityConstraintsViaGS).
__global__ __launch_bounds__(kGrid) void k_big_velocity(DeviceState d, BigState g, StepParams c, int it, int k) {
    if (velocity_done(g, c, it)) return;
    const Bodies<false> bs{d.s_bc_pos, d.s_bc_vel, d.hdr[0].nbodies};
    float inc = 0.0f;
    for (ColourRange r(g, k); r.e < r.e1; r.e += r.step) {
        float4* rec = d.s_vc + std::size_t(r.e) * prx::kVcStride;
        BodyC A, B;
        Vc vc = load_vc(rec, A, B);
        const float mi = solve_velocity_one(vc, A, B, bs);
        rec[4] = float4{vc.p[0].ni, vc.p[0].ti, vc.p[0].nm, vc.p[0].tm};
        rec[7] = float4{vc.p[1].ni, vc.p[1].ti, vc.p[1].nm, vc.p[1].tm};
        inc = fmax_std(inc, mi);
    }
    for (int off = kWave / 2; off > 0; off >>= 1) inc = fmax_std(inc, __shfl_xor(inc, off));
    if ((threadIdx.x & (kWave - 1)) == 0 && inc > 0.0f) atomicMax(&g.sweep[prx::kBigVelSlot + it], __float_as_int(inc));
}

// IntegratePositions (AabbTreeWorld.cpp:176-188).
__global__ __launch_bounds__(kGrid) void k_big_integrate(DeviceState d, StepParams c) {
    const int nb = d.hdr[0].nbodies;
    const Bodies<false> bs{d.s_bc_pos, d.s_bc_vel, nb};
    for (int b = int(blockIdx.x * blockDim.x + threadIdx.x); b < nb; b += int(gridDim.x * blockDim.x)) {
        const std::uint32_t fb = d.bflags[b];
        if (!((fb & prx::kSpeedable) && (fb & prx::kEnabled))) continue;
        float4 p = bs.pos(b);
        const float4 v = bs.vel(b);
        p.x = p.x + c.dt * v.x;
        p.y = p.y + c.dt * v.y;
        p.z = p.z + c.dt * v.z;
        bs.set_pos(b, p);
    }
}

// Position iteration `it` over colour k (SolvePositionConstraintsViaGS).
__global__ __launch_bounds__(kGrid) void k_big_position(DeviceState d, BigState g, StepParams c, int it, int k) {
    if (position_done(g, c, it)) return;
    const Bodies<false> bs{d.s_bc_pos, d.s_bc_vel, d.hdr[0].nbodies};
    float ms = __builtin_inff();
    for (ColourRange r(g, k); r.e < r.e1; r.e += r.step) {
        const Pc pc = load_pc(d.s_pc + std::size_t(r.e) * prx::kPcStride);
        float4 pa = bs.pos(pc.bA);
        float4 pb = bs.pos(pc.bB);
        const float sep = solve_pc(pc.m, pc.totalRadius, pc.A, pc.B, pa, pb, c);
        bs.set_pos(pc.bA, pa);
        bs.set_pos(pc.bB, pb);
        ms = fmin_std(ms, sep);
    }
    for (int off = kWave / 2; off > 0; off >>= 1) ms = fmin_std(ms, __shfl_xor(ms, off));
    if ((threadIdx.x & (kWave - 1)) == 0 && ms < __builtin_inff())
        atomicMin(&g.sweep[prx::kBigPosSlot + it], sep_key(ms));
}

// AssignImpulses (AabbTreeWorld.cpp:208-227) and the body writeback
// (AabbTreeWorld.cpp:1652-1664); records the smallest separation of the last
// position iteration run.
__global__ __launch_bounds__(kGrid) void k_big_finish(DeviceState d, BigState g, StepParams c) {
    const int m = g.counters[prx::kBigTouching];
    const int nb = d.hdr[0].nbodies;
    const int tid = int(blockIdx.x * blockDim.x + threadIdx.x), stride = int(gridDim.x * blockDim.x);
    if (tid == 0) {
        float last = __builtin_inff();
        for (int j = 0; j < int(c.posIters); ++j) {
            if (position_done(g, c, j)) break;
            last = sep_of_key(g.sweep[prx::kBigPosSlot + j]);
        }
        g.metrics[prx::kBigMinSeparation] = last;
    }
    for (int e = tid; e < m; e += stride) {
        const int gi = g.list2[e];
        const float4* rec = d.s_vc + std::size_t(e) * prx::kVcStride;
        const std::uint32_t n = __float_as_uint(rec[1].w);
        float4 m0 = d.m_p0[gi];
        m0.z = rec[4].x;
        m0.w = rec[4].y;
        d.m_p0[gi] = m0;
        if (n > 1) {
            float4 m1 = d.m_p1[gi];
            m1.z = rec[7].x;
            m1.w = rec[7].y;
            d.m_p1[gi] = m1;
        }
    }
    const Bodies<false> bs{d.s_bc_pos, d.s_bc_vel, nb};
    for (int b = tid; b < nb; b += stride) {
        const std::uint32_t f = d.bflags[b];
        if (!((f & prx::kSpeedable) && (f & prx::kEnabled))) continue;
        const float4 bv = bs.vel(b);
        const float4 bp = bs.pos(b);
        const float4 p1 = d.pos1[b];
        if (!(p1.x == bp.x && p1.y == bp.y) || p1.z != bp.z) {
            d.pos1[b] = float4{bp.x, bp.y, bp.z, 0.0f};
            const float4 l = d.lcinv[b];
            const Xf xf = get_xf(bp.x, bp.y, bp.z, V2{l.x, l.y});
            d.xf[b] = float4{xf.p.x, xf.p.y, xf.q.x, xf.q.y};
            d.bflags[b] = f | 0x100u;
        }
        float4 v = d.vel[b];
        v.x = bv.x;
        v.y = bv.y;
        v.z = bv.z;
        d.vel[b] = v;
    }
}

// Synchronize (AabbTreeWorld.cpp:1532-1544, 2564-2585) for every dynamic body.
__global__ void k_big_sync(DeviceState d, StepParams c) {
    const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= d.hdr[0].nbodies) return;
    const std::uint32_t f = d.bflags[b];
    if (!((f & prx::kSpeedable) && (f & prx::kEnabled))) return;
    const WorldCtx x = big_ctx(d);
    const std::uint32_t moved = synchronize_body(d, x, b, c);
    if (moved) atomicAdd(&d.hdr[0].stats[prx::kStRegProxiesMoved], moved);
}

__global__ void k_big_flag_contacts(DeviceState d) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= d.hdr[0].ncontacts) return;
    const int2 bd = d.c_body[i];
    if ((d.bflags[bd.x] | d.bflags[bd.y]) & 0x100u) d.c_flags[i] |= prx::kNeedsUpdating;
}

__global__ void k_big_clear_bodyflag(DeviceState d) {
    const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (b < d.hdr[0].nbodies) d.bflags[b] &= ~0x100u;
}

// Coloured mode with doToi: continuous collision of dynamic bodies against
// static shapes, after the regular phase.  Every enabled non-sensor contact
// between an awake dynamic body and a static body gets its time of impact
// over the step's sweep (GetToiViaSat, TimeOfImpact.cpp:35-251, as in the
// per-world TOI phase); each body keeps its earliest (ties: the lower contact
// index).  That body is advanced to its pose at the time of impact (Body::
// Advance, Body.hpp:1004-1009), the contact's velocity constraint is built
// and solved there against the static body with the TOI velocity iterations
// (VelocityConstraint, SolveVelocityConstraint with restitution and friction,
// as SolveToiViaGS does for the contact, AabbTreeWorld.cpp:1991-2057) after
// its position iterations at the TOI resolution rate, and the body is
// integrated over the rest of the step with the new velocity.  Its
// proxies and contacts are then synchronised and updated.  A second pass
// checks the sweep from the solved pose to the new end pose.  This is a
// reduced form of the reference's TOI phase: at most two events per body and
// step, the island is the body and the static body of its earliest contact,
// and contacts between dynamic bodies get no time of impact.
__global__ void k_big_toi_static(DeviceState d, BigState g, StepParams c) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= d.hdr[0].ncontacts) return;
    const std::uint32_t f = d.c_flags[i];
    if (!(f & prx::kCEnabled) || (f & prx::kSensor)) return;
    const int2 bd = d.c_body[i];
    const std::uint32_t fa = d.bflags[bd.x], fb = d.bflags[bd.y];
    const bool aDyn = (fa & prx::kAccelerable) && !(fb & prx::kSpeedable);
    const bool bDyn = (fb & prx::kAccelerable) && !(fa & prx::kSpeedable);
    if (!aDyn && !bDyn) return;
    const int dyn = aDyn ? bd.x : bd.y;
    if (!(d.bflags[dyn] & prx::kAwake)) return;
    const SweepD sA = normalized(sweep_of(d, bd.x));
    const SweepD sB = normalized(sweep_of(d, bd.y));
    const int2 sh = d.c_shape[i];
    const ToiRes r = toi_via_sat(shape_proxy(d, sh.x), sA, shape_proxy(d, sh.y), sB, c);
    if (r.time >= 0.0f && r.time < 1.0f)
        atomicMin(&g.toiKey[dyn], (static_cast<unsigned long long>(__float_as_uint(r.time)) << 32) |
                                      static_cast<unsigned long long>(std::uint32_t(i)));
}

__global__ void k_big_toi_reset(DeviceState d, BigState g) {
    const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (b < d.hdr[0].nbodies) g.toiLeft[b] = 1.0f;
}

__global__ void k_big_toi_apply(DeviceState d, BigState g, StepParams c) {
    const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= d.hdr[0].nbodies) return;
    const unsigned long long key = g.toiKey[b];
    if (key == ~0ull) return;
    g.toiKey[b] = ~0ull;
    const float t = __uint_as_float(std::uint32_t(key >> 32));  // a fraction of the sweep's time
    const int i = int(std::uint32_t(key & 0xffffffffull));
    const int2 bd = d.c_body[i];
    const int2 ss = d.c_shape[i];
    const int o = (bd.x == b) ? bd.y : bd.x;  // the static body
    float4 p = lerp_pos(d.pos0[b], d.pos1[b], t);
    const BodyC Bd = load_bodyc(d, b), Bo = load_bodyc(d, o);
    const float4 po = d.pos1[o];
    const Xf xfD = get_xf(p.x, p.y, p.z, Bd.lc);
    const Xf xfO = get_xf(po.x, po.y, po.z, Bo.lc);
    const bool dynIsA = (bd.x == b);
    const Xf& xfA = dynIsA ? xfD : xfO;
    const Xf& xfB = dynIsA ? xfO : xfD;
    Manifold m = collide_shapes(shape_proxy(d, ss.x), xfA, shape_proxy(d, ss.y), xfB, c.linearSlop, c.maxCirclesRatio);
    float4 v = d.vel[b];
    if (m.n > 0) {
        const float rA = d.shape_f[ss.x].x, rB = d.shape_f[ss.y].x;
        const BodyC& A = dynIsA ? Bd : Bo;
        const BodyC& B = dynIsA ? Bo : Bd;
        // Position iterations at the TOI resolution rate until the TOI minimum
        // separation (the TOI island's position solve, SolveToiViaGS).
        StepParams ps = c;
        ps.regResolutionRate = c.toiResolutionRate;
        float4 qo = po;
        for (std::uint32_t it = 0; it < c.toiPosIters; ++it) {
            float4& qa = dynIsA ? p : qo;
            float4& qb = dynIsA ? qo : p;
            const float sep = solve_pc(m, rA + rB, A, B, qa, qb, ps);
            if (sep >= c.toiMinSeparation) break;
        }
        // Velocity constraint at the solved pose, no warm start, TOI iterations.
        const float4 pA = dynIsA ? p : po, pB = dynIsA ? po : p;
        const Xf xA = get_xf(pA.x, pA.y, pA.z, A.lc), xB = get_xf(pB.x, pB.y, pB.z, B.lc);
        const WM wm = world_manifold(m, xA, rA, xB, rB);
        const float4 mt = d.c_mat[i];
        const float4 vo{0.0f, 0.0f, 0.0f, 0.0f};
        float4 vA = dynIsA ? v : vo, vB = dynIsA ? vo : v;
        StepParams vconf = c;
        vconf.warmStart = 0u;
        Vc vc = make_vc(wm, mt.x, mt.y, mt.z, bd.x, bd.y, A, pA, vA, B, pB, vB, vconf);
        for (std::uint32_t it = 0; it < c.toiVelIters; ++it) {
            float inc = solve_dir(vc, true, A, B, vA, vB);
            if (vc.n == 1 || (vc.K00 == 0.0f && vc.K11 == 0.0f && vc.K01 == 0.0f))
                inc = fmax_std(inc, solve_dir(vc, false, A, B, vA, vB));
            else
                inc = fmax_std(inc, block_solve(vc, A, B, vA, vB));
            if (inc <= c.toiMinMomentum) break;
        }
        const float4 vn = dynIsA ? vA : vB;
        v.x = vn.x;
        v.y = vn.y;
        v.z = vn.z;
        d.vel[b] = v;
    }
    // The rest of the step with the new velocity; the sweep from the solved
    // pose to the new end pose is what a further pass of this step checks.
    const float left = g.toiLeft[b] * (1.0f - t);
    g.toiLeft[b] = left;
    const float h = left * c.dt;
    d.pos0[b] = p;
    p.x = p.x + h * v.x;
    p.y = p.y + h * v.y;
    p.z = p.z + h * v.z;
    d.pos1[b] = p;
    const Xf xf = get_xf(p.x, p.y, p.z, Bd.lc);
    d.xf[b] = float4{xf.p.x, xf.p.y, xf.q.x, xf.q.y};
    d.bflags[b] |= 0x100u;  // contacts to update
    const WorldCtx x = big_ctx(d);
    const std::uint32_t moved = synchronize_body(d, x, b, c);
    atomicAdd(&d.hdr[0].stats[prx::kStToiIslandsSolved], 1u);
    if (moved) atomicAdd(&d.hdr[0].stats[prx::kStToiProxiesMoved], moved);
}

// Smallest separation over touching contacts at the current poses (a
// diagnostic; separation below zero is penetration).
__global__ void k_big_metrics(DeviceState d, BigState g) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= d.hdr[0].ncontacts) return;
    const std::uint32_t f = d.c_flags[i];
    if (!(f & prx::kTouching) || (f & prx::kSensor)) return;
    const int2 bd = d.c_body[i];
    const int2 ss = d.c_shape[i];
    const float4 xa = d.xf[bd.x], xb = d.xf[bd.y];
    const Xf xfA{V2{xa.x, xa.y}, V2{xa.z, xa.w}}, xfB{V2{xb.x, xb.y}, V2{xb.z, xb.w}};
    const Manifold mf = load_manifold(d, i);
    const float tr = d.shape_f[ss.x].x + d.shape_f[ss.y].x;
    float s = __builtin_inff();
#pragma unroll
    for (std::uint32_t k = 0; k < 2; ++k)
        if (k < mf.n) s = fmin_std(s, get_psm(mf, k, xfA, xfB).sep - tr);
    if (s < __builtin_inff()) {
        // Order-preserving integer image of the float for atomicMin.
        const int bits = __float_as_int(s);
        const int key = bits >= 0 ? bits : (bits ^ 0x7fffffff);
        atomicMin(reinterpret_cast<int*>(g.counters + prx::kBigMinSepBits), key);
    }
}

}  // namespace prxd

// ---------------------------------------------------------------------------
// Host-side launch sequences (compiled in this translation unit).

namespace {

unsigned Blocks(int n, int t) { return unsigned((n + t - 1) / t > 0 ? (n + t - 1) / t : 1); }

// FindContacts + AddContacts over the LBVH (statSlot: the stats word counting
// the added contacts), then the new contacts' narrowphase.
void FindAdd(const prx::DeviceState& d, const prx::BigState& g, const prx::StepParams& c, int statSlot,
             hipStream_t s) {
    using namespace prxd;
    const int P = g.P, cap = g.cap;
    hipLaunchKernelGGL(k_lbvh_bounds, dim3(1), dim3(kBlock), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_morton, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_sort, dim3(1), dim3(kBlock), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_build, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_refit, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_escape, dim3(Blocks(2 * P, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_hash_clear, dim3(Blocks(g.hashMask + 1, kGrid)), dim3(kGrid), 0, s, g);
    hipLaunchKernelGGL(k_hash_insert, dim3(Blocks(cap, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_pairs<0>, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_scan, dim3(1), dim3(kBlock), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_pairs<1>, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_pairs_filter, dim3(Blocks(g.pairCap, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_pairs_append, dim3(1), dim3(kBlock), 0, s, d, g, statSlot);
    hipLaunchKernelGGL(k_big_update, dim3(Blocks(cap, kGrid)), dim3(kGrid), 0, s, d, c, -1);
    hipLaunchKernelGGL(k_clear_moved, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d);
}

}  // namespace

// One step of the large world (P proxies, B bodies, contact capacity cap).
extern "C" hipError_t prx_big_launch_step(const prx::DeviceState* dp, const prx::BigState* gp,
                                          const prx::StepParams* c, hipStream_t s) {
    using namespace prxd;
    const prx::DeviceState& d = *dp;
    const prx::BigState& g = *gp;
    const int P = g.P, B = g.B, cap = g.cap;
    hipLaunchKernelGGL(k_big_pending, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d, *c);
    hipLaunchKernelGGL(k_big_destroy, dim3(1), dim3(kBlock), 0, s, d);
    hipLaunchKernelGGL(k_big_update, dim3(Blocks(cap, kGrid)), dim3(kGrid), 0, s, d, *c, prx::kStPreUpdated);
    FindAdd(d, g, *c, prx::kStPreAdded, s);
    if (c->dt != 0.0f) {
        hipLaunchKernelGGL(k_big_bodies, dim3(Blocks(B, kGrid)), dim3(kGrid), 0, s, d, *c);
        hipLaunchKernelGGL(k_big_touch_list, dim3(1), dim3(kBlock), 0, s, d, g);
        const unsigned colBlocks = Blocks(cap, kGrid) < 1024u ? Blocks(cap, kGrid) : 1024u;
        for (int r = 0; r < prx::kBigColorRounds; ++r) {
            hipLaunchKernelGGL(k_col_propose, dim3(colBlocks), dim3(kGrid), 0, s, d, g, r);
            hipLaunchKernelGGL(k_col_commit, dim3(colBlocks), dim3(kGrid), 0, s, d, g, r);
        }
        hipLaunchKernelGGL(k_big_buckets, dim3(1), dim3(kBlock), 0, s, d, g);
        hipLaunchKernelGGL(k_big_records, dim3(colBlocks), dim3(kGrid), 0, s, d, g, *c);
        const int ncol = prx::kBigMaxColors + 1;
        if (c->warmStart)
            for (int k = 0; k < ncol; ++k) hipLaunchKernelGGL(k_big_warm, dim3(kSweepBlocks), dim3(kGrid), 0, s, d, g, k);
        for (int it = 0; it < int(c->velIters); ++it)
            for (int k = 0; k < ncol; ++k)
                hipLaunchKernelGGL(k_big_velocity, dim3(kSweepBlocks), dim3(kGrid), 0, s, d, g, *c, it, k);
        hipLaunchKernelGGL(k_big_integrate, dim3(Blocks(B, kGrid)), dim3(kGrid), 0, s, d, *c);
        for (int it = 0; it < int(c->posIters); ++it)
            for (int k = 0; k < ncol; ++k)
                hipLaunchKernelGGL(k_big_position, dim3(kSweepBlocks), dim3(kGrid), 0, s, d, g, *c, it, k);
        hipLaunchKernelGGL(k_big_finish, dim3(colBlocks), dim3(kGrid), 0, s, d, g, *c);
        hipLaunchKernelGGL(k_big_sync, dim3(Blocks(B, kGrid)), dim3(kGrid), 0, s, d, *c);
        hipLaunchKernelGGL(k_big_flag_contacts, dim3(Blocks(cap, kGrid)), dim3(kGrid), 0, s, d);
        hipLaunchKernelGGL(k_big_clear_bodyflag, dim3(Blocks(B, kGrid)), dim3(kGrid), 0, s, d);
        hipLaunchKernelGGL(k_big_update, dim3(Blocks(cap, kGrid)), dim3(kGrid), 0, s, d, *c, -1);
        FindAdd(d, g, *c, prx::kStRegAdded, s);
        if (c->doToi) {
            // Two passes: a body's second event in a step (another static
            // shape reached while moving on) is found from the sweep the
            // first left.
            hipLaunchKernelGGL(k_big_toi_reset, dim3(Blocks(B, kGrid)), dim3(kGrid), 0, s, d, g);
            for (int pass = 0; pass < kBigToiPasses; ++pass) {
                hipLaunchKernelGGL(k_big_toi_static, dim3(Blocks(cap, kGrid)), dim3(kGrid), 0, s, d, g, *c);
                hipLaunchKernelGGL(k_big_toi_apply, dim3(Blocks(B, kGrid)), dim3(kGrid), 0, s, d, g, *c);
                hipLaunchKernelGGL(k_big_flag_contacts, dim3(Blocks(cap, kGrid)), dim3(kGrid), 0, s, d);
                hipLaunchKernelGGL(k_big_clear_bodyflag, dim3(Blocks(B, kGrid)), dim3(kGrid), 0, s, d);
                hipLaunchKernelGGL(k_big_update, dim3(Blocks(cap, kGrid)), dim3(kGrid), 0, s, d, *c, -1);
                FindAdd(d, g, *c, prx::kStToiAdded, s);
            }
        }
    }
    return hipGetLastError();
}

// Pair finding only (for tests of the LBVH against brute force): proxies'
// fat AABBs, bodies and moved flags must be set in d.
extern "C" hipError_t prx_big_launch_pairs(const prx::DeviceState* dp, const prx::BigState* gp, hipStream_t s) {
    using namespace prxd;
    const prx::DeviceState& d = *dp;
    const prx::BigState& g = *gp;
    const int P = g.P;
    hipLaunchKernelGGL(k_lbvh_bounds, dim3(1), dim3(kBlock), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_morton, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_sort, dim3(1), dim3(kBlock), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_build, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_refit, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_escape, dim3(Blocks(2 * P, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_pairs<0>, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_scan, dim3(1), dim3(kBlock), 0, s, d, g);
    hipLaunchKernelGGL(k_lbvh_pairs<1>, dim3(Blocks(P, kGrid)), dim3(kGrid), 0, s, d, g);
    return hipGetLastError();
}

extern "C" hipError_t prx_big_launch_metrics(const prx::DeviceState* dp, const prx::BigState* gp, hipStream_t s) {
    using namespace prxd;
    hipLaunchKernelGGL(k_big_metrics, dim3(Blocks(gp->cap, kGrid)), dim3(kGrid), 0, s, *dp, *gp);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pair finding without allocation.  The LBVH's working arrays for up to
// kNoAllocMaxProxies proxies live in one file-scope device array; the caller
// passes device buffers for the inputs and outputs and a stream.  Each device
// holds its own copy of the array, whose address is looked up once per
// device (hipGetSymbolAddress) and used for every kernel argument.

namespace {

constexpr int kNoAllocMaxProxies = 1 << 16;

constexpr std::size_t Align256(std::size_t n) { return (n + 255) & ~std::size_t(255); }

// Byte offsets of the working arrays inside the scratch array (sized for
// kNoAllocMaxProxies; a call with fewer proxies uses the front of each).
struct NoAllocLayout {
    static constexpr std::size_t N = std::size_t(kNoAllocMaxProxies);
    static constexpr std::size_t hdr = 0;
    static constexpr std::size_t paabb = Align256(hdr + sizeof(prx::WorldHeader));
    static constexpr std::size_t pflags = Align256(paabb + 16 * N);
    static constexpr std::size_t morton = Align256(pflags + 4 * N);
    static constexpr std::size_t morton2 = Align256(morton + 4 * N);
    static constexpr std::size_t order = Align256(morton2 + 4 * N);
    static constexpr std::size_t order2 = Align256(order + 4 * N);
    static constexpr std::size_t left = Align256(order2 + 4 * N);
    static constexpr std::size_t right = Align256(left + 4 * N);
    static constexpr std::size_t parent = Align256(right + 4 * N);
    static constexpr std::size_t escape = Align256(parent + 8 * N);
    static constexpr std::size_t visit = Align256(escape + 8 * N);
    static constexpr std::size_t naabb = Align256(visit + 4 * N);
    static constexpr std::size_t bounds = Align256(naabb + 32 * N);
    static constexpr std::size_t pcount = Align256(bounds + 16);
    static constexpr std::size_t counters = Align256(pcount + 4 * (N + 1));
    static constexpr std::size_t bytes = Align256(counters + 4 * prx::kBigCounters);
};

}  // namespace

// About 6.6 MB per device, held until the process ends.
static __device__ __align__(256) unsigned char scratch[NoAllocLayout::bytes];

namespace prxd {

// Inputs into the LBVH's proxy arrays: fat AABBs (4 floats each, any 4-byte
// alignment), moved flags; the header carries the proxy count and a cleared
// status (a fill by kernel, so the call can be captured in a graph).
__global__ void k_noalloc_prep(DeviceState d, const float* aabbs, const std::uint8_t* moved, int n) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i == 0) {
        d.hdr[0].nproxies = n;
        d.hdr[0].status = 0;
    }
    if (i >= n) return;
    d.paabb[i] = float4{aabbs[4 * i], aabbs[4 * i + 1], aabbs[4 * i + 2], aabbs[4 * i + 3]};
    d.pflags[i] = moved[i] ? prx::kProxyMoved : 0u;
}

// The pair count for the caller: the number found (above cap: only cap were
// written), or -1 when the tree walk exceeded its depth bound.
__global__ void k_noalloc_count(DeviceState d, BigState g, int* count) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int total = g.pcount[g.P];
    const bool depth = (d.hdr[0].status & prx::kStatusScratchOverflow) != 0 && total <= g.pairCap;
    *count = depth ? -1 : total;
}

}  // namespace prxd

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mutex;
unsigned char* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// The current device's copy of the scratch array, looked up on the first
// call on that device.
hipError_t NoAllocScratch(unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_noalloc_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_noalloc_ptr[dev];
    return hipSuccess;
}

}  // namespace

// Returns a PRX_* status; *err names the failure.
extern "C" int prx_lbvh_pairs_noalloc_impl(const float* aabbs, const std::int32_t* bodies, const std::uint8_t* moved,
                                           int n, std::int32_t* pairs, int cap, int* count, void* stream,
                                           const char** err) {
    using namespace prxd;
    if (!aabbs || !bodies || !moved || n <= 0 || !count || cap < 0 || (cap > 0 && !pairs)) {
        *err = "bad argument";
        return PRX_E_INVALID_ARGUMENT;
    }
    if (n > kNoAllocMaxProxies) {
        *err = "more proxies than the allocation-free pair finder holds (65536)";
        return PRX_E_LENGTH;
    }
    if ((reinterpret_cast<std::uintptr_t>(pairs) & 7u) || (reinterpret_cast<std::uintptr_t>(bodies) & 3u) ||
        (reinterpret_cast<std::uintptr_t>(aabbs) & 3u) || (reinterpret_cast<std::uintptr_t>(count) & 3u)) {
        *err = "misaligned buffer (pairs need 8-byte, the others 4-byte alignment)";
        return PRX_E_INVALID_ARGUMENT;
    }
    unsigned char* base = nullptr;
    if (NoAllocScratch(&base) != hipSuccess) {
        *err = "device scratch lookup failed";
        return PRX_E_DEVICE;
    }
    using L = NoAllocLayout;
    prx::DeviceState d{};
    d.nworlds = 1;
    d.hdr = reinterpret_cast<prx::WorldHeader*>(base + L::hdr);
    d.paabb = reinterpret_cast<float4*>(base + L::paabb);
    d.pflags = reinterpret_cast<std::uint32_t*>(base + L::pflags);
    d.pbody = const_cast<int*>(reinterpret_cast<const int*>(bodies));
    prx::BigState g{};
    g.P = n;
    g.B = 1;
    g.cap = 1;
    g.pairCap = cap;
    g.morton = reinterpret_cast<std::uint32_t*>(base + L::morton);
    g.morton2 = reinterpret_cast<std::uint32_t*>(base + L::morton2);
    g.order = reinterpret_cast<int*>(base + L::order);
    g.order2 = reinterpret_cast<int*>(base + L::order2);
    g.left = reinterpret_cast<int*>(base + L::left);
    g.right = reinterpret_cast<int*>(base + L::right);
    g.parent = reinterpret_cast<int*>(base + L::parent);
    g.escape = reinterpret_cast<int*>(base + L::escape);
    g.visit = reinterpret_cast<int*>(base + L::visit);
    g.naabb = reinterpret_cast<float4*>(base + L::naabb);
    g.bounds = reinterpret_cast<float4*>(base + L::bounds);
    g.pcount = reinterpret_cast<int*>(base + L::pcount);
    g.counters = reinterpret_cast<int*>(base + L::counters);
    g.pairs = reinterpret_cast<int2*>(pairs);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_noalloc_prep, dim3(Blocks(n, kGrid)), dim3(kGrid), 0, s, d, aabbs, moved, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = prx_big_launch_pairs(&d, &g, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_noalloc_count, dim3(1), dim3(64), 0, s, d, g, count);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        *err = hipGetErrorString(e);
        return PRX_E_DEVICE;
    }
    return PRX_OK;
}
