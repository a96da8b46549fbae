// Device-side building blocks of the step: per-contact narrowphase update,
// contact destruction/creation, island search, constraint records and the
// Gauss-Seidel solvers.  Shared by the per-world batched step kernel
// (step_kernel.hip) and the single-large-world kernels (big_world.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"
#include "narrowphase.hpp"

namespace prxd {

using prx::DeviceState;
using prx::StepParams;
using prx::WorldHeader;

constexpr int kWave = 64;
constexpr int kChunk = prx::kStageRecords;

// Phase timer: lane 0 records shader-clock deltas into phase_cycles[w].
struct PhaseTimer {
    unsigned long long* out;
    unsigned long long t0, start;
    __device__ __forceinline__ void begin(unsigned long long* o) {
        out = o;
        t0 = start = (o && threadIdx.x == 0) ? clock64() : 0ull;
    }
    __device__ __forceinline__ void mark(int phase) {
        if (out && threadIdx.x == 0) {
            const unsigned long long t = clock64();
            out[phase] += t - t0;
            t0 = t;
        }
    }
    __device__ __forceinline__ void end() {
        if (out && threadIdx.x == 0) out[prx::kPhTotal] += clock64() - start;
    }
};

__device__ __forceinline__ unsigned long long lanes_below(int lane) {
    return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

// Inclusive wave scan of a 64-lane value.
__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ Proxy shape_proxy(const DeviceState& d, int s) {
    const int4 si = d.shape_i[s];
    const float4 sf = d.shape_f[s];
    return Proxy{d.verts + si.y, d.norms + si.y, std::uint32_t(si.x), sf.x};
}

__device__ __forceinline__ Xf body_xf1(const DeviceState& d, int gb) {
    const float4 t = d.xf[gb];
    return Xf{V2{t.x, t.y}, V2{t.z, t.w}};
}
__device__ __forceinline__ Xf body_xf0(const DeviceState& d, int gb) {
    const float4 p = d.pos0[gb];
    const float4 l = d.lcinv[gb];
    return get_xf(p.x, p.y, p.z, V2{l.x, l.y});
}

// ComputeAABB(proxy, xf0, xf1) (AABB.cpp:51-62) with optional fattening by
// aabbExtension and displacement by displaceMultiplier * (p1 - p0)
// (detail/AABB.hpp:305-319, Interval.hpp:144-220).
__device__ __forceinline__ float4 proxy_aabb(const Proxy& px, const Xf& xf0, const Xf& xf1, const StepParams& c, bool fatten) {
    float mnx = __builtin_inff(), mny = __builtin_inff();
    float mxx = -__builtin_inff(), mxy = -__builtin_inff();
    for (std::uint32_t i = 0; i < px.n; ++i) {
        const V2 v = px.vert(i);
        const V2 a = transform(v, xf0);
        mnx = fmin_std(a.x, mnx);
        mxx = fmax_std(a.x, mxx);
        mny = fmin_std(a.y, mny);
        mxy = fmax_std(a.y, mxy);
        const V2 b = transform(v, xf1);
        mnx = fmin_std(b.x, mnx);
        mxx = fmax_std(b.x, mxx);
        mny = fmin_std(b.y, mny);
        mxy = fmax_std(b.y, mxy);
    }
    mnx = mnx - px.radius;
    mxx = mxx + px.radius;
    mny = mny - px.radius;
    mxy = mxy + px.radius;
    if (fatten) {
        mnx = mnx - c.aabbExtension;
        mxx = mxx + c.aabbExtension;
        mny = mny - c.aabbExtension;
        mxy = mxy + c.aabbExtension;
        const V2 disp = mul(c.displaceMultiplier, xf1.p - xf0.p);
        if (disp.x < 0.0f) mnx += disp.x;
        else mxx += disp.x;
        if (disp.y < 0.0f) mny += disp.y;
        else mxy += disp.y;
    }
    return float4{mnx, mny, mxx, mxy};
}

// Dynamic LDS of the step kernels (size StepParams::ldsBytes in k_step_worlds,
// none in k_toi_worlds).
extern __shared__ float4 g_smem[];

// IsIntersecting per axis (Interval.hpp:292-298).
__device__ __forceinline__ bool overlap(float4 a, float4 b) {
    const float mx0 = fmax_std(a.x, b.x);
    const float mn0 = fmin_std(a.z, b.z);
    if (!(mn0 >= mx0)) return false;
    const float mx1 = fmax_std(a.y, b.y);
    const float mn1 = fmin_std(a.w, b.w);
    return mn1 >= mx1;
}

__device__ __forceinline__ void body_set_awake(const DeviceState& d, int gb) {
    // Body::SetAwake (Body.cpp:163-171)
    const std::uint32_t f = d.bflags[gb];
    if (f & prx::kSpeedable) {
        d.bflags[gb] = f | prx::kAwake;
        d.vel[gb].w = 0.0f;
    }
}

__device__ __forceinline__ Manifold load_manifold(const DeviceState& d, int gi) {
    Manifold m;
    const uint4 h = d.m_hdr[gi];
    const float4 g = d.m_geo[gi];
    const float4 p0 = d.m_p0[gi];
    const float4 p1 = d.m_p1[gi];
    m.type = h.x & 0xffu;
    m.n = (h.x >> 8) & 0xffu;
    m.ln = V2{g.x, g.y};
    m.lpt = V2{g.z, g.w};
    m.p[0] = MPoint{V2{p0.x, p0.y}, h.y, p0.z, p0.w};
    m.p[1] = MPoint{V2{p1.x, p1.y}, h.z, p1.z, p1.w};
    return m;
}

__device__ __forceinline__ void store_manifold(const DeviceState& d, int gi, const Manifold& m) {
    d.m_hdr[gi] = uint4{m.type | (m.n << 8), m.n > 0 ? m.p[0].cf : kDefaultCf, m.n > 1 ? m.p[1].cf : kDefaultCf, 0u};
    d.m_geo[gi] = float4{m.ln.x, m.ln.y, m.lpt.x, m.lpt.y};
    d.m_p0[gi] = m.n > 0 ? float4{m.p[0].lp.x, m.p[0].lp.y, m.p[0].ni, m.p[0].ti} : float4{0, 0, 0, 0};
    d.m_p1[gi] = m.n > 1 ? float4{m.p[1].lp.x, m.p[1].lp.y, m.p[1].ni, m.p[1].ti} : float4{0, 0, 0, 0};
}

// Update (AabbTreeWorld.cpp:2587-2721): new manifold, warm-start impulses
// carried over by ContactFeature, else from the nearest old point.  Returns
// the listener calls the update makes (:2703-2719): bit 0 begin-contact, bit 1
// end-contact, bit 2 pre-solve with the manifold before the update (stored
// to oldOut when given).  mapOut (when given) receives which old point each
// new point took its impulses from (prx::Tape::cmap's low bits).
__device__ __forceinline__ std::uint32_t update_contact(const DeviceState& d, int bo, int gi, const StepParams& c,
                                                        Manifold* oldOut = nullptr, std::uint32_t* mapOut = nullptr) {
    std::uint32_t flags = d.c_flags[gi];
    const Manifold old = load_manifold(d, gi);
    const bool oldTouching = (flags & prx::kTouching) != 0;
    bool newTouching = false;
    Manifold nm = empty_manifold();
    std::uint32_t src = prx::kNoSource | (prx::kNoSource << 4);
    if (!(flags & prx::kSensor)) {
        const int2 bb = d.c_body[gi];
        const int2 ss = d.c_shape[gi];
        const Proxy pa = shape_proxy(d, ss.x);
        const Proxy pb = shape_proxy(d, ss.y);
        nm = collide_shapes(pa, body_xf1(d, bo + bb.x), pb, body_xf1(d, bo + bb.y), c.linearSlop, c.maxCirclesRatio);
        const std::uint32_t oldN = old.n, newN = nm.n;
        newTouching = newN > 0;
        bool found[2] = {false, newN < 2};
#pragma unroll
        for (std::uint32_t i = 0; i < 2; ++i) {
            if (i >= newN) continue;
            bool done = false;
#pragma unroll
            for (std::uint32_t j = 0; j < 2; ++j) {
                if (done || j >= oldN) continue;
                if (nm.p[i].cf == old.p[j].cf) {
                    found[i] = true;
                    nm.p[i].ni = old.p[j].ni;
                    nm.p[i].ti = old.p[j].ti;
                    src = (src & ~(0xfu << (4 * i))) | (j << (4 * i));
                    done = true;
                }
            }
        }
#pragma unroll
        for (std::uint32_t i = 0; i < 2; ++i) {
            if (i >= newN || found[i]) continue;
            float least = __builtin_inff();
#pragma unroll
            for (std::uint32_t j = 0; j < 2; ++j) {
                if (j >= oldN) continue;
                const float dd = magsq(old.p[j].lp - nm.p[i].lp);
                if (least > dd) {
                    least = dd;
                    nm.p[i].ni = old.p[j].ni;
                    nm.p[i].ti = old.p[j].ti;
                    src = (src & ~(0xfu << (4 * i))) | (j << (4 * i));
                }
            }
        }
    }
    store_manifold(d, gi, nm);
    if (mapOut) *mapOut = src | (nm.n << 8);
    flags &= ~prx::kNeedsUpdating;
    std::uint32_t calls = 0;
    if (!oldTouching && newTouching) {
        flags |= prx::kTouching;
        calls = 1u;
    } else if (oldTouching && !newTouching) {
        flags &= ~prx::kTouching;
        calls = 2u;
    }
    if (newTouching && !(flags & prx::kSensor)) calls |= 4u;
    d.c_flags[gi] = flags;
    if (oldOut) *oldOut = old;
    return calls;
}

struct WorldCtx {
    int w, bo, nb, po, np, co, cap, jo, nj;
    int eo, ecap;  // listener-call words of the world (ecap 0: not recorded)
    int lane;
};

struct Counters {
    int nc, nfree, pool;
    std::uint32_t status;
    int evw, evn;  // listener-call words and calls of the step so far
};

// Listener calls are appended wave-wide in call order: every lane passes the
// words and calls it appends and gets the first word of its own, behind those
// of the lanes below it, or -1 (nothing to write, or past the world's
// segment: the step then reports kStatusEventOverflow).
__device__ __forceinline__ int ev_claim(const WorldCtx& x, Counters& k, int words, int calls) {
    const int incl = wave_incl_scan(words, x.lane);
    const int cincl = wave_incl_scan(calls, x.lane);
    const int at = k.evw + incl - words;
    k.evw += __shfl(incl, kWave - 1);
    k.evn += __shfl(cincl, kWave - 1);
    const bool over = words > 0 && at + words > x.ecap;
    if (__ballot(over)) k.status |= prx::kStatusEventOverflow;
    return (words > 0 && !over) ? x.eo + at : -1;
}

// The listener calls of one contact update (update_contact's bits), the
// pre-solve call with the manifold as it was before the update.
__device__ __forceinline__ void ev_update(const DeviceState& d, const WorldCtx& x, Counters& k, std::uint32_t calls,
                                          std::uint32_t cid, const Manifold& old) {
    const bool touch = (calls & 3u) != 0, pre = (calls & 4u) != 0;
    const int at = ev_claim(x, k, (touch ? 2 : 0) + (pre ? prx::kEvPreWords : 0), (touch ? 1 : 0) + (pre ? 1 : 0));
    if (at < 0) return;
    std::uint32_t* o = d.ev + at;
    if (touch) {
        o[0] = (calls & 1u) ? prx::kEvBegin : prx::kEvEnd;
        o[1] = cid;
        o += 2;
    }
    if (pre) {
        const bool face = old.type == kFaceA || old.type == kFaceB;
        const bool set = old.type != kUnset;
        o[0] = prx::kEvPreSolve;
        o[1] = cid;
        o[2] = old.type | (old.n << 8);
        o[3] = old.n > 0 ? old.p[0].cf : kDefaultCf;
        o[4] = old.n > 1 ? old.p[1].cf : kDefaultCf;
        o[5] = face ? __float_as_uint(old.ln.x) : 0u;
        o[6] = face ? __float_as_uint(old.ln.y) : 0u;
        o[7] = set ? __float_as_uint(old.lpt.x) : 0u;
        o[8] = set ? __float_as_uint(old.lpt.y) : 0u;
#pragma unroll
        for (std::uint32_t q = 0; q < 2; ++q) {
            const bool has = q < old.n;
            o[9 + 4 * q] = has ? __float_as_uint(old.p[q].lp.x) : 0u;
            o[10 + 4 * q] = has ? __float_as_uint(old.p[q].lp.y) : 0u;
            o[11 + 4 * q] = has ? __float_as_uint(old.p[q].ni) : 0u;
            o[12 + 4 * q] = has ? __float_as_uint(old.p[q].ti) : 0u;
        }
    }
}

// UpdateContacts (AabbTreeWorld.cpp:2333-2421).  Returns contacts updated.
// The contacts that need an update, grouped by the pair of their shapes'
// vertex counts (disk 1, edge 2, polygons 3..16), so that the lanes of a
// wave take the same narrowphase path: each contact's update depends on that
// contact alone, so the grouping leaves every result as it is.  A counting
// sort in LDS (ldsAvail bytes); returns -1 where it does not fit, else the
// number of listed contacts.
constexpr int kUpdKeys = 17 * 17;

__device__ __forceinline__ int update_key(const DeviceState& d, int gi) {
    if (d.c_flags[gi] & prx::kSensor) return 0;
    const int2 ss = d.c_shape[gi];
    const int na = d.shape_i[ss.x].x, nb = d.shape_i[ss.y].x;
    const int lo = (na < nb) ? na : nb, hi = (na < nb) ? nb : na;
    return lo * 17 + hi;
}

__device__ __forceinline__ int group_updates(const DeviceState& d, const WorldCtx& x, int nc, int ldsAvail) {
    if (4 * (kUpdKeys + 1) + 6 * nc > ldsAvail) return -1;
    int* start = reinterpret_cast<int*>(g_smem);   // kUpdKeys + 1 counters
    int* list = start + kUpdKeys + 1;               // nc contact indices
    unsigned short* key = reinterpret_cast<unsigned short*>(list + nc);  // nc keys
    for (int k = x.lane; k <= kUpdKeys; k += kWave) start[k] = 0;
    __syncthreads();
    for (int i = x.lane; i < nc; i += kWave) {
        int kk = kUpdKeys;  // no update
        if (d.c_flags[x.co + i] & prx::kNeedsUpdating) {
            kk = update_key(d, x.co + i);
            atomicAdd(&start[kk], 1);
        }
        key[i] = static_cast<unsigned short>(kk);
    }
    __syncthreads();
    int carry = 0;
    for (int base = 0; base < kUpdKeys; base += kWave) {
        const int k = base + x.lane;
        const int v = (k < kUpdKeys) ? start[k] : 0;
        const int incl = wave_incl_scan(v, x.lane);
        if (k < kUpdKeys) start[k] = carry + incl - v;
        carry += __shfl(incl, kWave - 1);
    }
    __syncthreads();
    for (int i = x.lane; i < nc; i += kWave) {
        const int kk = key[i];
        if (kk < kUpdKeys) list[atomicAdd(&start[kk], 1)] = i;
    }
    __syncthreads();
    return carry;
}

// cmap (when given: a tape's carry map of the world) receives each updated
// contact's carry map, by pool id, keeping its kCarryFresh bit.
__device__ __forceinline__ void update_mapped(const DeviceState& d, const WorldCtx& x, int gi, const StepParams& c,
                                              std::uint32_t* cmap) {
    std::uint32_t m = 0;
    update_contact(d, x.bo, gi, c, nullptr, &m);
    std::uint32_t& e = cmap[d.c_id[gi]];
    e = (e & prx::kCarryFresh) | m;
}

__device__ __forceinline__ std::uint32_t phase_update(const DeviceState& d, const WorldCtx& x, Counters& k,
                                                      const StepParams& c, int ldsAvail,
                                                      std::uint32_t* cmap = nullptr) {
    const int nc = k.nc;
    if (x.ecap == 0) {
        const int n = group_updates(d, x, nc, ldsAvail);
        if (n >= 0) {
            const int* list = reinterpret_cast<const int*>(g_smem) + kUpdKeys + 1;
            for (int r = x.lane; r < n; r += kWave) {
                if (cmap) update_mapped(d, x, x.co + list[r], c, cmap);
                else update_contact(d, x.bo, x.co + list[r], c);
            }
            __syncthreads();
            return std::uint32_t(n);
        }
    }
    std::uint32_t updated = 0;
    for (int base = 0; base < nc; base += kWave) {
        const int i = base + x.lane;
        bool need = false;
        if (i < nc) need = (d.c_flags[x.co + i] & prx::kNeedsUpdating) != 0;
        updated += std::uint32_t(__popcll(__ballot(need)));
        if (x.ecap > 0) {
            Manifold old = empty_manifold();
            std::uint32_t calls = 0;
            std::uint32_t m = 0;
            if (need) calls = update_contact(d, x.bo, x.co + i, c, &old, cmap ? &m : nullptr);
            if (need && cmap) {
                std::uint32_t& e = cmap[d.c_id[x.co + i]];
                e = (e & prx::kCarryFresh) | m;
            }
            ev_update(d, x, k, calls, need ? d.c_id[x.co + i] : 0u, old);
        } else if (need) {
            if (cmap) update_mapped(d, x, x.co + i, c, cmap);
            else update_contact(d, x.bo, x.co + i, c);
        }
    }
    return updated;
}

// DestroyContacts (AabbTreeWorld.cpp:2290-2331) + InternalDestroy (:2263-2288).
__device__ __forceinline__ std::uint32_t phase_destroy(const DeviceState& d, const WorldCtx& x, Counters& k) {
    int out = 0;
    std::uint32_t destroyedTotal = 0;
    const int nc = k.nc;
    for (int base = 0; base < nc; base += kWave) {
        const int i = base + x.lane;
        const bool valid = i < nc;
        bool keep = false;
        if (valid) {
            const int2 pr = d.c_prox[x.co + i];
            keep = overlap(d.paabb[x.po + pr.x], d.paabb[x.po + pr.y]);
        }
        const bool destroyed = valid && !keep;
        const unsigned long long dm = __ballot(destroyed);
        const unsigned long long km = __ballot(valid && keep);
        const unsigned long long below = lanes_below(x.lane);
        if (destroyed) {
            const int r = __popcll(dm & below);
            d.c_free[x.co + k.nfree + r] = d.c_id[x.co + i];
            const std::uint32_t mh = d.m_hdr[x.co + i].x;
            const std::uint32_t fl = d.c_flags[x.co + i];
            if (((mh >> 8) & 0xffu) > 0 && !(fl & prx::kSensor)) {
                const int2 bd = d.c_body[x.co + i];
                body_set_awake(d, x.bo + bd.x);
                body_set_awake(d, x.bo + bd.y);
            }
        }
        if (x.ecap > 0) {
            // End-contact for a destroyed touching contact (AabbTreeWorld.cpp:2249-2252).
            const bool end = destroyed && (d.c_flags[x.co + i] & prx::kTouching) != 0;
            const int at = ev_claim(x, k, end ? 2 : 0, end ? 1 : 0);
            if (at >= 0) {
                d.ev[at] = prx::kEvEnd;
                d.ev[at + 1] = d.c_id[x.co + i];
            }
        }
        const int dst = out + __popcll(km & below);
        const bool move = valid && keep && (dst != i);
        if (__ballot(move)) {
            // Stable compaction: every lane loads its record, then the wave
            // synchronises, then the moved records are stored (dst <= i).
            int2 pr{0, 0}, bd{0, 0}, sh{0, 0};
            std::uint32_t id = 0, fl = 0;
            float4 mt{0, 0, 0, 0}, mg{0, 0, 0, 0}, p0{0, 0, 0, 0}, p1{0, 0, 0, 0};
            uint4 mh{0, 0, 0, 0};
            if (move) {
                const int src = x.co + i;
                pr = d.c_prox[src];
                bd = d.c_body[src];
                sh = d.c_shape[src];
                id = d.c_id[src];
                fl = d.c_flags[src];
                mt = d.c_mat[src];
                mh = d.m_hdr[src];
                mg = d.m_geo[src];
                p0 = d.m_p0[src];
                p1 = d.m_p1[src];
            }
            __syncthreads();
            if (move) {
                const int g = x.co + dst;
                d.c_prox[g] = pr;
                d.c_body[g] = bd;
                d.c_shape[g] = sh;
                d.c_id[g] = id;
                d.c_flags[g] = fl;
                d.c_mat[g] = mt;
                d.m_hdr[g] = mh;
                d.m_geo[g] = mg;
                d.m_p0[g] = p0;
                d.m_p1[g] = p1;
            }
        }
        k.nfree += __popcll(dm);
        out += __popcll(km);
        destroyedTotal += std::uint32_t(__popcll(dm));
        __syncthreads();
    }
    k.nc = out;
    return destroyedTotal;
}

// ShouldCollide for two shapes' filters (Filter.hpp:81-88).
__device__ __forceinline__ bool should_collide(const DeviceState& d, int sa, int sb) {
    const int ga = d.shape_i[sa].w, gb = d.shape_i[sb].w;
    if (ga == gb && ga != 0) return ga > 0;
    const std::uint32_t cma = d.shape_catmask[sa], cmb = d.shape_catmask[sb];
    return (((cma >> 16) & (cmb & 0xffffu)) != 0) && (((cmb >> 16) & (cma & 0xffffu)) != 0);
}

// ShouldCollide for two bodies' joints (AabbTreeWorld.cpp:479-490): false if
// a joint between them does not let its bodies collide.
__device__ __forceinline__ bool joint_allows(const DeviceState& d, const WorldCtx& x, int ba, int bb) {
    if (x.nj == 0) return true;
    const int2 bj = d.bjoint[x.bo + ba];
    for (int k = 0; k < bj.y; ++k) {
        const int2 e = d.bjoint_list[2 * x.jo + bj.x + k];
        if (e.y == bb && !(__float_as_uint(d.j_cfg2[x.jo + e.x].y) & prx::kJointCollideConnected)) return false;
    }
    return true;
}

// FindContacts (AabbTreeWorld.cpp:709-751) + AddContacts (:2424-2547).
// Candidate keys pack world-local proxy indices (min << 16 | max); proxy
// indices increase with tree leaf ids, so key order equals ContactKey order.
__device__ __forceinline__ int readlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// ldsAvail: bytes of dynamic LDS the caller's kernel has (free during this
// phase).  When the candidates fit, they are sorted and looked up there.
__device__ __forceinline__ std::uint32_t phase_find_add(const DeviceState& d, const WorldCtx& x, Counters& k, const StepParams& c,
                                                      int ldsAvail) {
    // Moved proxies, in index order.
    int nmoved = 0;
    for (int base = 0; base < x.np; base += kWave) {
        const int p = base + x.lane;
        bool mv = false;
        if (p < x.np) mv = (d.pflags[x.po + p] & prx::kProxyMoved) != 0;
        const unsigned long long m = __ballot(mv);
        if (mv) {
            const int r = nmoved + __popcll(m & lanes_below(x.lane));
            d.s_isl_body[x.bo + x.co + r] = p;  // scratch list (segment is nbodies + cap >= nproxies)
        }
        nmoved += __popcll(m);
    }
    __syncthreads();
    if (nmoved == 0) return 0;
    int* moved = d.s_isl_body + x.bo + x.co;

    // Candidate pairs: every proxy q (one per lane) against every moved proxy
    // m.  The moved proxies come in chunks of one per lane (index, body and
    // fat AABB in registers) and the wave reads them one at a time from the
    // holding lane, so the inner loop does no memory access.
    int ncand = 0;
    bool overflow = false;
    for (int base = 0; base < x.np; base += kWave) {
        const int q = base + x.lane;
        const bool qv = q < x.np;
        float4 qa = float4{0, 0, 0, 0};
        int qb = -1;
        bool qmoved = false;
        if (qv) {
            qa = d.paabb[x.po + q];
            qb = d.pbody[x.po + q];
            qmoved = (d.pflags[x.po + q] & prx::kProxyMoved) != 0;
        }
        for (int mb = 0; mb < nmoved; mb += kWave) {
            const int cnt = (nmoved - mb < kWave) ? nmoved - mb : kWave;
            int mi = 0, mbody = -1;
            float4 ma = float4{0, 0, 0, 0};
            if (x.lane < cnt) {
                mi = moved[mb + x.lane];
                mbody = d.pbody[x.po + mi];
                ma = d.paabb[x.po + mi];
            }
            for (int j = 0; j < cnt; ++j) {
                const int m = readlane_i(mi, j);
                const int mbd = readlane_i(mbody, j);
                const float4 am{readlane_f(ma.x, j), readlane_f(ma.y, j), readlane_f(ma.z, j), readlane_f(ma.w, j)};
                // A pair of two moved proxies is produced once, from the smaller.
                const bool hit = qv && q != m && !(qmoved && q < m) && qb != mbd && overlap(am, qa);
                const unsigned long long hm = __ballot(hit);
                if (hit) {
                    const int r = ncand + __popcll(hm & lanes_below(x.lane));
                    if (r < x.cap) {
                        const int lo = q < m ? q : m, hi = q < m ? m : q;
                        d.s_keys[x.co + r] = (std::uint32_t(lo) << 16) | std::uint32_t(hi);
                    }
                }
                ncand += __popcll(hm);
            }
        }
    }
    if (ncand > x.cap) {
        overflow = true;
        ncand = x.cap;
    }
    __syncthreads();
    // Clear the moved flags (proxiesForContacts is consumed).
    for (int p = x.lane; p < x.np; p += kWave) d.pflags[x.po + p] &= ~prx::kProxyMoved;
    if (ncand == 0) {
        if (overflow) k.status |= prx::kStatusScratchOverflow;
        return 0;
    }
    // The (unique) candidate keys in ascending order, and a mark on those
    // that are contacts already.
    const bool ldsSort = x.np <= 0x7fff && 4 * (ncand + 2 * x.np) <= ldsAvail;
    std::uint32_t* skeys = reinterpret_cast<std::uint32_t*>(g_smem);  // sorted keys, bit 31: exists
    if (ldsSort) {
        // Counting sort by the lower proxy index in LDS (histogram, prefix
        // sum, scatter), then each lower index's bucket (the few partners of
        // one proxy) sorted on one lane.  Keys have bit 31 clear (indices
        // below 2^15).
        int* bcnt = reinterpret_cast<int*>(skeys + ncand);
        int* bend = bcnt + x.np;
        for (int i = x.lane; i < x.np; i += kWave) bcnt[i] = 0;
        __syncthreads();
        for (int i = x.lane; i < ncand; i += kWave) atomicAdd(&bcnt[d.s_keys[x.co + i] >> 16], 1);
        __syncthreads();
        {
            int carry = 0;
            for (int b0 = 0; b0 < x.np; b0 += kWave) {
                const int i = b0 + x.lane;
                const int v = (i < x.np) ? bcnt[i] : 0;
                const int incl = wave_incl_scan(v, x.lane);
                if (i < x.np) bend[i] = carry + incl - v;
                carry += __shfl(incl, kWave - 1);
            }
        }
        __syncthreads();
        for (int i = x.lane; i < ncand; i += kWave) {
            const std::uint32_t key = d.s_keys[x.co + i];
            skeys[atomicAdd(&bend[key >> 16], 1)] = key;
        }
        __syncthreads();
        for (int b = x.lane; b < x.np; b += kWave) {
            const int e = bend[b], s0 = e - bcnt[b];
            for (int i = s0 + 1; i < e; ++i) {
                const std::uint32_t key = skeys[i];
                int j = i - 1;
                while (j >= s0 && skeys[j] > key) {
                    skeys[j + 1] = skeys[j];
                    --j;
                }
                skeys[j + 1] = key;
            }
        }
        __syncthreads();
        // Existing contacts: binary search in LDS, mark with bit 31 (each
        // contact's key is unique, so no two lanes write one slot).
        for (int i = x.lane; i < k.nc; i += kWave) {
            const int2 pr = d.c_prox[x.co + i];
            const std::uint32_t key = (std::uint32_t(pr.x) << 16) | std::uint32_t(pr.y);
            int lo = 0, hi = ncand;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if ((skeys[mid] & 0x7fffffffu) < key) lo = mid + 1;
                else hi = mid;
            }
            if (lo < ncand && skeys[lo] == key) skeys[lo] = key | 0x80000000u;
        }
        __syncthreads();
    } else {
        // Rank sort in global scratch.
        for (int i = x.lane; i < ncand; i += kWave) {
            const std::uint32_t key = d.s_keys[x.co + i];
            int rank = 0;
            for (int j = 0; j < ncand; ++j) rank += (d.s_keys[x.co + j] < key) ? 1 : 0;
            d.s_keys2[x.co + rank] = key;
        }
        __syncthreads();
        // Existing contacts: mark candidates already present (binary search).
        for (int i = x.lane; i < ncand; i += kWave) d.s_isl_c[x.co + i] = 0;
        __syncthreads();
        for (int i = x.lane; i < k.nc; i += kWave) {
            const int2 pr = d.c_prox[x.co + i];
            const std::uint32_t key = (std::uint32_t(pr.x) << 16) | std::uint32_t(pr.y);
            int lo = 0, hi = ncand;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (d.s_keys2[x.co + mid] < key) lo = mid + 1;
                else hi = mid;
            }
            if (lo < ncand && d.s_keys2[x.co + lo] == key) d.s_isl_c[x.co + lo] = 1;
        }
        __syncthreads();
    }
    auto sorted_key = [&](int i, bool& exists) -> std::uint32_t {
        if (ldsSort) {
            const std::uint32_t v = skeys[i];
            exists = (v >> 31) != 0;
            return v & 0x7fffffffu;
        }
        exists = d.s_isl_c[x.co + i] != 0;
        return d.s_keys2[x.co + i];
    };
    // Filter, then append in key order.
    std::uint32_t added = 0;
    const int nc0 = k.nc;
    for (int base = 0; base < ncand; base += kWave) {
        const int i = base + x.lane;
        bool acc = false;
        int pa = 0, pb = 0, ba = 0, bb = 0, sa = 0, sb = 0;
        if (i < ncand) {
            bool exists = false;
            const std::uint32_t key = sorted_key(i, exists);
            pa = int(key >> 16);
            pb = int(key & 0xffffu);
            ba = d.pbody[x.po + pa];
            bb = d.pbody[x.po + pb];
            sa = d.pshape[x.po + pa];
            sb = d.pshape[x.po + pb];
            const std::uint32_t fa = d.bflags[x.bo + ba], fb = d.bflags[x.bo + bb];
            acc = ((fa | fb) & prx::kAccelerable) != 0 && should_collide(d, sa, sb) && !exists &&
                  joint_allows(d, x, ba, bb);
        }
        const unsigned long long am = __ballot(acc);
        if (acc) {
            const int r = int(added) + __popcll(am & lanes_below(x.lane));
            const int idx = nc0 + r;
            if (idx < x.cap) {
                std::uint32_t id;
                if (r < k.nfree) id = d.c_free[x.co + k.nfree - 1 - r];
                else id = std::uint32_t(k.pool + (r - k.nfree));
                const int gi = x.co + idx;
                const std::uint32_t fa = d.bflags[x.bo + ba], fb = d.bflags[x.bo + bb];
                const int4 sia = d.shape_i[sa], sib = d.shape_i[sb];
                const float4 sfa = d.shape_f[sa], sfb = d.shape_f[sb];
                std::uint32_t flags = prx::kCEnabled | prx::kNeedsUpdating;
                if ((fa | fb) & prx::kImpenetrable) flags |= prx::kCImpenetrable;
                const bool sensor = sia.z != 0 || sib.z != 0;
                if (sensor) flags |= prx::kSensor;
                d.c_prox[gi] = int2{pa, pb};
                d.c_body[gi] = int2{ba, bb};
                d.c_shape[gi] = int2{sa, sb};
                d.c_id[gi] = id;
                d.c_flags[gi] = flags;
                // MixFriction / MixRestitution (Contact.hpp:57-69)
                d.c_mat[gi] = float4{__builtin_sqrtf(sfa.y * sfb.y), (sfa.z > sfb.z) ? sfa.z : sfb.z, 0.0f, 0.0f};
                d.m_hdr[gi] = uint4{0u, kDefaultCf, kDefaultCf, 0u};
                d.m_geo[gi] = float4{0, 0, 0, 0};
                d.m_p0[gi] = float4{0, 0, 0, 0};
                d.m_p1[gi] = float4{0, 0, 0, 0};
            }
        }
        added += std::uint32_t(__popcll(am));
    }
    __syncthreads();
    // Awake flags (SetAwakeFlag for speedable bodies of non-sensor contacts):
    // idempotent, so done per new contact in a second pass without races on
    // the flags word of a body touched by two lanes.
    int nadd = int(added);
    if (nc0 + nadd > x.cap) {
        k.status |= prx::kStatusContactOverflow;
        nadd = x.cap - nc0;
    }
    if (overflow) k.status |= prx::kStatusScratchOverflow;
    if (x.lane == 0) {
        for (int r = 0; r < nadd; ++r) {
            const int gi = x.co + nc0 + r;
            if (d.c_flags[gi] & prx::kSensor) continue;
            const int2 bd = d.c_body[gi];
            const std::uint32_t fa = d.bflags[x.bo + bd.x];
            if (fa & prx::kSpeedable) d.bflags[x.bo + bd.x] = fa | prx::kAwake;
            const std::uint32_t fb = d.bflags[x.bo + bd.y];
            if (fb & prx::kSpeedable) d.bflags[x.bo + bd.y] = fb | prx::kAwake;
        }
    }
    const int consumedFree = nadd < k.nfree ? nadd : k.nfree;
    k.pool += nadd - consumedFree;
    k.nfree -= consumedFree;
    k.nc = nc0 + nadd;
    __syncthreads();
    // Narrowphase of the new contacts.
    for (int base = nc0; base < k.nc; base += kWave) {
        const int i = base + x.lane;
        const bool v = i < k.nc;
        if (x.ecap > 0) {
            Manifold old = empty_manifold();
            std::uint32_t calls = 0;
            if (v) calls = update_contact(d, x.bo, x.co + i, c, &old);
            ev_update(d, x, k, calls, v ? d.c_id[x.co + i] : 0u, old);
        } else if (v) {
            update_contact(d, x.bo, x.co + i, c);
        }
    }
    __syncthreads();
    return std::uint32_t(nadd);
}

// ---------------------------------------------------------------------------
// Island solve.  Velocity-constraint record (kVcStride float4, stored in
// island order so a sweep reads consecutive records):
//   v[0] normal.x, normal.y, friction, tangentSpeed
//   v[1] K00, K11, K01, count
//   v[2] M00, M11, M01, bodyA            (normal mass = K^-1)
//   v[3] p0 relA.x, relA.y, relB.x, relB.y
//   v[4] p0 ni, ti, normalMass, tangentMass
//   v[5] p0 velocityBias, p1 velocityBias, bodyB, restitution
//   v[6] p1 relA.x, relA.y, relB.x, relB.y
//   v[7] p1 ni, ti, normalMass, tangentMass
//   v[8] invMassA, invRotIA, invMassB, invRotIB
// Position-constraint record (kPcStride float4, island order):
//   u[0] type | count << 8, bodyA, bodyB, totalRadius
//   u[1] local normal x, y, local point x, y
//   u[2] point 0 local x, y, point 1 local x, y
//   u[3] invMassA, invRotIA, localCenterA x, y
//   u[4] invMassB, invRotIB, localCenterB x, y

struct BodyC {
    float invMass, invRotI;
    V2 lc;
};

__device__ __forceinline__ V2 rel_vel(V2 vA, float wA, V2 relA, V2 vB, float wB, V2 relB) {
    // GetContactRelVelocity (d2/Math.cpp:73-103)
    const V2 velBrot = mul(revperp(relB), wB / 1.0f);
    const V2 velArot = mul(revperp(relA), wA / 1.0f);
    return (vB + velBrot) - (vA + velArot);
}

struct VcPoint {
    V2 relA, relB;
    float ni, ti, nm, tm, bias;
};

struct Vc {
    V2 normal;
    float friction, restitution, tangentSpeed;
    std::uint32_t n;
    float K00, K11, K01, M00, M11, M01;
    int bA, bB;
    VcPoint p[2];
};

__device__ __forceinline__ void store_vc(float4* v, const Vc& vc, const BodyC& A, const BodyC& B) {
    v[0] = float4{vc.normal.x, vc.normal.y, vc.friction, vc.tangentSpeed};
    v[1] = float4{vc.K00, vc.K11, vc.K01, __uint_as_float(vc.n)};
    v[2] = float4{vc.M00, vc.M11, vc.M01, __int_as_float(vc.bA)};
    v[3] = float4{vc.p[0].relA.x, vc.p[0].relA.y, vc.p[0].relB.x, vc.p[0].relB.y};
    v[4] = float4{vc.p[0].ni, vc.p[0].ti, vc.p[0].nm, vc.p[0].tm};
    v[5] = float4{vc.p[0].bias, vc.p[1].bias, __int_as_float(vc.bB), vc.restitution};
    v[6] = float4{vc.p[1].relA.x, vc.p[1].relA.y, vc.p[1].relB.x, vc.p[1].relB.y};
    v[7] = float4{vc.p[1].ni, vc.p[1].ti, vc.p[1].nm, vc.p[1].tm};
    v[8] = float4{A.invMass, A.invRotI, B.invMass, B.invRotI};
}
__device__ __forceinline__ Vc load_vc(const float4* v, BodyC& A, BodyC& B) {
    Vc vc;
    const float4 a = v[0], b = v[1], e = v[2], p0a = v[3], p0b = v[4], f = v[5], p1a = v[6], p1b = v[7],
                 m = v[8];
    vc.normal = V2{a.x, a.y};
    vc.friction = a.z;
    vc.tangentSpeed = a.w;
    vc.K00 = b.x;
    vc.K11 = b.y;
    vc.K01 = b.z;
    vc.n = __float_as_uint(b.w);
    vc.M00 = e.x;
    vc.M11 = e.y;
    vc.M01 = e.z;
    vc.bA = __float_as_int(e.w);
    vc.p[0] = VcPoint{V2{p0a.x, p0a.y}, V2{p0a.z, p0a.w}, p0b.x, p0b.y, p0b.z, p0b.w, f.x};
    vc.p[1] = VcPoint{V2{p1a.x, p1a.y}, V2{p1a.z, p1a.w}, p1b.x, p1b.y, p1b.z, p1b.w, f.y};
    vc.bB = __float_as_int(f.z);
    vc.restitution = f.w;
    A = BodyC{m.x, m.y, V2{0, 0}};
    B = BodyC{m.z, m.w, V2{0, 0}};
    return vc;
}

// World manifold point data (WorldManifold.cpp:37-130).
struct WM {
    V2 normal;
    std::uint32_t n;
    V2 p[2];
    float ni[2], ti[2];
};

__device__ __forceinline__ WM world_manifold(const Manifold& m, const Xf& xfA, float rA, const Xf& xfB, float rB) {
    WM w;
    w.n = 0;
    w.normal = V2{0, 0};
    if (m.type == kCircles) {
        const V2 pA = transform(m.lpt, xfA);
        const V2 pB = transform(m.p[0].lp, xfB);
        const V2 normal = unit_vector(pB - pA, V2{1.0f, 0.0f});
        const V2 cA = pA + mul(rA, normal);
        const V2 cB = pB - mul(rB, normal);
        w.normal = normal;
        w.n = 1;
        w.p[0] = divs(cA + cB, 2.0f);
        w.ni[0] = m.p[0].ni;
        w.ti[0] = m.p[0].ti;
    } else if (m.type == kFaceA) {
        const V2 normal = rotate(m.ln, xfA.q);
        const V2 plane = transform(m.lpt, xfA);
        w.normal = normal;
        w.n = m.n;
#pragma unroll
        for (std::uint32_t i = 0; i < 2; ++i) {
            if (i >= m.n) continue;
            const V2 clip = transform(m.p[i].lp, xfB);
            const V2 cA = clip + mul(rA - dot(clip - plane, normal), normal);
            const V2 cB = clip - mul(rB, normal);
            w.p[i] = divs(cA + cB, 2.0f);
            w.ni[i] = m.p[i].ni;
            w.ti[i] = m.p[i].ti;
        }
    } else if (m.type == kFaceB) {
        const V2 normal = rotate(m.ln, xfB.q);
        const V2 plane = transform(m.lpt, xfB);
        w.normal = -normal;
        w.n = m.n;
#pragma unroll
        for (std::uint32_t i = 0; i < 2; ++i) {
            if (i >= m.n) continue;
            const V2 clip = transform(m.p[i].lp, xfA);
            const V2 cB = clip + mul(rB - dot(clip - plane, normal), normal);
            const V2 cA = clip - mul(rA, normal);
            w.p[i] = divs(cA + cB, 2.0f);
            w.ni[i] = m.p[i].ni;
            w.ti[i] = m.p[i].ti;
        }
    }
    return w;
}

// VelocityConstraint ctor + GetPoint + ComputeK (VelocityConstraint.cpp:38-185).
__device__ __forceinline__ Vc make_vc(const WM& wm, float friction, float restitution, float tangentSpeed, int bA, int bB,
                      const BodyC& A, float4 posA, float4 velA, const BodyC& B, float4 posB, float4 velB,
                      const StepParams& c) {
    Vc vc;
    vc.normal = wm.normal;
    vc.friction = friction;
    vc.restitution = restitution;
    vc.tangentSpeed = tangentSpeed;
    vc.n = 0;
    vc.K00 = vc.K11 = vc.K01 = 0.0f;
    vc.M00 = vc.M11 = vc.M01 = 0.0f;
    vc.bA = bA;
    vc.bB = bB;
    vc.p[0] = VcPoint{V2{0, 0}, V2{0, 0}, 0, 0, 0, 0, 0};
    vc.p[1] = vc.p[0];
    const float dtRatio = c.warmStart ? c.dtRatio : 0.0f;
    const float invMass = A.invMass + B.invMass;
    const V2 tangent = fwdperp(vc.normal);
#pragma unroll
    for (std::uint32_t j = 0; j < 2; ++j) {
        if (j >= wm.n) continue;
        VcPoint p;
        const V2 relA = wm.p[j] - V2{posA.x, posA.y};
        const V2 relB = wm.p[j] - V2{posB.x, posB.y};
        p.ni = wm.ni[j] * dtRatio;
        p.ti = wm.ti[j] * dtRatio;
        {
            const V2 dv = rel_vel(V2{velA.x, velA.y}, velA.z, relA, V2{velB.x, velB.y}, velB.z, relB);
            const float vn = dot(dv, vc.normal);
            p.bias = (vn < -c.velocityThreshold) ? -vc.restitution * vn : 0.0f;
        }
        p.relA = relA;
        p.relB = relB;
        {
            const float cA = cross(relA, vc.normal);
            const float cB = cross(relB, vc.normal);
            const float rmA = A.invRotI * (cA * cA) / 1.0f;
            const float rmB = B.invRotI * (cB * cB) / 1.0f;
            const float value = invMass + rmA + rmB;
            p.nm = (value != 0.0f) ? 1.0f / value : 0.0f;
        }
        {
            const float cA = cross(relA, tangent);
            const float cB = cross(relB, tangent);
            const float rmA = A.invRotI * (cA * cA) / 1.0f;
            const float rmB = B.invRotI * (cB * cB) / 1.0f;
            const float value = invMass + rmA + rmB;
            p.tm = (value != 0.0f) ? 1.0f / value : 0.0f;
        }
        // wm points fill vc points in order, so slot j receives point j.
        vc.p[j] = p;
        ++vc.n;
    }
    if (c.blockSolve && vc.n == 2) {
        const float rnA0 = cross(vc.p[0].relA, vc.normal);
        const float rnB0 = cross(vc.p[0].relB, vc.normal);
        const float rnA1 = cross(vc.p[1].relA, vc.normal);
        const float rnB1 = cross(vc.p[1].relB, vc.normal);
        const float irmA0 = (A.invRotI * (rnA0 * rnA0)) / 1.0f;
        const float irmA1 = (A.invRotI * (rnA1 * rnA1)) / 1.0f;
        const float irmB0 = (B.invRotI * (rnB0 * rnB0)) / 1.0f;
        const float irmB1 = (B.invRotI * (rnB1 * rnB1)) / 1.0f;
        const float irmA = (A.invRotI * rnA0 * rnA1) / 1.0f;
        const float irmB = (B.invRotI * rnB0 * rnB1) / 1.0f;
        const float k00 = invMass + irmA0 + irmB0;
        const float k11 = invMass + irmA1 + irmB1;
        const float k01 = invMass + irmA + irmB;
        const float k00sq = k00 * k00;
        const float kdiff = k00 * k11 - k01 * k01;
        if (k00sq < 1000.0f * kdiff) {
            vc.K00 = k00;
            vc.K11 = k11;
            vc.K01 = k01;
            // Invert (Math.hpp:428-442): cp = Cross(row0, row1)
            const float cpm = k00 * k11;
            const float cps = k01 * k01;
            const float cp = cpm - cps;
            if (!almost_zero(cp)) {
                const float inv = 1.0f / cp;
                vc.M00 = k11 * inv;
                vc.M11 = k00 * inv;
                vc.M01 = -k01 * inv;
            }
        } else {
            vc.n = 1;
        }
    }
    return vc;
}

// SolveTangentConstraint / SeqSolveNormalConstraint (ContactSolver.cpp:337-454).
__device__ __forceinline__ float solve_dir(Vc& vc, bool tangent, const BodyC& A, const BodyC& B, float4& velA,
                                           float4& velB) {
    float maxInc = 0.0f;
    const V2 dir = tangent ? fwdperp(vc.normal) : vc.normal;
    V2 vA{velA.x, velA.y}, vB{velB.x, velB.y};
    float wA = velA.z, wB = velB.z;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
        const int k = 1 - kk;  // point 1 first, then point 0
        if (k >= int(vc.n)) continue;
        VcPoint& p = vc.p[k];
        const V2 closing = rel_vel(vA, wA, p.relA, vB, wB, p.relB);
        float oldI, newI;
        if (tangent) {
            const float dirVel = vc.tangentSpeed - dot(closing, dir);
            const float lambda = p.tm * dirVel;
            const float maxI = vc.friction * p.ni;
            oldI = p.ti;
            newI = clamp_std(oldI + lambda, -maxI, maxI);
        } else {
            const float dirVel = dot(closing, dir);
            const float lambda = p.nm * (p.bias - dirVel);
            oldI = p.ni;
            newI = fmax_std(oldI + lambda, 0.0f);
        }
        const float inc = newI - oldI;
        const V2 P = mul(inc, dir);
        const float LA = cross(p.relA, P) / 1.0f;
        const float LB = cross(p.relB, P) / 1.0f;
        vA = vA - mul(A.invMass, P);
        wA = wA - A.invRotI * LA;
        vB = vB + mul(B.invMass, P);
        wB = wB + B.invRotI * LB;
        maxInc = fmax_std(maxInc, __builtin_fabsf(inc));
        if (tangent) p.ti = oldI + inc;
        else p.ni = oldI + inc;
    }
    velA.x = vA.x;
    velA.y = vA.y;
    velA.z = wA;
    velB.x = vB.x;
    velB.y = vB.y;
    velB.z = wB;
    return maxInc;
}

// BlockSolveNormalConstraint (ContactSolver.cpp:65-335).  The four cases
// (both points, point 0 alone, point 1 alone, neither) are all evaluated and
// the first valid one selected, in the reference's order; with no valid case
// the velocities and impulses stay as they are.
__device__ __forceinline__ float block_solve(Vc& vc, const BodyC& A, const BodyC& B, float4& velA, float4& velB) {
    const V2 normal = vc.normal;
    V2 bp;
    {
        const V2 vA{velA.x, velA.y}, vB{velB.x, velB.y};
        const V2 dv0 = rel_vel(vA, velA.z, vc.p[0].relA, vB, velB.z, vc.p[0].relB);
        const V2 dv1 = rel_vel(vA, velA.z, vc.p[1].relA, vB, velB.z, vc.p[1].relB);
        const float vn0 = dot(dv0, normal);
        const float vn1 = dot(dv1, normal);
        const V2 b{vn0 - vc.p[0].bias, vn1 - vc.p[1].bias};
        float t0 = 0.0f;
        t0 += vc.K00 * vc.p[0].ni;
        t0 += vc.K01 * vc.p[1].ni;
        float t1 = 0.0f;
        t1 += vc.K01 * vc.p[0].ni;
        t1 += vc.K11 * vc.p[1].ni;
        bp = V2{b.x - t0, b.y - t1};
    }
    // Case 1: both points (x = -M b').
    float x0 = 0.0f;
    x0 += vc.M00 * bp.x;
    x0 += vc.M01 * bp.y;
    float x1 = 0.0f;
    x1 += vc.M01 * bp.x;
    x1 += vc.M11 * bp.y;
    const float c1n0 = -x0, c1n1 = -x1;
    const bool ok1 = c1n0 >= 0.0f && c1n1 >= 0.0f;
    // Case 2: point 0 alone.
    const float c2n0 = -vc.p[0].nm * bp.x;
    const float vn2 = vc.K01 * c2n0 + bp.y;
    const bool ok2 = c2n0 >= 0.0f && vn2 >= 0.0f;
    // Case 3: point 1 alone.
    const float c3n1 = -vc.p[1].nm * bp.y;
    const float vn1 = vc.K01 * c3n1 + bp.x;
    const bool ok3 = c3n1 >= 0.0f && vn1 >= 0.0f;
    // Case 4: neither.
    const bool ok4 = bp.x >= 0.0f && bp.y >= 0.0f;
    const float n0 = ok1 ? c1n0 : (ok2 ? c2n0 : 0.0f);
    const float n1 = ok1 ? c1n1 : (ok2 ? 0.0f : (ok3 ? c3n1 : 0.0f));
    const bool ok = ok1 || ok2 || ok3 || ok4;
    const float d0 = n0 - vc.p[0].ni;
    const float d1 = n1 - vc.p[1].ni;
    const V2 P0 = mul(d0, normal);
    const V2 P1 = mul(d1, normal);
    const V2 P = P0 + P1;
    const float LA = (cross(vc.p[0].relA, P0) + cross(vc.p[1].relA, P1)) / 1.0f;
    const float LB = (cross(vc.p[0].relB, P0) + cross(vc.p[1].relB, P1)) / 1.0f;
    const V2 dAl = -mul(A.invMass, P);
    const float dAa = -(A.invRotI * LA);
    const V2 dBl = mul(B.invMass, P);
    const float dBa = B.invRotI * LB;
    velA.x = ok ? velA.x + dAl.x : velA.x;
    velA.y = ok ? velA.y + dAl.y : velA.y;
    velA.z = ok ? velA.z + dAa : velA.z;
    velB.x = ok ? velB.x + dBl.x : velB.x;
    velB.y = ok ? velB.y + dBl.y : velB.y;
    velB.z = ok ? velB.z + dBa : velB.z;
    vc.p[0].ni = ok ? n0 : vc.p[0].ni;
    vc.p[1].ni = ok ? n1 : vc.p[1].ni;
    return ok ? fmax_std(__builtin_fabsf(n0), __builtin_fabsf(n1)) : 0.0f;
}

// Position solver manifold (PositionSolverManifold.cpp:42-121).
struct PSM {
    V2 normal, point;
    float sep;
};
// The three manifold types share one form, so that lanes of different types
// run the same instructions: with X the transformation of the body holding
// the manifold's local point (A for circles and face A, B for face B) and Y
// the other, P1 = X(local point), P2 = Y(the manifold point) and the
// separation is dot(P2 - P1, normal).  Circles take the normal from
// P2 - P1 and the midpoint as the point; the face types take the rotated
// local normal and P2, and face B reports the normal negated.
__device__ __forceinline__ PSM get_psm(const Manifold& m, std::uint32_t idx, const Xf& xfA, const Xf& xfB) {
    const V2 plp = idx ? m.p[1].lp : m.p[0].lp;
    const bool faceB = m.type == kFaceB;
    const bool circles = m.type == kCircles;
    const Xf& X = faceB ? xfB : xfA;
    const Xf& Y = faceB ? xfA : xfB;
    const V2 p1 = transform(m.lpt, X);
    const V2 p2 = transform(plp, Y);
    const V2 delta = p2 - p1;
    V2 normal = rotate(m.ln, X.q);
    if (circles) normal = unit_vector(delta, V2{0, 0});
    const float sep = dot(delta, normal);
    const V2 point = circles ? divs(p1 + p2, 2.0f) : p2;
    return PSM{faceB ? -normal : normal, point, sep};
}

// GaussSeidel::SolvePositionConstraint (ContactSolver.cpp:494-643).
__device__ __forceinline__ float solve_pc(const Manifold& m, float totalRadius, const BodyC& A, const BodyC& B, float4& posA,
                          float4& posB, const StepParams& c) {
    const float invMassTotal = A.invMass + B.invMass;
    struct Sol {
        V2 la;
        float aa;
        V2 lb;
        float ab;
        float sep;
    };
    auto solve = [&](const PSM& psm, V2 pA, V2 pB) {
        const float separation = psm.sep - totalRadius;
        const V2 rA = psm.point - pA;
        const V2 rB = psm.point - pB;
        const float crA = cross(rA, psm.normal) / 1.0f;
        const float crB = cross(rB, psm.normal) / 1.0f;
        const float K = invMassTotal + A.invRotI * (crA * crA) + B.invRotI * (crB * crB);
        const float C = -clamp_std(c.regResolutionRate * (separation + c.linearSlop), -c.maxLinearCorrection, 0.0f);
        const V2 P = divs(mul(psm.normal, C), K);
        const float LA = cross(rA, P) / 1.0f;
        const float LB = cross(rB, P) / 1.0f;
        return Sol{-mul(A.invMass, P), -(A.invRotI * LA), mul(B.invMass, P), B.invRotI * LB, separation};
    };
    // One path for every lane: the first point solved is point 0 (one
    // point, equal separations, or point 0 the deeper) or point 1 (point 1
    // the deeper); the second point is solved at the positions and
    // transformations before the first's correction when the separations are
    // almost equal, else after it.
    const V2 oAl{posA.x, posA.y}, oBl{posB.x, posB.y};
    const float oAa = posA.z, oBa = posB.z;
    const Xf xfA = get_xf(oAl.x, oAl.y, oAa, A.lc);
    const Xf xfB = get_xf(oBl.x, oBl.y, oBa, B.lc);
    const PSM psm0 = get_psm(m, 0, xfA, xfB);
    const PSM psm1 = get_psm(m, 1, xfA, xfB);
    const bool two = m.n == 2;
    const bool eq = two && almost_equal(psm0.sep, psm1.sep);
    const bool lt = two && !eq && psm0.sep < psm1.sep;
    const bool gt = two && !eq && psm1.sep < psm0.sep;
    if (!(m.n == 1 || eq || lt || gt)) return __builtin_inff();
    const Sol s0 = solve(gt ? psm1 : psm0, oAl, oBl);
    V2 pAl = oAl + s0.la;
    float pAa = oAa + s0.aa;
    V2 pBl = oBl + s0.lb;
    float pBa = oBa + s0.ab;
    const float minSep = two ? s0.sep : 0.0f + s0.sep;
    if (two) {
        const PSM next = get_psm(m, gt ? 0u : 1u, get_xf(pAl.x, pAl.y, pAa, A.lc), get_xf(pBl.x, pBl.y, pBa, B.lc));
        const Sol s1 = solve(eq ? psm1 : next, eq ? oAl : pAl, eq ? oBl : pBl);
        pAl = pAl + s1.la;
        pAa = pAa + s1.aa;
        pBl = pBl + s1.lb;
        pBa = pBa + s1.ab;
    }
    posA.x = pAl.x;
    posA.y = pAl.y;
    posA.z = pAa;
    posB.x = pBl.x;
    posB.y = pBl.y;
    posB.z = pBa;
    return minSep;
}

// ---------------------------------------------------------------------------
// Revolute joints (RevoluteJointConf.cpp:127-433).  A joint's solver record in
// island order (kJcStride float4):
//   u[0] rA x, y, rB x, y               u[4] invMassA, invRotIA, invMassB, invRotIB
//   u[1] mass exx, eyx, ezx, eyy        u[5] motor speed, dt * max motor torque, body A, body B
//   u[2] ezy, ezz, angular mass, flags | limit state << 8
//   u[3] impulse x, y, z, angular motor impulse (updated by the sweeps)
//   u[6] local anchor A - local center A, local anchor B - local center B
//   u[7] reference angle, lower angle, upper angle, -
struct Jc {
    V2 rA, rB;
    float exx, eyx, ezx, eyy, ezy, ezz, angularMass;
    std::uint32_t flags, limit;
    float ix, iy, iz, motorImp;
    BodyC A, B;
    float motorSpeed, maxImpulse;
    int bA, bB;
};

__device__ __forceinline__ Jc load_jc_velocity(const float4* u) {
    const float4 r = u[0], m = u[1], n = u[2], im = u[3], ms = u[4], mo = u[5];
    Jc j;
    j.rA = V2{r.x, r.y};
    j.rB = V2{r.z, r.w};
    j.exx = m.x;
    j.eyx = m.y;
    j.ezx = m.z;
    j.eyy = m.w;
    j.ezy = n.x;
    j.ezz = n.y;
    j.angularMass = n.z;
    j.flags = __float_as_uint(n.w) & 0xffu;
    j.limit = __float_as_uint(n.w) >> 8;
    j.ix = im.x;
    j.iy = im.y;
    j.iz = im.z;
    j.motorImp = im.w;
    j.A = BodyC{ms.x, ms.y, V2{0, 0}};
    j.B = BodyC{ms.z, ms.w, V2{0, 0}};
    j.motorSpeed = mo.x;
    j.maxImpulse = mo.y;
    j.bA = __float_as_int(mo.z);
    j.bB = __float_as_int(mo.w);
    return j;
}

// Solve33 / Solve22 of the symmetric joint mass (Math.hpp:446-474); vectors
// GetX = (exx, eyx, ezx), GetY = (eyx, eyy, ezy), GetZ = (ezx, ezy, ezz).
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    float r = 0.0f;
    r += ax * bx;
    r += ay * by;
    r += az * bz;
    return r;
}
__device__ __forceinline__ void cross3(float ax, float ay, float az, float bx, float by, float bz, float& cx,
                                       float& cy, float& cz) {
    cx = ay * bz - az * by;
    cy = az * bx - ax * bz;
    cz = ax * by - ay * bx;
}
__device__ __forceinline__ void solve33(const Jc& j, float bx, float by, float bz, float& ox, float& oy, float& oz) {
    float yzx, yzy, yzz;
    cross3(j.eyx, j.eyy, j.ezy, j.ezx, j.ezy, j.ezz, yzx, yzy, yzz);
    const float dp = dot3(j.exx, j.eyx, j.ezx, yzx, yzy, yzz);
    const float det = (dp != 0.0f) ? (1.0f / dp) : dp;
    ox = det * dot3(bx, by, bz, yzx, yzy, yzz);
    float bzx, bzy, bzz;
    cross3(bx, by, bz, j.ezx, j.ezy, j.ezz, bzx, bzy, bzz);
    oy = det * dot3(j.exx, j.eyx, j.ezx, bzx, bzy, bzz);
    float ybx, yby, ybz;
    cross3(j.eyx, j.eyy, j.ezy, bx, by, bz, ybx, yby, ybz);
    oz = det * dot3(j.exx, j.eyx, j.ezx, ybx, yby, ybz);
}
__device__ __forceinline__ V2 solve22(const Jc& j, V2 b) {
    const float xx = j.exx, xy = j.eyx, yx = j.eyx, yy = j.eyy;
    const float cp = xx * yy - xy * yx;
    const float det = (cp != 0.0f) ? (1.0f / cp) : cp;
    return V2{det * (yy * b.x - xy * b.y), det * (xx * b.y - yx * b.x)};
}

// Warm-start half of InitVelocity (RevoluteJointConf.cpp:199-224).
__device__ __forceinline__ void joint_warm_start(const Jc& j, float4& velA, float4& velB) {
    const V2 P{j.ix, j.iy};
    const float L = j.motorImp + j.iz;
    const float LA = cross(j.rA, P) / 1.0f + L;
    const float LB = cross(j.rB, P) / 1.0f + L;
    velA.x = velA.x - j.A.invMass * P.x;
    velA.y = velA.y - j.A.invMass * P.y;
    velA.z = velA.z - j.A.invRotI * LA;
    velB.x = velB.x + j.B.invMass * P.x;
    velB.y = velB.y + j.B.invMass * P.y;
    velB.z = velB.z + j.B.invRotI * LB;
}

// SolveVelocity (RevoluteJointConf.cpp:228-331).  Returns true if the bodies'
// velocities were left unchanged.
__device__ __forceinline__ bool joint_solve_velocity(Jc& j, float4& velA, float4& velB) {
    const float4 oldA = velA, oldB = velB;
    const float mA = j.A.invMass, iA = j.A.invRotI, mB = j.B.invMass, iB = j.B.invRotI;
    const bool fixedRotation = (iA + iB == 0.0f);
    if ((j.flags & prx::kJointMotor) && j.limit != 3u && !fixedRotation) {
        const float impulse = -j.angularMass * (velB.z - velA.z - j.motorSpeed);
        const float oldImpulse = j.motorImp;
        j.motorImp = clamp_std(j.motorImp + impulse, -j.maxImpulse, j.maxImpulse);
        const float inc = j.motorImp - oldImpulse;
        velA.z = velA.z - iA * inc;
        velB.z = velB.z + iB * inc;
    }
    const V2 vb = V2{velB.x, velB.y} + mul(revperp(j.rB), velB.z / 1.0f);
    const V2 va = V2{velA.x, velA.y} + mul(revperp(j.rA), velA.z / 1.0f);
    const V2 vDelta = vb - va;
    V2 P;
    float L = 0.0f;
    if ((j.flags & prx::kJointLimit) && j.limit != 0u && !fixedRotation) {
        float sx, sy, sz;
        solve33(j, vDelta.x / 1.0f, vDelta.y / 1.0f, (velB.z - velA.z) / 1.0f, sx, sy, sz);
        float ix = -sx, iy = -sy, iz = -sz;
        const bool lowerCase = j.limit == 1u, upperCase = j.limit == 2u;
        const float newZ = j.iz + iz;
        if ((lowerCase && newZ < 0.0f) || (upperCase && newZ > 0.0f)) {
            const V2 rhs = -V2{vDelta.x / 1.0f, vDelta.y / 1.0f} + mul(j.iz, V2{j.ezx, j.ezy});
            const V2 reduced = solve22(j, rhs);
            ix = reduced.x;
            iy = reduced.y;
            iz = -j.iz;
            j.ix = j.ix + reduced.x;
            j.iy = j.iy + reduced.y;
            j.iz = 0.0f;
        } else {
            j.ix = j.ix + ix;
            j.iy = j.iy + iy;
            j.iz = j.iz + iz;
        }
        P = V2{ix, iy};
        L = iz;
        const float LA = cross(j.rA, P) / 1.0f + L;
        const float LB = cross(j.rB, P) / 1.0f + L;
        velA.x = velA.x - mA * P.x;
        velA.y = velA.y - mA * P.y;
        velA.z = velA.z - iA * LA;
        velB.x = velB.x + mB * P.x;
        velB.y = velB.y + mB * P.y;
        velB.z = velB.z + iB * LB;
    } else {
        const V2 imp = solve22(j, -V2{vDelta.x / 1.0f, vDelta.y / 1.0f});
        j.ix = j.ix + imp.x;
        j.iy = j.iy + imp.y;
        P = imp;
        const float LA = cross(j.rA, P) / 1.0f;
        const float LB = cross(j.rB, P) / 1.0f;
        velA.x = velA.x - mA * P.x;
        velA.y = velA.y - mA * P.y;
        velA.z = velA.z - iA * LA;
        velB.x = velB.x + mB * P.x;
        velB.y = velB.y + mB * P.y;
        velB.z = velB.z + iB * LB;
    }
    const bool same = velA.x == oldA.x && velA.y == oldA.y && velA.z == oldA.z && velB.x == oldB.x &&
                      velB.y == oldB.y && velB.z == oldB.z;
    if (same) {
        velA = oldA;
        velB = oldB;
    }
    return same;
}

// SolvePosition (RevoluteJointConf.cpp:333-433).  Returns whether the joint
// is within tolerance.
__device__ __forceinline__ bool joint_solve_position(const float4* u, float4& posA, float4& posB, const StepParams& c) {
    const float4 n = u[2], ms = u[4], la = u[6], ang = u[7];
    const float angularMass = n.z;
    const std::uint32_t flags = __float_as_uint(n.w) & 0xffu, limit = __float_as_uint(n.w) >> 8;
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const bool fixedRotation = ((iA + iB) == 0.0f);
    float angularError = 0.0f;
    if ((flags & prx::kJointLimit) && limit != 0u && !fixedRotation) {
        const float angle = posB.z - posA.z - ang.x;
        float limitImpulse = 0.0f;
        if (limit == 1u) {
            float C = angle - ang.y;
            angularError = -C;
            C = clamp_std(C + c.angularSlop, -c.maxAngularCorrection, 0.0f);
            limitImpulse = -angularMass * C;
        } else if (limit == 2u) {
            float C = angle - ang.z;
            angularError = C;
            C = clamp_std(C - c.angularSlop, 0.0f, c.maxAngularCorrection);
            limitImpulse = -angularMass * C;
        } else {
            const float C = clamp_std(angle - ang.y, -c.maxAngularCorrection, c.maxAngularCorrection);
            limitImpulse = -angularMass * C;
            angularError = __builtin_fabsf(C);
        }
        posA.z = posA.z - iA * limitImpulse;
        posB.z = posB.z + iB * limitImpulse;
    }
    float positionError;
    {
        const V2 qA = unit_from_angle(posA.z);
        const V2 qB = unit_from_angle(posB.z);
        const V2 rA = rotate(V2{la.x, la.y}, qA);
        const V2 rB = rotate(V2{la.z, la.w}, qB);
        const V2 C = (V2{posB.x, posB.y} + rB) - (V2{posA.x, posA.y} + rA);
        positionError = magsq(C);
        const float exx = mA + (iA * (rA.y * rA.y) / 1.0f) + mB + (iB * (rB.y * rB.y) / 1.0f);
        const float exy = (-iA * rA.x * rA.y / 1.0f) + (-iB * rB.x * rB.y / 1.0f);
        const float eyy = mA + (iA * (rA.x * rA.x) / 1.0f) + mB + (iB * (rB.x * rB.x) / 1.0f);
        V2 sol{0.0f, 0.0f};
        const float cp = cross(V2{exx, exy}, V2{exy, eyy});
        if (!almost_zero(cp)) {
            const float inverse = 1.0f / cp;
            sol = V2{(eyy * C.x - exy * C.y) * inverse, (exx * C.y - exy * C.x) * inverse};
        }
        const V2 P = -sol;
        posA.x = posA.x - mA * P.x;
        posA.y = posA.y - mA * P.y;
        posA.z = posA.z - iA * cross(rA, P) / 1.0f;
        posB.x = posB.x + mB * P.x;
        posB.y = posB.y + mB * P.y;
        posB.z = posB.z + iB * cross(rB, P) / 1.0f;
    }
    return (positionError <= c.linearSlop * c.linearSlop) && (angularError <= c.angularSlop);
}

// ---------------------------------------------------------------------------
// Distance, weld, rope, friction, motor, wheel and pulley joints.  Their records share the revolute record's slots
// u[0] (rA, rB; a wheel holds its axes there), u[3] (impulses), u[4] (inverse masses), u[5].zw (bodies),
// u[6] (anchors less local centers); u[2].w holds flags | limit state << 8 |
// kind << 16 (the revolute record's word, kind 0), limit state 0 but for ropes.
//   distance: u[1] direction x, y, mass, invGamma   u[5].x bias
//             u[7] -, frequency, length, kind        impulse in u[3].x
//   weld:     u[1] mass exx, eyx, ezx, eyy           u[2] ezy, ezz, -, flags
//             u[5].xy gamma, bias                    u[7] reference angle, frequency, -, kind
__device__ __forceinline__ std::uint32_t jc_kind(const float4* u) { return __float_as_uint(u[2].w) >> 16; }

__device__ __forceinline__ void apply_joint_impulse(const float4 ms, V2 P, float LA, float LB, float4& velA,
                                                    float4& velB) {
    velA.x = velA.x - ms.x * P.x;
    velA.y = velA.y - ms.x * P.y;
    velA.z = velA.z - ms.y * LA;
    velB.x = velB.x + ms.z * P.x;
    velB.y = velB.y + ms.z * P.y;
    velB.z = velB.z + ms.w * LB;
}

// GetMat33 (WeldJointConf.cpp:33-60), the same matrix as the revolute one.
struct K33 {
    float exx, eyx, ezx, eyy, ezy, ezz;
};
__device__ __forceinline__ K33 joint_k33(float mA, V2 rA, float iA, float mB, V2 rB, float iB) {
    K33 k;
    k.exx = mA + ((rA.y * rA.y) * iA / 1.0f) + mB + ((rB.y * rB.y) * iB / 1.0f);
    k.eyx = (-rA.y * rA.x * iA / 1.0f) + (-rB.y * rB.x * iB / 1.0f);
    k.ezx = (-rA.y * iA) + (-rB.y * iB);
    k.eyy = mA + ((rA.x * rA.x) * iA / 1.0f) + mB + ((rB.x * rB.x) * iB / 1.0f);
    k.ezy = (rA.x * iA) + (rB.x * iB);
    k.ezz = iA + iB;
    return k;
}

// InitVelocity(DistanceJointConf&) less its warm start (DistanceJointConf.cpp:81-141).
__device__ __forceinline__ void build_jc_distance(float4* u, const float4 cf, float4 im, V2 rA, V2 rB, const float4& pA,
                                                  const float4& pB, float mA, float iA, float mB, float iB,
                                                  const StepParams& c) {
    float length;
    const V2 dir = unit_polar((V2{pB.x, pB.y} + rB) - (V2{pA.x, pA.y} + rA), length);
    const float crAu = cross(rA, dir) / 1.0f;
    const float crBu = cross(rB, dir) / 1.0f;
    float invMass = mA + iA * (crAu * crAu) + mB + iB * (crBu * crBu);
    float mass = (invMass != 0.0f) ? 1.0f / invMass : 0.0f;
    float invGamma = 0.0f, bias = 0.0f;
    if (cf.y > 0.0f) {
        const float C = length - cf.x;
        const float omega = 2.0f * prx::kPiF * cf.y;
        const float d = 2.0f * mass * cf.z * omega;
        const float k = mass * (omega * omega);
        const float h = c.dt;
        const float gamma = h * (d + h * k);
        invGamma = (gamma != 0.0f) ? 1.0f / gamma : 0.0f;
        bias = C * h * k * invGamma;
        invMass += invGamma;
        mass = (invMass != 0.0f) ? 1.0f / invMass : 0.0f;
    }
    im = c.warmStart ? float4{im.x * c.dtRatio, 0.0f, 0.0f, 0.0f} : float4{0.0f, 0.0f, 0.0f, 0.0f};
    u[1] = float4{dir.x, dir.y, mass, invGamma};
    u[3] = im;
    u[5].x = bias;
    u[5].y = 0.0f;
    u[7] = float4{0.0f, cf.y, cf.x, __uint_as_float(prx::kJointDistance)};
}

// InitVelocity(WeldJointConf&) less its warm start (WeldJointConf.cpp:115-195).
__device__ __forceinline__ void build_jc_weld(float4* u, const float4 cf, float4 im, V2 rA, V2 rB, const float4& pA,
                                              const float4& pB, float mA, float iA, float mB, float iB,
                                              std::uint32_t flags, const StepParams& c) {
    const K33 K = joint_k33(mA, rA, iA, mB, rB, iB);
    K33 m;
    float gamma = 0.0f, bias = 0.0f;
    if (cf.y > 0.0f || K.ezz == 0.0f) {
        // GetInverse22 (Math.hpp:478-490); K is symmetric.
        float det = (K.exx * K.eyy) - (K.eyx * K.eyx);
        if (det != 0.0f) det = 1.0f / det;
        m = K33{det * K.eyy, -det * K.eyx, 0.0f, det * K.exx, 0.0f, 0.0f};
        if (cf.y > 0.0f) {
            float invRotI = iA + iB;
            const float rotI = (invRotI > 0.0f) ? 1.0f / invRotI : 0.0f;
            const float C = pB.z - pA.z - cf.x;
            const float omega = 2.0f * prx::kPiF * cf.y;
            const float d = 2.0f * rotI * cf.z * omega;
            const float k = rotI * omega * omega;
            const float h = c.dt;
            const float invGamma = h * (d + h * k);
            gamma = (invGamma != 0.0f) ? 1.0f / invGamma : 0.0f;
            bias = C * h * k * gamma;
            invRotI += gamma;
            m.ezz = (invRotI != 0.0f) ? 1.0f / invRotI : 0.0f;
        }
    } else {
        // GetSymInverse33 (Math.hpp:493-510).
        float yzx, yzy, yzz;
        cross3(K.eyx, K.eyy, K.ezy, K.ezx, K.ezy, K.ezz, yzx, yzy, yzz);
        const float invDet = dot3(K.exx, K.eyx, K.ezx, yzx, yzy, yzz);
        const float det = (invDet != 0.0f) ? (1.0f / invDet) : invDet;
        const float a11 = K.exx, a12 = K.eyx, a13 = K.ezx, a22 = K.eyy, a23 = K.ezy, a33 = K.ezz;
        m.exx = det * (a22 * a33 - a23 * a23);
        m.eyx = det * (a13 * a23 - a12 * a33);
        m.ezx = det * (a12 * a23 - a13 * a22);
        m.eyy = det * (a11 * a33 - a13 * a13);
        m.ezy = det * (a13 * a12 - a11 * a23);
        m.ezz = det * (a11 * a22 - a12 * a12);
    }
    im = c.warmStart ? float4{im.x * c.dtRatio, im.y * c.dtRatio, im.z * c.dtRatio, 0.0f}
                     : float4{0.0f, 0.0f, 0.0f, 0.0f};
    u[1] = float4{m.exx, m.eyx, m.ezx, m.eyy};
    u[2] = float4{m.ezy, m.ezz, 0.0f, __uint_as_float(flags | (prx::kJointWeld << 16))};
    u[3] = im;
    u[5].x = gamma;
    u[5].y = bias;
    u[7] = float4{cf.x, cf.y, 0.0f, __uint_as_float(prx::kJointWeld)};
}

// Rope record: u[1] direction x, y, mass, length; u[2].w flags | limit << 8 | kind << 16;
// u[5].x 1 / dt, u[5].y 1 when too short to act; u[7].z maximum length;
// impulse in u[3].x.
// InitVelocity(RopeJointConf&) less its warm start (RopeJointConf.cpp:59-113).
__device__ __forceinline__ void build_jc_rope(float4* u, const float4 cf, float4 im, V2 rA, V2 rB, const float4& pA,
                                              const float4& pB, float mA, float iA, float mB, float iB,
                                              std::uint32_t flags, const StepParams& c) {
    float length;
    const V2 uv = unit_polar((V2{pB.x, pB.y} + rB) - (V2{pA.x, pA.y} + rA), length);
    const float C = length - cf.x;
    const std::uint32_t limit = (C > 0.0f) ? 2u : 0u;
    V2 dir{0.0f, 0.0f};
    float mass = 0.0f;
    if (length > c.linearSlop) {
        dir = uv;
        const float crA = cross(rA, dir) / 1.0f;
        const float crB = cross(rB, dir) / 1.0f;
        const float invMass = mA + mB + iA * (crA * crA) + iB * (crB * crB);
        mass = (invMass != 0.0f) ? 1.0f / invMass : 0.0f;
        im = c.warmStart ? float4{im.x * c.dtRatio, 0.0f, 0.0f, 0.0f} : float4{0.0f, 0.0f, 0.0f, 0.0f};
    } else {
        im = float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    u[1] = float4{dir.x, dir.y, mass, length};
    u[2] = float4{0.0f, 0.0f, 0.0f, __uint_as_float(flags | (limit << 8) | (prx::kJointRope << 16))};
    u[3] = im;
    u[5].x = (c.dt != 0.0f) ? 1.0f / c.dt : 0.0f;
    u[5].y = (length > c.linearSlop) ? 0.0f : 1.0f;  // 1: InitVelocity returned early, no warm start
    u[7] = float4{0.0f, 0.0f, cf.x, __uint_as_float(prx::kJointRope)};
}

// Friction record: u[1] linear mass (Invert of the 2x2 K) m00, m01, m10, m11;
// u[2].x angular mass; u[5].xy dt * maximum force, dt * maximum torque;
// impulses linear x, y, angular in u[3].
// InitVelocity(FrictionJointConf&) less its warm start (FrictionJointConf.cpp:76-123).
__device__ __forceinline__ void build_jc_friction(float4* u, const float4 cf, float4 im, V2 rA, V2 rB, float mA,
                                                  float iA, float mB, float iB, std::uint32_t flags,
                                                  const StepParams& c) {
    const float exx = mA + iA * (rA.y * rA.y) / 1.0f + mB + iB * (rB.y * rB.y) / 1.0f;
    const float exy = -iA * rA.x * rA.y / 1.0f + -iB * rB.x * rB.y / 1.0f;
    const float eyy = mA + iA * (rA.x * rA.x) / 1.0f + mB + iB * (rB.x * rB.x) / 1.0f;
    // Invert(Matrix22) (Math.hpp:429-442).
    float4 lm{0.0f, 0.0f, 0.0f, 0.0f};
    const float cp = cross(V2{exx, exy}, V2{exy, eyy});
    if (!almost_zero(cp)) {
        const float inverse = 1.0f / cp;
        lm = float4{eyy * inverse, -exy * inverse, -exy * inverse, exx * inverse};
    }
    const float invRotI = iA + iB;
    const float angularMass = (invRotI > 0.0f) ? 1.0f / invRotI : 0.0f;
    im = c.warmStart ? float4{im.x * c.dtRatio, im.y * c.dtRatio, im.z * c.dtRatio, 0.0f}
                     : float4{0.0f, 0.0f, 0.0f, 0.0f};
    u[1] = lm;
    u[2] = float4{angularMass, 0.0f, 0.0f, __uint_as_float(flags | (prx::kJointFriction << 16))};
    u[3] = im;
    u[5].x = c.dt * cf.x;
    u[5].y = c.dt * cf.y;
    u[7] = float4{0.0f, 0.0f, 0.0f, __uint_as_float(prx::kJointFriction)};
}

// Wheel record: u[0] ay x, y, ax x, y; u[1] sAy, sBy, sAx, sBx; u[2].xyz
// point-to-line mass, spring mass, angular (motor) mass; u[3] impulse, spring
// impulse, angular impulse; u[5].xy bias, gamma; u[7] motor speed,
// dt * maximum motor torque, local x axis of A.
// InitVelocity(WheelJointConf&) less its warm start (WheelJointConf.cpp:105-211).
__device__ __forceinline__ void build_jc_wheel(float4* u, const float4 cf, const float4 cf2, float4 im, V2 rA, V2 rB,
                                               V2 qA, const float4& pA, const float4& pB, float mA, float iA,
                                               float mB, float iB, std::uint32_t flags, const StepParams& c) {
    const V2 axis{cf2.z, cf2.w};
    const V2 dd = (V2{pB.x, pB.y} + rB) - (V2{pA.x, pA.y} + rA);
    const V2 ay = rotate(revperp(axis), qA);
    const float sAy = cross(dd + rA, ay);
    const float sBy = cross(rB, ay);
    const float invMass = mA + mB + iA * (sAy * sAy) / 1.0f + iB * (sBy * sBy) / 1.0f;
    const float mass = (invMass > 0.0f) ? 1.0f / invMass : 0.0f;
    float springMass = 0.0f, bias = 0.0f, gamma = 0.0f, sAx = 0.0f, sBx = 0.0f;
    V2 ax{0.0f, 0.0f};
    if (cf.x > 0.0f) {
        ax = rotate(axis, qA);
        sAx = cross(dd + rA, ax);
        sBx = cross(rB, ax);
        const float invMass2 = mA + mB + iA * (sAx * sAx) / 1.0f + iB * (sBx * sBx) / 1.0f;
        if (invMass2 > 0.0f) {
            springMass = 1.0f / invMass2;
            const float C = dot(dd, ax);
            const float omega = 2.0f * prx::kPiF * cf.x;
            const float d = 2.0f * springMass * cf.y * omega;
            const float k = springMass * omega * omega;
            const float h = c.dt;
            const float invGamma = h * (d + h * k);
            gamma = (invGamma > 0.0f) ? 1.0f / invGamma : 0.0f;
            bias = C * h * k * gamma;
            const float total = invMass2 + gamma;
            springMass = (total > 0.0f) ? 1.0f / total : 0.0f;
        }
    } else {
        im.y = 0.0f;
    }
    float angularMass = 0.0f;
    if (flags & prx::kJointMotor) {
        const float invRotI = iA + iB;
        angularMass = (invRotI > 0.0f) ? 1.0f / invRotI : 0.0f;
    } else {
        im.z = 0.0f;
    }
    im = c.warmStart ? float4{im.x * c.dtRatio, im.y * c.dtRatio, im.z * c.dtRatio, 0.0f}
                     : float4{0.0f, 0.0f, 0.0f, 0.0f};
    u[0] = float4{ay.x, ay.y, ax.x, ax.y};
    u[1] = float4{sAy, sBy, sAx, sBx};
    u[2] = float4{mass, springMass, angularMass, __uint_as_float(flags | (prx::kJointWheel << 16))};
    u[3] = im;
    u[5].x = bias;
    u[5].y = gamma;
    u[7] = float4{cf.w, c.dt * cf2.x, axis.x, axis.y};
}

// SolveVelocity(WheelJointConf&) (WheelJointConf.cpp:232-307): returns whether
// the bodies' velocities were left unchanged.
__device__ __forceinline__ bool wheel_solve_velocity(float4* u, float4& velA, float4& velB) {
    const float4 a = u[0], s = u[1], n = u[2], ms = u[4], gb = u[5], mo = u[7];
    const V2 ay{a.x, a.y}, ax{a.z, a.w};
    const float4 oldA = velA, oldB = velB;
    float4 im = u[3];
    {
        const float dv = dot(ax, V2{velB.x, velB.y} - V2{velA.x, velA.y});
        const float Cdot = dv + s.w * velB.z / 1.0f - s.z * velA.z / 1.0f;
        const float impulse = -n.y * (Cdot + gb.x + gb.y * im.y);
        im.y = im.y + impulse;
        apply_joint_impulse(ms, mul(impulse, ax), impulse * s.z / 1.0f, impulse * s.w / 1.0f, velA, velB);
    }
    {
        const float Cdot = velB.z - velA.z - mo.x;
        float impulse = -n.z * Cdot;
        const float old = im.z;
        im.z = clamp_std(im.z + impulse, -mo.y, mo.y);
        impulse = im.z - old;
        velA.z = velA.z - ms.y * impulse;
        velB.z = velB.z + ms.w * impulse;
    }
    {
        const float dv = dot(ay, V2{velB.x, velB.y} - V2{velA.x, velA.y});
        const float Cdot = dv + s.y * velB.z / 1.0f - s.x * velA.z / 1.0f;
        const float impulse = -n.x * Cdot;
        im.x = im.x + impulse;
        apply_joint_impulse(ms, mul(impulse, ay), impulse * s.x / 1.0f, impulse * s.y / 1.0f, velA, velB);
    }
    u[3] = im;
    const bool same = velA.x == oldA.x && velA.y == oldA.y && velA.z == oldA.z && velB.x == oldB.x &&
                      velB.y == oldB.y && velB.z == oldB.z;
    if (same) {
        velA = oldA;
        velB = oldB;
    }
    return same;
}

// Target record (body A is body B: its copy is left unchanged and written back
// first): u[0] rB x, y, C x, y; u[1] effective mass m00, m01, m10, m11;
// u[2].xyz gamma, dt * maximum force, 1 when warm starting; impulse in
// u[3].xy.  InitVelocity(TargetJointConf&) less its velocity changes
// (TargetJointConf.cpp:85-132, GetEffectiveMassMatrix :60-83).
__device__ __forceinline__ void build_jc_target(float4* u, const float4 cf, const float4 cf2, float4 im, V2 rB,
                                                const float4& pB, float mB, float iB, std::uint32_t flags,
                                                const StepParams& c) {
    const float mass = (mB != 0.0f) ? (1.0f / mB) : 0.0f;
    const float omega = 2.0f * prx::kPiF * cf.w;
    const float dd = 2.0f * mass * cf2.x * omega;
    const float k = mass * (omega * omega);
    const float h = c.dt;
    const float tmp = dd + h * k;
    const float invGamma = h * tmp;
    const float gamma = (invGamma != 0.0f) ? 1.0f / invGamma : 0.0f;
    const float beta = h * k * gamma;
    const float exx = mB + (iB * (rB.y * rB.y) / 1.0f) + gamma;
    const float exy = -iB * rB.x * rB.y / 1.0f;
    const float eyy = mB + (iB * (rB.x * rB.x) / 1.0f) + gamma;
    // Invert(Matrix22) (Math.hpp:429-442) of columns (exx, exy), (exy, eyy).
    const float cp = cross(V2{exx, exy}, V2{exy, eyy});
    float4 m{0.0f, 0.0f, 0.0f, 0.0f};
    if (!almost_zero(cp)) {
        const float inverse = 1.0f / cp;
        m = float4{eyy * inverse, -exy * inverse, -exy * inverse, exx * inverse};
    }
    const V2 C = mul((V2{pB.x, pB.y} + rB) - V2{cf.x, cf.y}, beta);
    im = c.warmStart ? float4{im.x * c.dtRatio, im.y * c.dtRatio, 0.0f, 0.0f} : float4{0.0f, 0.0f, 0.0f, 0.0f};
    u[0] = float4{rB.x, rB.y, C.x, C.y};
    u[1] = m;
    u[2] = float4{gamma, c.dt * cf.z, c.warmStart ? 1.0f : 0.0f,
                  __uint_as_float(flags | (prx::kJointTarget << 16))};
    u[3] = im;
}

// The velocity half of InitVelocity(TargetJointConf&) (TargetJointConf.cpp:
// 125-140): the spin damping, then the warm start when warm starting.
__device__ __forceinline__ void target_warm_start(const float4* u, float4& velB) {
    const float4 r = u[0], im = u[3], ms = u[4];
    velB.z = velB.z * 0.98f;
    if (u[2].z == 0.0f) return;
    const V2 P{im.x, im.y};
    const float crossBP = cross(V2{r.x, r.y}, P) / 1.0f;
    velB.x = velB.x + ms.z * P.x;
    velB.y = velB.y + ms.z * P.y;
    velB.z = velB.z + ms.w * crossBP;
}

// SolveVelocity(TargetJointConf&) (TargetJointConf.cpp:143-178): returns
// whether the impulse increment was zero.
__device__ __forceinline__ bool target_solve_velocity(float4* u, float4& velB) {
    const float4 r = u[0], m = u[1], g = u[2], ms = u[4];
    const V2 rB{r.x, r.y}, C{r.z, r.w};
    float4 im = u[3];
    const V2 Cdot = V2{velB.x, velB.y} + mul(revperp(rB), velB.z / 1.0f);
    const V2 ev = Cdot + (C + mul(g.x, V2{im.x, im.y}));
    const V2 old{im.x, im.y};
    // Transform(-ev, mass): the matrix's vectors dotted with -ev from zero.
    const V2 nev = -ev;
    float tx = 0.0f, ty = 0.0f;
    tx += m.x * nev.x;
    tx += m.y * nev.y;
    ty += m.z * nev.x;
    ty += m.w * nev.y;
    V2 imp{old.x + tx, old.y + ty};
    const float maxImpulse = g.y;
    if (magsq(imp) > maxImpulse * maxImpulse) {
        const V2 dir = unit_vector(imp, V2{0.0f, 0.0f});
        imp = V2{dir.x * maxImpulse, dir.y * maxImpulse};
    }
    im.x = imp.x;
    im.y = imp.y;
    u[3] = im;
    const V2 inc = imp - old;
    const float angImpulseB = cross(rB, inc) / 1.0f;
    velB.x = velB.x + ms.z * inc.x;
    velB.y = velB.y + ms.z * inc.y;
    velB.z = velB.z + ms.w * angImpulseB;
    return inc.x == 0.0f && inc.y == 0.0f;
}

// Prismatic record: u[0] axis x, y, perp x, y; u[1] s1, s2, a1, a2; u[2].xy
// motor speed, dt * maximum motor force, u[2].w flags | limit state << 8 |
// kind << 16; u[3] impulse x, y, z, motor impulse; u[5].x upper translation;
// u[7] local x axis of A, reference angle, lower translation.  K and the motor
// mass are recomputed from u[1] and u[4] where used, by InitVelocity's own
// operations (PrismaticJointConf.cpp:196-230).
struct PrisK {
    float k11, k12, k13, k22, k23, k33;
};
__device__ __forceinline__ PrisK prismatic_k(float4 ms, float s1, float s2, float a1, float a2) {
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    const float invRotMassA = iA * (a1 * a1) / 1.0f;
    const float invRotMassB = iB * (a2 * a2) / 1.0f;
    const float invRotMassA2 = iA * (s1 * s1) / 1.0f;
    const float invRotMassB2 = iB * (s2 * s2) / 1.0f;
    PrisK k;
    k.k11 = mA + mB + invRotMassA2 + invRotMassB2;
    k.k12 = iA * s1 + iB * s2;
    k.k13 = (iA * s1 * a1 + iB * s2 * a2) / 1.0f;
    const float tot = iA + iB;
    k.k22 = (tot == 0.0f) ? 1.0f : tot;
    k.k23 = iA * a1 + iB * a2;
    k.k33 = mA + mB + invRotMassA + invRotMassB;
    return k;
}
__device__ __forceinline__ Jc prismatic_jc(const PrisK& k) {
    Jc j;
    j.exx = k.k11;
    j.eyx = k.k12;
    j.ezx = k.k13;
    j.eyy = k.k22;
    j.ezy = k.k23;
    j.ezz = k.k33;
    return j;
}

// InitVelocity(PrismaticJointConf&) less its warm start (PrismaticJointConf.cpp:155-264).
__device__ __forceinline__ void build_jc_prismatic(float4* u, const float4 cf, const float4 cf2, float4 im, V2 rA,
                                                   V2 rB, V2 qA, const float4& pA, const float4& pB,
                                                   std::uint32_t flags, std::uint32_t limit, const StepParams& c) {
    const V2 axisA{cf2.z, cf2.w};
    const V2 d = ((V2{pB.x, pB.y} - V2{pA.x, pA.y}) + rB) - rA;
    const V2 axis = rotate(axisA, qA);
    const float a1 = cross(d + rA, axis);
    const float a2 = cross(rB, axis);
    const V2 perp = rotate(revperp(axisA), qA);
    const float s1 = cross(d + rA, perp);
    const float s2 = cross(rB, perp);
    if (flags & prx::kJointLimit) {
        const float translation = dot(axis, d);
        if (__builtin_fabsf(cf.z - cf.y) < (c.linearSlop * 2.0f)) {
            limit = 3u;
        } else if (translation <= cf.y) {
            if (limit != 1u) {
                limit = 1u;
                im.z = 0.0f;
            }
        } else if (translation >= cf.z) {
            if (limit != 2u) {
                limit = 2u;
                im.z = 0.0f;
            }
        } else {
            limit = 0u;
            im.z = 0.0f;
        }
    } else {
        limit = 0u;
        im.z = 0.0f;
    }
    if (!(flags & prx::kJointMotor)) im.w = 0.0f;
    im = c.warmStart ? float4{im.x * c.dtRatio, im.y * c.dtRatio, im.z * c.dtRatio, im.w * c.dtRatio}
                     : float4{0.0f, 0.0f, 0.0f, 0.0f};
    u[0] = float4{axis.x, axis.y, perp.x, perp.y};
    u[1] = float4{s1, s2, a1, a2};
    u[2] = float4{cf.w, c.dt * cf2.x, 0.0f, __uint_as_float(flags | (limit << 8) | (prx::kJointPrismatic << 16))};
    u[3] = im;
    u[5].x = cf.z;
    u[7] = float4{axisA.x, axisA.y, cf.x, cf.y};
}

// Warm start of InitVelocity (PrismaticJointConf.cpp:241-262).
__device__ __forceinline__ void prismatic_warm_start(const float4* u, float4& velA, float4& velB) {
    const float4 a = u[0], s = u[1], im = u[3];
    const V2 axis{a.x, a.y}, perp{a.z, a.w};
    const V2 Px = mul(im.x, perp);
    const float Pxs1 = im.x * s.x;
    const float Pxs2 = im.x * s.y;
    const float PzLength = im.w + im.z;
    const V2 Pz = mul(PzLength, axis);
    const V2 P = Px + Pz;
    const float L = im.y;
    const float LA = L + (Pxs1 + PzLength * s.z);
    const float LB = L + (Pxs2 + PzLength * s.w);
    apply_joint_impulse(u[4], P, LA, LB, velA, velB);
}

// SolveVelocity(PrismaticJointConf&) (PrismaticJointConf.cpp:270-392):
// returns whether the bodies' velocities were left unchanged.
__device__ __forceinline__ bool prismatic_solve_velocity(float4* u, float4& velA, float4& velB) {
    const float4 a = u[0], s = u[1], mo = u[2], ms = u[4];
    const V2 axis{a.x, a.y}, perp{a.z, a.w};
    const float s1 = s.x, s2 = s.y, a1 = s.z, a2 = s.w;
    const std::uint32_t word = __float_as_uint(mo.w);
    const std::uint32_t flags = word & 0xffu, limit = (word >> 8) & 0xffu;
    const PrisK K = prismatic_k(ms, s1, s2, a1, a2);
    const Jc kj = prismatic_jc(K);
    const float4 oldA = velA, oldB = velB;
    float4 im = u[3];
    if ((flags & prx::kJointMotor) && limit != 3u) {
        const float vDot = dot(axis, V2{velB.x, velB.y} - V2{velA.x, velA.y});
        const float Cdot = vDot + (a2 * velB.z - a1 * velA.z);
        const float motorMass = (K.k33 > 0.0f) ? 1.0f / K.k33 : 0.0f;
        float impulse = motorMass * (mo.x - Cdot);
        const float old = im.w;
        im.w = clamp_std(im.w + impulse, -mo.y, mo.y);
        impulse = im.w - old;
        apply_joint_impulse(ms, mul(impulse, axis), impulse * a1, impulse * a2, velA, velB);
    }
    const V2 velDelta = V2{velB.x, velB.y} - V2{velA.x, velA.y};
    const float sRotSpeed = (s2 * velB.z - s1 * velA.z);
    const V2 Cdot1{dot(perp, velDelta) + sRotSpeed, velB.z - velA.z};
    if ((flags & prx::kJointLimit) && limit != 0u) {
        const float deltaDot = dot(axis, velDelta);
        const float aRotSpeed = (a2 * velB.z - a1 * velA.z);
        const float Cdot2 = deltaDot + aRotSpeed;
        const float4 f1 = im;
        float sx, sy, sz;
        solve33(kj, -Cdot1.x, -Cdot1.y, -Cdot2, sx, sy, sz);
        im.x = im.x + sx;
        im.y = im.y + sy;
        im.z = im.z + sz;
        if (limit == 1u) im.z = fmax_std(im.z, 0.0f);
        else if (limit == 2u) im.z = fmin_std(im.z, 0.0f);
        const float dz = im.z - f1.z;
        const V2 b{-Cdot1.x - dz * K.k13, -Cdot1.y - dz * K.k23};
        const V2 f2r = solve22(kj, b) + V2{f1.x, f1.y};
        im.x = f2r.x;
        im.y = f2r.y;
        const float dfx = im.x - f1.x, dfy = im.y - f1.y, dfz = im.z - f1.z;
        const V2 P = mul(dfx, perp) + mul(dfz, axis);
        const float LA = (dfx * s1 + dfy + dfz * a1);
        const float LB = (dfx * s2 + dfy + dfz * a2);
        apply_joint_impulse(ms, P, LA, LB, velA, velB);
    } else {
        const V2 df = solve22(kj, V2{-Cdot1.x, -Cdot1.y});
        im.x = im.x + df.x;
        im.y = im.y + df.y;
        const V2 P = mul(df.x, perp);
        const float LA = (df.x * s1 + df.y);
        const float LB = (df.x * s2 + df.y);
        apply_joint_impulse(ms, P, LA, LB, velA, velB);
    }
    u[3] = im;
    const bool same = velA.x == oldA.x && velA.y == oldA.y && velA.z == oldA.z && velB.x == oldB.x &&
                      velB.y == oldB.y && velB.z == oldB.z;
    if (same) {
        velA = oldA;
        velB = oldB;
    }
    return same;
}

// Pulley record: u[1] uA x, y, uB x, y; u[2].xy mass, ratio; u[5].xy lengths
// A, B; u[7] ground anchors A, B; impulse in u[3].x.
// InitVelocity(PulleyJointConf&) less its warm start (PulleyJointConf.cpp:92-141).
__device__ __forceinline__ void build_jc_pulley(float4* u, const float4 cf, const float4 cf2, float4 im, V2 rA, V2 rB,
                                                const float4& pA, const float4& pB, float mA, float iA, float mB,
                                                float iB, std::uint32_t flags, const StepParams& c) {
    const float ratio = cf2.x;
    const V2 uA = unit_vector((V2{pA.x, pA.y} + rA) - V2{cf.x, cf.y}, V2{0.0f, 0.0f});
    const V2 uB = unit_vector((V2{pB.x, pB.y} + rB) - V2{cf.z, cf.w}, V2{0.0f, 0.0f});
    const float ruA = cross(rA, uA);
    const float ruB = cross(rB, uB);
    const float tA = mA + (iA * (ruA * ruA)) / 1.0f;
    const float tB = mB + (iB * (ruB * ruB)) / 1.0f;
    const float total = tA + ratio * ratio * tB;
    const float mass = (total > 0.0f) ? 1.0f / total : 0.0f;
    im = c.warmStart ? float4{im.x * c.dtRatio, 0.0f, 0.0f, 0.0f} : float4{0.0f, 0.0f, 0.0f, 0.0f};
    u[1] = float4{uA.x, uA.y, uB.x, uB.y};
    u[2] = float4{mass, ratio, 0.0f, __uint_as_float(flags | (prx::kJointPulley << 16))};
    u[3] = im;
    u[5].x = cf2.z;
    u[5].y = cf2.w;
    u[7] = cf;
}

// Applies a pulley impulse: both bodies gain (PulleyJointConf.cpp:135-140,
// 178-186).
__device__ __forceinline__ void pulley_apply(const float4* u, float impulse, float4& velA, float4& velB) {
    const float4 r = u[0], uv = u[1], ms = u[4];
    const float ratio = u[2].y;
    const V2 PA = mul(-impulse, V2{uv.x, uv.y});
    const V2 PB = mul(-ratio * impulse, V2{uv.z, uv.w});
    velA.x = velA.x + ms.x * PA.x;
    velA.y = velA.y + ms.x * PA.y;
    velA.z = velA.z + ms.y * cross(V2{r.x, r.y}, PA) / 1.0f;
    velB.x = velB.x + ms.z * PB.x;
    velB.y = velB.y + ms.z * PB.y;
    velB.z = velB.z + ms.w * cross(V2{r.z, r.w}, PB) / 1.0f;
}

// SolveVelocity(PulleyJointConf&) (PulleyJointConf.cpp:155-190): returns
// whether the impulse was zero.
__device__ __forceinline__ bool pulley_solve_velocity(float4* u, float4& velA, float4& velB) {
    const float4 r = u[0], uv = u[1], n = u[2];
    const V2 vpA = V2{velA.x, velA.y} + mul(revperp(V2{r.x, r.y}), velA.z / 1.0f);
    const V2 vpB = V2{velB.x, velB.y} + mul(revperp(V2{r.z, r.w}), velB.z / 1.0f);
    const float Cdot = -dot(V2{uv.x, uv.y}, vpA) - n.y * dot(V2{uv.z, uv.w}, vpB);
    const float impulse = -n.x * Cdot;
    u[3].x = u[3].x + impulse;
    pulley_apply(u, impulse, velA, velB);
    return impulse == 0.0f;
}

// Motor record: u[1] linear mass m00, m01, m10, m11; u[2].xy angular mass,
// angular error; u[5].xy dt * maximum force, dt * maximum torque; u[7] linear
// error x, y, (1 / dt) * correction factor; impulses as friction.
// InitVelocity(MotorJointConf&) less its warm start (MotorJointConf.cpp:79-145).
__device__ __forceinline__ void build_jc_motor(float4* u, const float4 cf, float4 im, V2 rA, V2 rB, const float4& pA,
                                               const float4& pB, float mA, float iA, float mB, float iB,
                                               std::uint32_t flags, const StepParams& c) {
    const float exx = mA + mB + iA * (rA.y * rA.y) / 1.0f + iB * (rB.y * rB.y) / 1.0f;
    const float exy = -iA * rA.x * rA.y / 1.0f + -iB * rB.x * rB.y / 1.0f;
    const float eyy = mA + mB + iA * (rA.x * rA.x) / 1.0f + iB * (rB.x * rB.x) / 1.0f;
    float4 lm{0.0f, 0.0f, 0.0f, 0.0f};
    const float cp = cross(V2{exx, exy}, V2{exy, eyy});
    if (!almost_zero(cp)) {
        const float inverse = 1.0f / cp;
        lm = float4{eyy * inverse, -exy * inverse, -exy * inverse, exx * inverse};
    }
    const float invRotI = iA + iB;
    const float angularMass = (invRotI > 0.0f) ? 1.0f / invRotI : 0.0f;
    const V2 linearError = (V2{pB.x, pB.y} + rB) - (V2{pA.x, pA.y} + rA);
    const float angularError = (pB.z - pA.z) - cf.x;
    im = c.warmStart ? float4{im.x * c.dtRatio, im.y * c.dtRatio, im.z * c.dtRatio, 0.0f}
                     : float4{0.0f, 0.0f, 0.0f, 0.0f};
    const float invH = (c.dt != 0.0f) ? 1.0f / c.dt : 0.0f;
    u[1] = lm;
    u[2] = float4{angularMass, angularError, 0.0f, __uint_as_float(flags | (prx::kJointMotorKind << 16))};
    u[3] = im;
    u[5].x = c.dt * cf.y;
    u[5].y = c.dt * cf.z;
    u[7] = float4{linearError.x, linearError.y, invH * cf.w, __uint_as_float(prx::kJointMotorKind)};
}

// SolveVelocity(MotorJointConf&) (MotorJointConf.cpp:160-229): returns
// whether neither impulse changed (velocities are then left as they were).
__device__ __forceinline__ bool motor_solve_velocity(float4* u, float4& velA, float4& velB) {
    const float4 r = u[0], lm = u[1], n = u[2], ms = u[4], mx = u[5], e = u[7];
    const V2 rA{r.x, r.y}, rB{r.z, r.w};
    float4 im = u[3];
    bool solved = true;
    {
        const float Cdot = (velB.z - velA.z) + e.z * n.y;
        const float angularImpulse = -n.x * Cdot;
        const float old = im.z;
        const float now = clamp_std(old + angularImpulse, -mx.y, mx.y);
        im.z = now;
        const float inc = now - old;
        if (inc != 0.0f) {
            // Velocity{LinearVelocity2{}, ...}: the linear parts take a zero.
            velA.x = velA.x - 0.0f;
            velA.y = velA.y - 0.0f;
            velA.z = velA.z - ms.y * inc;
            velB.x = velB.x + 0.0f;
            velB.y = velB.y + 0.0f;
            velB.z = velB.z + ms.w * inc;
            solved = false;
        }
    }
    {
        const V2 vb = V2{velB.x, velB.y} + mul(revperp(rB), velB.z / 1.0f);
        const V2 va = V2{velA.x, velA.y} + mul(revperp(rA), velA.z / 1.0f);
        const V2 Cdot = (vb - va) + V2{e.z * e.x, e.z * e.y};
        float tx = 0.0f, ty = 0.0f;
        tx += lm.x * Cdot.x;
        tx += lm.y * Cdot.y;
        ty += lm.z * Cdot.x;
        ty += lm.w * Cdot.y;
        const V2 old{im.x, im.y};
        V2 li{old.x + -tx, old.y + -ty};
        const float maxImpulse = mx.x;
        if (magsq(li) > maxImpulse * maxImpulse) {
            const V2 dir = unit_vector(li, V2{0.0f, 0.0f});
            li = V2{dir.x * maxImpulse, dir.y * maxImpulse};
        }
        im.x = li.x;
        im.y = li.y;
        const V2 inc = li - old;
        const float angA = cross(rA, inc) / 1.0f;
        const float angB = cross(rB, inc) / 1.0f;
        if (inc.x != 0.0f || inc.y != 0.0f || angA != 0.0f || angB != 0.0f) {
            apply_joint_impulse(ms, inc, angA, angB, velA, velB);
            solved = false;
        }
    }
    u[3] = im;
    return solved;
}

// SolveVelocity(RopeJointConf&) (RopeJointConf.cpp:130-167): returns whether
// the impulse did not change.
__device__ __forceinline__ bool rope_solve_velocity(float4* u, float4& velA, float4& velB) {
    const float4 r = u[0], m = u[1], ms = u[4];
    const V2 rA{r.x, r.y}, rB{r.z, r.w}, dir{m.x, m.y};
    const V2 vpA = V2{velA.x, velA.y} + mul(revperp(rA), velA.z / 1.0f);
    const V2 vpB = V2{velB.x, velB.y} + mul(revperp(rB), velB.z / 1.0f);
    const float C = m.w - u[7].z;
    const float Cdot = dot(dir, vpB - vpA) + ((C < 0.0f) ? u[5].x * C : 0.0f);
    float local = -m.z * Cdot;
    const float old = u[3].x;
    const float acc = fmin_std(0.0f, old + local);
    u[3].x = acc;
    local = acc - old;
    const V2 P = mul(local, dir);
    apply_joint_impulse(ms, P, cross(rA, P) / 1.0f, cross(rB, P) / 1.0f, velA, velB);
    return local == 0.0f;
}

// SolveVelocity(FrictionJointConf&) (FrictionJointConf.cpp:141-231): returns
// whether neither impulse changed.
__device__ __forceinline__ bool friction_solve_velocity(float4* u, float4& velA, float4& velB) {
    const float4 r = u[0], lm = u[1], ms = u[4], mx = u[5];
    const V2 rA{r.x, r.y}, rB{r.z, r.w};
    float4 im = u[3];
    bool solved = true;
    {
        const float angularImpulse = -u[2].x * (velB.z - velA.z);
        const float old = im.z;
        im.z = clamp_std(im.z + angularImpulse, -mx.y, mx.y);
        const float inc = im.z - old;
        if (inc != 0.0f) solved = false;
        velA.z = velA.z - ms.y * inc;
        velB.z = velB.z + ms.w * inc;
    }
    {
        const V2 vb = V2{velB.x, velB.y} + mul(revperp(rB), velB.z / 1.0f);
        const V2 va = V2{velA.x, velA.y} + mul(revperp(rA), velA.z / 1.0f);
        const V2 dv = vb - va;
        // Transform(v, Mat22), each row dotted from zero (Vector.hpp:455-470).
        float tx = 0.0f, ty = 0.0f;
        tx += lm.x * dv.x;
        tx += lm.y * dv.y;
        ty += lm.z * dv.x;
        ty += lm.w * dv.y;
        const V2 old{im.x, im.y};
        V2 li{old.x + -tx, old.y + -ty};
        const float maxImpulse = mx.x;
        if (magsq(li) > maxImpulse * maxImpulse) {
            const V2 dir = unit_vector(li, V2{0.0f, 0.0f});
            li = V2{dir.x * maxImpulse, dir.y * maxImpulse};
        }
        im.x = li.x;
        im.y = li.y;
        const V2 inc = li - old;
        if (inc.x != 0.0f || inc.y != 0.0f) solved = false;
        apply_joint_impulse(ms, inc, cross(rA, inc) / 1.0f, cross(rB, inc) / 1.0f, velA, velB);
    }
    u[3] = im;
    return solved;
}

// Warm start of a distance or weld joint (DistanceJointConf.cpp:143-158,
// WeldJointConf.cpp:197-210), a rope (RopeJointConf.cpp:114-125) or a friction
// joint (FrictionJointConf.cpp:124-134).
__device__ __forceinline__ void spring_joint_warm_start(const float4* u, std::uint32_t kind, float4& velA,
                                                        float4& velB) {
    const float4 r = u[0], im = u[3], ms = u[4];
    const V2 rA{r.x, r.y}, rB{r.z, r.w};
    if (kind == prx::kJointRope && u[5].y != 0.0f) return;
    if (kind == prx::kJointDistance || kind == prx::kJointRope) {
        const V2 P = mul(im.x, V2{u[1].x, u[1].y});
        apply_joint_impulse(ms, P, cross(rA, P) / 1.0f, cross(rB, P) / 1.0f, velA, velB);
    } else if (kind == prx::kJointPulley) {
        pulley_apply(u, im.x, velA, velB);
    } else if (kind == prx::kJointPrismatic) {
        prismatic_warm_start(u, velA, velB);
    } else if (kind == prx::kJointTarget) {
        target_warm_start(u, velB);
    } else if (kind == prx::kJointWheel) {
        // WheelJointConf.cpp:212-223.
        const float4 a = u[0], s = u[1];
        const V2 P{im.x * a.x + im.y * a.z, im.x * a.y + im.y * a.w};
        apply_joint_impulse(ms, P, (im.x * s.x + im.y * s.z) / 1.0f + im.z, (im.x * s.y + im.y * s.w) / 1.0f + im.z,
                            velA, velB);
    } else if (kind == prx::kJointFriction || kind == prx::kJointMotorKind) {
        const V2 P{im.x, im.y};
        apply_joint_impulse(ms, P, cross(rA, P) / 1.0f + im.z, cross(rB, P) / 1.0f + im.z, velA, velB);
    } else {
        const V2 P{im.x, im.y};
        apply_joint_impulse(ms, P, im.z + cross(rA, P) / 1.0f, im.z + cross(rB, P) / 1.0f, velA, velB);
    }
}

// SolveVelocity(DistanceJointConf&) (DistanceJointConf.cpp:164-194): returns
// whether the impulse was zero.
__device__ __forceinline__ bool distance_solve_velocity(float4* u, float4& velA, float4& velB) {
    const float4 r = u[0], m = u[1], ms = u[4];
    const V2 rA{r.x, r.y}, rB{r.z, r.w}, dir{m.x, m.y};
    const V2 vpA = V2{velA.x, velA.y} + mul(revperp(rA), velA.z / 1.0f);
    const V2 vpB = V2{velB.x, velB.y} + mul(revperp(rB), velB.z / 1.0f);
    const float Cdot = dot(dir, vpB - vpA);
    const float acc = u[3].x;
    const float impulse = -m.z * (Cdot + u[5].x + m.w * acc);
    u[3].x = acc + impulse;
    const V2 P = mul(impulse, dir);
    apply_joint_impulse(ms, P, cross(rA, P) / 1.0f, cross(rB, P) / 1.0f, velA, velB);
    return impulse == 0.0f;
}

// SolveVelocity(WeldJointConf&) (WeldJointConf.cpp:215-303): returns whether
// the bodies' velocities were left unchanged.
__device__ __forceinline__ bool weld_solve_velocity(float4* u, float4& velA, float4& velB) {
    const float4 r = u[0], m1 = u[1], m2 = u[2], ms = u[4], gb = u[5], cf = u[7];
    const V2 rA{r.x, r.y}, rB{r.z, r.w};
    const float4 oldA = velA, oldB = velB;
    float4 im = u[3];
    if (cf.y > 0.0f) {
        const float Cdot2 = velB.z - velA.z;
        const float g = gb.x * im.z;
        const float impulse2 = -m2.y * (Cdot2 + gb.y + g);
        im.z = im.z + impulse2;
        velA.z = velA.z - ms.y * impulse2;
        velB.z = velB.z + ms.w * impulse2;
        const V2 vb = V2{velB.x, velB.y} + mul(revperp(rB), velB.z / 1.0f);
        const V2 va = V2{velA.x, velA.y} + mul(revperp(rA), velA.z / 1.0f);
        const V2 Cdot1 = vb - va;
        // Transform(Vec2, Mat33) (Math.hpp:522-527).
        const V2 t{m1.x * Cdot1.x + m1.y * Cdot1.y, m1.y * Cdot1.x + m1.w * Cdot1.y};
        const V2 impulse1 = -t;
        im.x = im.x + impulse1.x;
        im.y = im.y + impulse1.y;
        apply_joint_impulse(ms, impulse1, cross(rA, impulse1) / 1.0f, cross(rB, impulse1) / 1.0f, velA, velB);
    } else {
        const V2 vb = V2{velB.x, velB.y} + mul(revperp(rB), velB.z / 1.0f);
        const V2 va = V2{velA.x, velA.y} + mul(revperp(rA), velA.z / 1.0f);
        const V2 Cdot1 = vb - va;
        const float Cdot2 = (velB.z - velA.z) / 1.0f;
        // Mat33 times vector, row by row (Vector.hpp:455-470).
        const float tx = dot3(m1.x, m1.y, m1.z, Cdot1.x, Cdot1.y, Cdot2);
        const float ty = dot3(m1.y, m1.w, m2.x, Cdot1.x, Cdot1.y, Cdot2);
        const float tz = dot3(m1.z, m2.x, m2.y, Cdot1.x, Cdot1.y, Cdot2);
        const V2 P{-tx, -ty};
        const float L = -tz;
        im.x = im.x + P.x;
        im.y = im.y + P.y;
        im.z = im.z + L;
        apply_joint_impulse(ms, P, L + cross(rA, P) / 1.0f, L + cross(rB, P) / 1.0f, velA, velB);
    }
    u[3] = im;
    const bool same = velA.x == oldA.x && velA.y == oldA.y && velA.z == oldA.z && velB.x == oldB.x &&
                      velB.y == oldB.y && velB.z == oldB.z;
    if (same) {
        velA = oldA;
        velB = oldB;
    }
    return same;
}

// SolvePosition of a distance joint (DistanceJointConf.cpp:196-245) or a weld
// joint (WeldJointConf.cpp:305-375): returns whether it is within tolerance.
__device__ __forceinline__ bool spring_joint_solve_position(const float4* u, std::uint32_t kind, float4& posA,
                                                            float4& posB, const StepParams& c) {
    const float4 ms = u[4], la = u[6], cf = u[7];
    const float mA = ms.x, iA = ms.y, mB = ms.z, iB = ms.w;
    if (kind == prx::kJointFriction) return true;    // FrictionJointConf.cpp:233-237
    if (kind == prx::kJointTarget) return true;      // TargetJointConf.cpp:180-184
    if (kind == prx::kJointPulley) {
        // PulleyJointConf.cpp:192-253.
        const float ratio = u[2].y;
        const V2 rA = rotate(V2{la.x, la.y}, unit_from_angle(posA.z));
        const V2 rB = rotate(V2{la.z, la.w}, unit_from_angle(posB.z));
        float lengthA, lengthB;
        const V2 uA = unit_polar((V2{posA.x, posA.y} + rA) - V2{cf.x, cf.y}, lengthA);
        const V2 uB = unit_polar((V2{posB.x, posB.y} + rB) - V2{cf.z, cf.w}, lengthB);
        const float ruA = cross(rA, uA);
        const float ruB = cross(rB, uB);
        const float tA = mA + iA * (ruA * ruA) / 1.0f;
        const float tB = mB + iB * (ruB * ruB) / 1.0f;
        const float total = tA + (ratio * ratio) * tB;
        const float mass = (total > 0.0f) ? 1.0f / total : 0.0f;
        const float C = (u[5].x + ratio * u[5].y) - (lengthA + ratio * lengthB);
        const float impulse = -mass * C;
        const V2 PA = mul(-impulse, uA);
        const V2 PB = mul(-ratio * impulse, uB);
        posA.x = posA.x + mA * PA.x;
        posA.y = posA.y + mA * PA.y;
        posA.z = posA.z + iA * cross(rA, PA) / 1.0f;
        posB.x = posB.x + mB * PB.x;
        posB.y = posB.y + mB * PB.y;
        posB.z = posB.z + iB * cross(rB, PB) / 1.0f;
        return __builtin_fabsf(C) < c.linearSlop;
    }
    if (kind == prx::kJointMotorKind) return true;   // MotorJointConf.cpp:231-234
    if (kind == prx::kJointDistance && cf.y > 0.0f) return true;
    const V2 qA = unit_from_angle(posA.z);
    const V2 rA = rotate(V2{la.x, la.y}, qA);
    const V2 rB = rotate(V2{la.z, la.w}, unit_from_angle(posB.z));
    const V2 d = (V2{posB.x, posB.y} + rB) - (V2{posA.x, posA.y} + rA);
    V2 P;
    float LA, LB;
    bool ok;
    if (kind == prx::kJointPrismatic) {
        // PrismaticJointConf.cpp:409-531.
        const V2 axisA{cf.x, cf.y};
        const float lower = cf.w, upper = u[5].x;
        const V2 axis = rotate(axisA, qA);
        const float a1 = cross(d + rA, axis);
        const float a2 = cross(rB, axis);
        const V2 perp = rotate(revperp(axisA), qA);
        const float s1 = cross(d + rA, perp);
        const float s2 = cross(rB, perp);
        const V2 C1{dot(perp, d), posB.z - posA.z - cf.z};
        float linearError = __builtin_fabsf(C1.x);
        const float angularError = __builtin_fabsf(C1.y);
        bool active = false;
        float C2 = 0.0f;
        if (__float_as_uint(u[2].w) & prx::kJointLimit) {
            const float translation = dot(axis, d);
            if (__builtin_fabsf(upper - lower) < (2.0f * c.linearSlop)) {
                C2 = clamp_std(translation, -c.maxLinearCorrection, c.maxLinearCorrection);
                linearError = fmax_std(linearError, __builtin_fabsf(translation));
                active = true;
            } else if (translation <= lower) {
                C2 = clamp_std(translation - lower + c.linearSlop, -c.maxLinearCorrection, 0.0f);
                linearError = fmax_std(linearError, lower - translation);
                active = true;
            } else if (translation >= upper) {
                C2 = clamp_std(translation - upper - c.linearSlop, 0.0f, c.maxLinearCorrection);
                linearError = fmax_std(linearError, translation - upper);
                active = true;
            }
        }
        float ix, iy, iz;
        if (active) {
            Jc kj;
            kj.exx = mA + iA * (s1 * s1) / 1.0f + mB + iB * (s2 * s2) / 1.0f;
            kj.eyx = iA * s1 + iB * s2;
            kj.ezx = iA * s1 * a1 / 1.0f + iB * s2 * a2 / 1.0f;
            const float k22 = iA + iB;
            kj.eyy = (k22 == 0.0f) ? 1.0f : k22;
            kj.ezy = iA * a1 + iB * a2;
            kj.ezz = mA + iA * (a1 * a1) / 1.0f + mB + iB * (a2 * a2) / 1.0f;
            solve33(kj, -C1.x, -C1.y, -C2, ix, iy, iz);
        } else {
            const float k11 = mA + iA * s1 * s1 / 1.0f + mB + iB * s2 * s2 / 1.0f;
            const float k12 = iA * s1 + iB * s2;
            float k22 = iA + iB;
            if (k22 == 0.0f) k22 = 1.0f;
            // Solve(Mat22, Vec2) (Math.hpp:413-426), columns (k11, k12) and (k12, k22).
            const float cp = k11 * k22 - k12 * k12;
            const V2 bb{-C1.x, -C1.y};
            ix = 0.0f;
            iy = 0.0f;
            if (!almost_zero(cp)) {
                const float inverse = 1.0f / cp;
                ix = (k22 * bb.x - k12 * bb.y) * inverse;
                iy = (k11 * bb.y - k12 * bb.x) * inverse;
            }
            iz = 0.0f;
        }
        P = mul(ix, perp) + mul(iz, axis);
        LA = (ix * s1 + iy + iz * a1);
        LB = (ix * s2 + iy + iz * a2);
        ok = (linearError <= c.linearSlop) && (angularError <= c.angularSlop);
    } else if (kind == prx::kJointWheel) {
        // WheelJointConf.cpp:309-359; the mass uses InitVelocity's sAy, sBy.
        const V2 dw = (V2{posB.x, posB.y} - V2{posA.x, posA.y}) + (rB - rA);
        const V2 ay = rotate(revperp(V2{cf.z, cf.w}), qA);
        const float sAy = cross(dw + rA, ay);
        const float sBy = cross(rB, ay);
        const float C = dot(dw, ay);
        const float4 s = u[1];
        const float k = mA + mB + iA * (s.x * s.x) / 1.0f + iB * (s.y * s.y) / 1.0f;
        const float impulse = (k != 0.0f) ? -(C / k) : 0.0f;
        P = mul(impulse, ay);
        LA = impulse * sAy / 1.0f;
        LB = impulse * sBy / 1.0f;
        ok = __builtin_fabsf(C) <= c.linearSlop;
    } else if (kind == prx::kJointRope) {
        // RopeJointConf.cpp:169-212.
        float length;
        const V2 dir = unit_polar(d, length);
        const float C = clamp_std(length - cf.z, 0.0f, c.maxLinearCorrection);
        const float impulse = -u[1].z * C;
        P = mul(impulse, dir);
        LA = cross(rA, P) / 1.0f;
        LB = cross(rB, P) / 1.0f;
        ok = (length - cf.z) < c.linearSlop;
    } else if (kind == prx::kJointDistance) {
        float length;
        const V2 dir = unit_polar(d, length);
        const float C = clamp_std(length - cf.z, -c.maxLinearCorrection, c.maxLinearCorrection);
        const float impulse = -u[1].z * C;
        P = mul(impulse, dir);
        LA = cross(rA, P) / 1.0f;
        LB = cross(rB, P) / 1.0f;
        ok = __builtin_fabsf(C) < c.linearSlop;
    } else {
        const K33 K = joint_k33(mA, rA, iA, mB, rB, iB);
        const float positionError = __builtin_sqrtf(magsq(d));
        float angularError = 0.0f, L = 0.0f;
        Jc kj;
        kj.exx = K.exx;
        kj.eyx = K.eyx;
        kj.ezx = K.ezx;
        kj.eyy = K.eyy;
        kj.ezy = K.ezy;
        kj.ezz = K.ezz;
        if (cf.y > 0.0f) {
            P = -solve22(kj, d);
            LA = cross(rA, P) / 1.0f;
            LB = cross(rB, P) / 1.0f;
        } else {
            const float C2 = posB.z - posA.z - cf.x;
            angularError = __builtin_fabsf(C2);
            if (K.ezz > 0.0f) {
                float sx, sy, sz;
                solve33(kj, d.x, d.y, C2, sx, sy, sz);
                P = V2{-sx, -sy};
                L = -sz;
            } else {
                P = -solve22(kj, d);
                L = 0.0f;
            }
            LA = L + cross(rA, P) / 1.0f;
            LB = L + cross(rB, P) / 1.0f;
        }
        ok = (positionError <= c.linearSlop) && (angularError <= c.angularSlop);
    }
    posA.x = posA.x - mA * P.x;
    posA.y = posA.y - mA * P.y;
    posA.z = posA.z - iA * LA;
    posB.x = posB.x + mB * P.x;
    posB.y = posB.y + mB * P.y;
    posB.z = posB.z + iB * LB;
    return ok;
}

// Gear record (kJcStride float4): u[0] (JvAC, JvBD); u[1] (JwA, JwB, JwC,
// JwD); u[2] (mass, ratio, constant, flags | kind << 16); u[3] (impulse, 0,
// 0, 0); u[4] (mA, iA, mB, iB); u[5] (mC, iC, body A, body B); u[6] (mD, iD,
// body C, body D); u[7] the A-C side's anchors less the local centres (local
// anchor A - centre of C, local anchor B - centre of A; a revolute side's
// reference angle in .x); u[8] the B-D side's (anchor A - centre of D, anchor
// B - centre of B); u[9] the sides' local x axes.  Revolute sides have no
// linear terms and angular terms 1 (A-C) or the ratio (B-D).

// A prismatic side's Jacobian at the given angles (GearJointConf.cpp:188-204,
// 213-229): the axis rotated by the other body (C or D) scaled by s, and the
// crosses of both arms with it (scaled by s on the B-D side only).
struct GearSide {
    V2 Jv;
    float JwSelf, JwOther;
};
__device__ __forceinline__ GearSide gear_side_prismatic(V2 axis, V2 lOther, V2 lSelf, float angSelf, float angOther,
                                                       float s, bool scaled) {
    const V2 qSelf = unit_from_angle(angSelf);
    const V2 qOther = unit_from_angle(angOther);
    const V2 u = rotate(axis, qOther);
    const V2 rOther = rotate(lOther, qOther);
    const V2 rSelf = rotate(lSelf, qSelf);
    GearSide g;
    if (scaled) {
        g.Jv = V2{s * u.x, s * u.y};
        g.JwOther = s * cross(rOther, u);
        g.JwSelf = s * cross(rSelf, u);
    } else {
        g.Jv = V2{1.0f * u.x, 1.0f * u.y};
        g.JwOther = cross(rOther, u);
        g.JwSelf = cross(rSelf, u);
    }
    return g;
}

// InitVelocity(GearJointConf&) less its warm start (GearJointConf.cpp:171-237).
__device__ __forceinline__ void build_jc_gear(const DeviceState& d, int jg, float4* u, const BodyC& A,
                                              const float4& pA, const BodyC& B, const float4& pB, const BodyC& C,
                                              const float4& pC, const BodyC& D, const float4& pD, int bA, int bB,
                                              int bC, int bD, std::uint32_t flags, const StepParams& c) {
    const float4 g0 = d.j_gear[2 * std::size_t(jg)], ax = d.j_gear[2 * std::size_t(jg) + 1];
    const float4 ac = d.j_anchor[jg], bd = d.j_cfg[jg];
    const float ratio = g0.x;
    float invMass = 0.0f;
    float4 sideAC = ac, sideBD = bd;
    V2 JvAC{0.0f, 0.0f}, JvBD{0.0f, 0.0f};
    float JwA, JwB, JwC, JwD;
    if (flags & prx::kGearPrisAC) {
        const V2 lC = V2{ac.x, ac.y} - C.lc, lA = V2{ac.z, ac.w} - A.lc;
        sideAC = float4{lC.x, lC.y, lA.x, lA.y};
        const GearSide g = gear_side_prismatic(V2{ax.x, ax.y}, lC, lA, pA.z, pC.z, 1.0f, false);
        JvAC = g.Jv;
        JwC = g.JwOther;
        JwA = g.JwSelf;
        const float invRotMassC = C.invRotI * (JwC * JwC) / 1.0f;
        const float invRotMassA = A.invRotI * (JwA * JwA) / 1.0f;
        invMass += C.invMass + A.invMass + invRotMassC + invRotMassA;
    } else {
        JwA = 1.0f;
        JwC = 1.0f;
        invMass += A.invRotI + C.invRotI;
    }
    if (flags & prx::kGearPrisBD) {
        const V2 lD = V2{bd.x, bd.y} - D.lc, lB = V2{bd.z, bd.w} - B.lc;
        sideBD = float4{lD.x, lD.y, lB.x, lB.y};
        const GearSide g = gear_side_prismatic(V2{ax.z, ax.w}, lD, lB, pB.z, pD.z, ratio, true);
        JvBD = g.Jv;
        JwD = g.JwOther;
        JwB = g.JwSelf;
        const float invRotMassD = D.invRotI * (JwD * JwD) / 1.0f;
        const float invRotMassB = B.invRotI * (JwB * JwB) / 1.0f;
        invMass += (ratio * ratio) * (D.invMass + B.invMass) + invRotMassD + invRotMassB;
    } else {
        JwB = ratio * 1.0f;
        JwD = ratio * 1.0f;
        invMass += (ratio * ratio) * (B.invRotI + D.invRotI);
    }
    const float mass = (invMass > 0.0f) ? 1.0f / invMass : 0.0f;
    u[0] = float4{JvAC.x, JvAC.y, JvBD.x, JvBD.y};
    u[1] = float4{JwA, JwB, JwC, JwD};
    u[2] = float4{mass, ratio, g0.y, __uint_as_float((flags & 0xffu) | (prx::kJointGear << 16))};
    u[3] = float4{c.warmStart ? d.j_imp[jg].x : 0.0f, 0.0f, 0.0f, 0.0f};
    u[4] = float4{A.invMass, A.invRotI, B.invMass, B.invRotI};
    u[5] = float4{C.invMass, C.invRotI, __int_as_float(bA), __int_as_float(bB)};
    u[6] = float4{D.invMass, D.invRotI, __int_as_float(bC), __int_as_float(bD)};
    u[7] = sideAC;
    u[8] = sideBD;
    u[9] = ax;
}

// The velocity change of an impulse on the four bodies (GearJointConf.cpp:
// 226-235 and 281-288): A and B gain, C and D lose.
__device__ __forceinline__ void gear_apply(const float4* u, float imp, float4& vA, float4& vB, float4& vC,
                                           float4& vD) {
    const float4 J = u[0], W = u[1], mab = u[4], mc = u[5], md = u[6];
    const float lA = mab.x * imp, lB = mab.z * imp, lC = mc.x * imp, lD = md.x * imp;
    vA.x = vA.x + lA * J.x;
    vA.y = vA.y + lA * J.y;
    vA.z = vA.z + mab.y * imp * W.x / 1.0f;
    vB.x = vB.x + lB * J.z;
    vB.y = vB.y + lB * J.w;
    vB.z = vB.z + mab.w * imp * W.y / 1.0f;
    vC.x = vC.x - lC * J.x;
    vC.y = vC.y - lC * J.y;
    vC.z = vC.z - mc.y * imp * W.z / 1.0f;
    vD.x = vD.x - lD * J.z;
    vD.y = vD.y - lD * J.w;
    vD.z = vD.z - md.y * imp * W.w / 1.0f;
}

// SolveVelocity(GearJointConf&) (GearJointConf.cpp:257-296): returns whether
// the impulse was zero.
__device__ __forceinline__ bool gear_solve_velocity(float4* u, float4& vA, float4& vB, float4& vC, float4& vD) {
    const float4 J = u[0], W = u[1];
    const float acDot = dot(V2{J.x, J.y}, V2{vA.x - vC.x, vA.y - vC.y});
    const float bdDot = dot(V2{J.z, J.w}, V2{vB.x - vD.x, vB.y - vD.y});
    const float Cdot = acDot + bdDot + (W.x * vA.z - W.z * vC.z) / 1.0f + (W.y * vB.z - W.w * vD.z) / 1.0f;
    const float impulse = -u[2].x * 1.0f * Cdot;
    u[3].x = u[3].x + impulse;
    gear_apply(u, impulse, vA, vB, vC, vD);
    return impulse == 0.0f;
}

// Cap(Position, ConstraintSolverConf) (d2/Math.cpp:63-71).
__device__ __forceinline__ float4 cap_position(float4 p, const StepParams& c) {
    const float lsquared = magsq(V2{p.x, p.y});
    if (lsquared > c.maxLinearCorrection * c.maxLinearCorrection) {
        const float s = c.maxLinearCorrection / __builtin_sqrtf(lsquared);
        p.x = p.x * s;
        p.y = p.y * s;
    }
    p.z = clamp_std(p.z, -c.maxAngularCorrection, +c.maxAngularCorrection);
    return p;
}

// SolvePosition(const GearJointConf&) (GearJointConf.cpp:298-471): the
// coordinates at the positions now, one capped correction on all four
// bodies; returns whether it is within tolerance.
__device__ __forceinline__ bool gear_solve_position(const float4* u, float4& pA, float4& pB, float4& pC, float4& pD,
                                                    const StepParams& c) {
    const float4 w2 = u[2], mab = u[4], mc = u[5], md = u[6], sAC = u[7], sBD = u[8], ax = u[9];
    const float ratio = w2.y, constant = w2.z;
    const std::uint32_t flags = __float_as_uint(w2.w);
    if (!__builtin_isfinite(constant)) return true;
    V2 JvAC{0.0f, 0.0f}, JvBD{0.0f, 0.0f};
    float JwA, JwB, JwC, JwD, coordinateA, coordinateB;
    float invMass = 0.0f;
    if (flags & prx::kGearPrisAC) {
        const V2 lC{sAC.x, sAC.y}, lA{sAC.z, sAC.w}, axis{ax.x, ax.y};
        const V2 qA = unit_from_angle(pA.z);
        const V2 qC = unit_from_angle(pC.z);
        const V2 uu = rotate(axis, qC);
        const V2 rC = rotate(lC, qC);
        const V2 rA = rotate(lA, qA);
        JvAC = V2{uu.x * 1.0f, uu.y * 1.0f};
        JwC = cross(rC, uu);
        JwA = cross(rA, uu);
        const float invLinMass = mc.x + mab.x;
        const float sc = JwC * 1.0f / 1.0f, sa = JwA * 1.0f / 1.0f;
        const float invRotMassC = mc.y * (sc * sc);
        const float invRotMassA = mab.y * (sa * sa);
        invMass += invLinMass + invRotMassC + invRotMassA;
        const V2 p = invrotate(rA + (V2{pA.x, pA.y} - V2{pC.x, pC.y}), qC);
        coordinateA = dot(p - lC, axis) / 1.0f;
    } else {
        JwA = 1.0f;
        JwC = 1.0f;
        invMass += mab.y + mc.y;
        coordinateA = ((pA.z - pC.z) - sAC.x) / 1.0f;
    }
    if (flags & prx::kGearPrisBD) {
        const V2 lD{sBD.x, sBD.y}, lB{sBD.z, sBD.w}, axis{ax.z, ax.w};
        const V2 qB = unit_from_angle(pB.z);
        const V2 qD = unit_from_angle(pD.z);
        const V2 uu = rotate(axis, qD);
        const V2 rD = rotate(lD, qD);
        const V2 rB = rotate(lB, qB);
        JvBD = V2{ratio * uu.x, ratio * uu.y};
        JwD = ratio * cross(rD, uu);
        JwB = ratio * cross(rB, uu);
        const float invLinMass = (ratio * ratio) * (md.x + mab.z);
        const float sd = JwD * 1.0f / 1.0f, sb = JwB * 1.0f / 1.0f;
        const float invRotMassD = md.y * (sd * sd);
        const float invRotMassB = mab.w * (sb * sb);
        invMass += invLinMass + invRotMassD + invRotMassB;
        const V2 p = invrotate(rB + (V2{pB.x, pB.y} - V2{pD.x, pD.y}), qD);
        coordinateB = dot(p - lD, axis) / 1.0f;
    } else {
        JwB = ratio;
        JwD = ratio;
        invMass += (ratio * ratio) * (mab.w + md.y);
        coordinateB = ((pB.z - pD.z) - sBD.x) / 1.0f;
    }
    const float C = ((coordinateA + ratio * coordinateB) - constant) * c.regResolutionRate;
    const float impulse = ((invMass > 0.0f) ? -C / invMass : 0.0f) * 1.0f * 1.0f;
    auto delta = [&](float m, float i, V2 Jv, float Jw) {
        const float l = m * impulse;
        return cap_position(float4{l * Jv.x, l * Jv.y, i * impulse * Jw * 1.0f / 1.0f, 0.0f}, c);
    };
    const float4 dA = delta(mab.x, mab.y, JvAC, JwA);
    const float4 dC = delta(mc.x, mc.y, JvAC, JwC);
    const float4 dB = delta(mab.z, mab.w, JvBD, JwB);
    const float4 dD = delta(md.x, md.y, JvBD, JwD);
    pA = float4{pA.x + dA.x, pA.y + dA.y, pA.z + dA.z, pA.w};
    pB = float4{pB.x + dB.x, pB.y + dB.y, pB.z + dB.z, pB.w};
    pC = float4{pC.x - dC.x, pC.y - dC.y, pC.z - dC.z, pC.w};
    pD = float4{pD.x - dD.x, pD.y - dD.y, pD.z - dD.z, pD.w};
    float linearError = 0.0f, angularError = 0.0f;
    linearError = fmax_std(linearError, magsq(V2{dA.x, dA.y}));
    linearError = fmax_std(linearError, magsq(V2{dB.x, dB.y}));
    linearError = fmax_std(linearError, magsq(V2{dC.x, dC.y}));
    linearError = fmax_std(linearError, magsq(V2{dD.x, dD.y}));
    angularError = fmax_std(angularError, __builtin_fabsf(dA.z));
    angularError = fmax_std(angularError, __builtin_fabsf(dB.z));
    angularError = fmax_std(angularError, __builtin_fabsf(dC.z));
    angularError = fmax_std(angularError, __builtin_fabsf(dD.z));
    return (linearError <= c.linearSlop * c.linearSlop) && (angularError <= c.angularSlop);
}

// Per-step joint set-up half of InitVelocity (RevoluteJointConf.cpp:127-197),
// writes the joint's solver record.
__device__ __forceinline__ void build_jc(const DeviceState& d, const WorldCtx& x, int jg, float4* u, const BodyC& A,
                                         const float4& pA, const BodyC& B, const float4& pB, int bA, int bB,
                                         const StepParams& c) {
    const float4 an = d.j_anchor[jg], cf = d.j_cfg[jg], cf2 = d.j_cfg2[jg];
    float4 im = d.j_imp[jg];
    std::uint32_t limit = d.j_limit[jg];
    const std::uint32_t flags = __float_as_uint(cf2.y);
    const V2 laA = V2{an.x, an.y} - A.lc;
    const V2 laB = V2{an.z, an.w} - B.lc;
    const V2 qA = unit_from_angle(pA.z);
    const V2 qB = unit_from_angle(pB.z);
    const V2 rA = rotate(laA, qA);
    const V2 rB = rotate(laB, qB);
    const float mA = A.invMass, iA = A.invRotI, mB = B.invMass, iB = B.invRotI;
    const std::uint32_t kind = flags >> 8;
    if (kind != prx::kJointRevolute) {
        u[0] = float4{rA.x, rA.y, rB.x, rB.y};
        u[2] = float4{0.0f, 0.0f, 0.0f, __uint_as_float((flags & 0xffu) | (kind << 16))};
        u[4] = float4{mA, iA, mB, iB};
        u[5] = float4{0.0f, 0.0f, __int_as_float(bA), __int_as_float(bB)};
        u[6] = float4{laA.x, laA.y, laB.x, laB.y};
        if (kind == prx::kJointDistance) build_jc_distance(u, cf, im, rA, rB, pA, pB, mA, iA, mB, iB, c);
        else if (kind == prx::kJointWeld) build_jc_weld(u, cf, im, rA, rB, pA, pB, mA, iA, mB, iB, flags & 0xffu, c);
        else if (kind == prx::kJointRope) build_jc_rope(u, cf, im, rA, rB, pA, pB, mA, iA, mB, iB, flags & 0xffu, c);
        else if (kind == prx::kJointFriction) build_jc_friction(u, cf, im, rA, rB, mA, iA, mB, iB, flags & 0xffu, c);
        else if (kind == prx::kJointWheel)
            build_jc_wheel(u, cf, cf2, im, rA, rB, qA, pA, pB, mA, iA, mB, iB, flags & 0xffu, c);
        else if (kind == prx::kJointPulley)
            build_jc_pulley(u, cf, cf2, im, rA, rB, pA, pB, mA, iA, mB, iB, flags & 0xffu, c);
        else if (kind == prx::kJointPrismatic)
            build_jc_prismatic(u, cf, cf2, im, rA, rB, qA, pA, pB, flags & 0xffu, limit, c);
        else if (kind == prx::kJointTarget)
            build_jc_target(u, cf, cf2, im, rB, pB, mB, iB, flags & 0xffu, c);
        else {
            // Body B's anchor is its origin: rB = Rotate(-localCenterB, qB).
            const V2 rBm = rotate(V2{-B.lc.x, -B.lc.y}, qB);
            u[0] = float4{rA.x, rA.y, rBm.x, rBm.y};
            build_jc_motor(u, cf, im, rA, rBm, pA, pB, mA, iA, mB, iB, flags & 0xffu, c);
        }
        return;
    }
    const float totInvI = iA + iB;
    const bool fixedRotation = (totInvI == 0.0f);
    // GetMat33 (RevoluteJointConf.cpp:36-62).
    const float exx = mA + ((rA.y * rA.y) * iA / 1.0f) + mB + ((rB.y * rB.y) * iB / 1.0f);
    const float eyx = (-rA.y * rA.x * iA / 1.0f) + (-rB.y * rB.x * iB / 1.0f);
    const float ezx = (-rA.y * iA) + (-rB.y * iB);
    const float eyy = mA + ((rA.x * rA.x) * iA / 1.0f) + mB + ((rB.x * rB.x) * iB / 1.0f);
    const float ezy = (rA.x * iA) + (rB.x * iB);
    const float angularMass = (totInvI > 0.0f) ? (1.0f / totInvI) : 0.0f;
    if (!(flags & prx::kJointMotor) || fixedRotation) im.w = 0.0f;
    if ((flags & prx::kJointLimit) && !fixedRotation) {
        const float jointAngle = pB.z - pA.z - cf.x;
        if (__builtin_fabsf(cf.z - cf.y) < (c.angularSlop * 2.0f)) {
            limit = 3u;
        } else if (jointAngle <= cf.y) {
            if (limit != 1u) {
                limit = 1u;
                im.z = 0.0f;
            }
        } else if (jointAngle >= cf.z) {
            if (limit != 2u) {
                limit = 2u;
                im.z = 0.0f;
            }
        } else {
            limit = 0u;
            im.z = 0.0f;
        }
    } else {
        limit = 0u;
    }
    if (c.warmStart) {
        im.x = im.x * c.dtRatio;
        im.y = im.y * c.dtRatio;
        im.z = im.z * c.dtRatio;
        im.w = im.w * c.dtRatio;
    } else {
        im = float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    u[0] = float4{rA.x, rA.y, rB.x, rB.y};
    u[1] = float4{exx, eyx, ezx, eyy};
    u[2] = float4{ezy, totInvI, angularMass, __uint_as_float(flags | (limit << 8))};
    u[3] = im;
    u[4] = float4{mA, iA, mB, iB};
    u[5] = float4{cf.w, c.dt * cf2.x, __int_as_float(bA), __int_as_float(bB)};
    u[6] = float4{laA.x, laA.y, laB.x, laB.y};
    u[7] = float4{cf.x, cf.y, cf.z, __uint_as_float(prx::kJointRevolute)};
}

__device__ __forceinline__ BodyC load_bodyc(const DeviceState& d, int gb) {
    const float4 l = d.lcinv[gb];
    return BodyC{l.z, l.w, V2{l.x, l.y}};
}

// GetVelocity(body, h) + Cap (Body.cpp:290-309, Velocity.cpp:27-45).
__device__ __forceinline__ float4 integrated_velocity(const DeviceState& d, int gb, const StepParams& c) {
    const float4 v = d.vel[gb];
    V2 lin{v.x, v.y};
    float ang = v.z;
    const float h = c.dt;
    if (d.bflags[gb] & prx::kAccelerable) {
        const float4 a = d.acc[gb];
        const float2 dm = d.damp[gb];
        lin = lin + mul(h, V2{a.x, a.y});
        ang = ang + h * a.z;
        const float linv = 1.0f / (1.0f + h * dm.x);
        lin = V2{lin.x * linv, lin.y * linv};
        ang = ang / (1.0f + h * dm.y);
    }
    const V2 tr = mul(h, lin);
    const float lsq = magsq(tr);
    if (lsq > c.maxTranslation * c.maxTranslation) {
        const float ratio = c.maxTranslation / __builtin_sqrtf(lsq);
        lin = V2{lin.x * ratio, lin.y * ratio};
    }
    const float absRot = __builtin_fabsf(h * ang);
    if (absRot > c.maxRotation) {
        const float ratio = c.maxRotation / absRot;
        ang = ang * ratio;
    }
    return float4{lin.x, lin.y, ang, 0.0f};
}

// ---------------------------------------------------------------------------
// SolveReg (AabbTreeWorld.cpp:1472-1561) and SolveRegIslandViaGS (:1563-1682).
//
// Only two things in an island solve depend on order: the island's contact
// order (the DFS) and the Gauss-Seidel sweeps over it.  Everything else is
// independent per body or per contact and runs lane-parallel over all islands
// of the world at once.  Static bodies are left out of island body lists: the
// reference adds them but they never propagate the search, their solver
// updates are exact no-ops (zero inverse mass and velocity) and they are not
// written back, so islands become disjoint in every body.
//
// Body-constraint positions and velocities live in LDS when the world fits
// (32 bytes per body), else in global scratch.

template <bool kLds>
struct Bodies {
    float4* gpos;
    float4* gvel;
    int nb;
    __device__ __forceinline__ float4 pos(int b) const {
        if constexpr (kLds) return g_smem[b];
        else return gpos[b];
    }
    __device__ __forceinline__ float4 vel(int b) const {
        if constexpr (kLds) return g_smem[nb + b];
        else return gvel[b];
    }
    __device__ __forceinline__ void set_pos(int b, float4 v) const {
        if constexpr (kLds) g_smem[b] = v;
        else gpos[b] = v;
    }
    __device__ __forceinline__ void set_vel(int b, float4 v) const {
        if constexpr (kLds) g_smem[nb + b] = v;
        else gvel[b] = v;
    }
};

// Exclusive prefix sum of a[0..n) in place (wave-parallel, carry across chunks);
// returns the total.
__device__ __forceinline__ int wave_exclusive_scan(int* a, int n, int lane) {
    int carry = 0;
    for (int base = 0; base < n; base += kWave) {
        const int i = base + lane;
        const int v = (i < n) ? a[i] : 0;
        const int inc = wave_incl_scan(v, lane);
        if (i < n) a[i] = carry + inc - v;
        carry += __shfl(inc, kWave - 1);
    }
    return carry;
}

struct IslandScratch {
    int* off;         // per-body CSR offsets (nb + 1)
    int2* csr;        // per-body touching contacts: contact index, other body
    int* ibody;       // island body lists
    int* icont;       // island contact lists (contact indices)
    int* ijoint;      // island joint lists (joint indices)
    int2* istart;     // island i: first body entry, first contact entry
    int* jstart;      // island i: first joint entry
    int* bisland;     // island of each body (-1: none)
    std::uint32_t* isl_word;  // per island: solved position iteration + 1 (0: not solved)
    std::uint32_t* isl_under;
};

// Per-body CSR of touching, enabled, non-sensor contacts in contact order
// (each body's contact list order), entries carry the other body.
// kAll: every contact of each body (the TOI phase's per-body contact lists).
template <bool kAll = false>
__device__ __forceinline__ void build_adjacency(const DeviceState& d, const WorldCtx& x, int nc, const IslandScratch& S) {
    for (int b = x.lane; b <= x.nb; b += kWave) S.off[b] = 0;
    __syncthreads();
    for (int i = x.lane; i < nc; i += kWave) {
        const std::uint32_t f = d.c_flags[x.co + i];
        if (kAll || ((f & prx::kCEnabled) && (f & prx::kTouching) && !(f & prx::kSensor))) {
            const int2 bd = d.c_body[x.co + i];
            atomicAdd(&S.off[bd.x], 1);
            atomicAdd(&S.off[bd.y], 1);
        }
    }
    __syncthreads();
    wave_exclusive_scan(S.off, x.nb + 1, x.lane);
    __syncthreads();
    int* cur = d.s_stack + 2 * x.bo;
    for (int b = x.lane; b < x.nb; b += kWave) cur[b] = S.off[b];
    __syncthreads();
    for (int i = x.lane; i < nc; i += kWave) {
        const std::uint32_t f = d.c_flags[x.co + i];
        if (kAll || ((f & prx::kCEnabled) && (f & prx::kTouching) && !(f & prx::kSensor))) {
            const int2 bd = d.c_body[x.co + i];
            S.csr[atomicAdd(&cur[bd.x], 1)] = int2{i, bd.y};
            S.csr[atomicAdd(&cur[bd.y], 1)] = int2{i, bd.x};
        }
    }
    __syncthreads();
    for (int b = x.lane; b < x.nb; b += kWave) {
        const int s0 = S.off[b], e0 = S.off[b + 1];
        for (int i = s0 + 1; i < e0; ++i) {
            const int2 v = S.csr[i];
            int j = i - 1;
            while (j >= s0 && S.csr[j].x > v.x) {
                S.csr[j + 1] = S.csr[j];
                --j;
            }
            S.csr[j + 1] = v;
        }
    }
    __syncthreads();
}

// Islands (AddToIsland, AabbTreeWorld.cpp:1358-1444; seeds in body order
// :1495-1523).  Awake speedable bodies without touching contacts are
// one-body islands and are emitted lane-parallel; the rest are found by a
// depth-first search on one lane.  kMode 2: adjacency and flags in LDS,
// 1: flags in LDS, 0: everything in global scratch.  Flag bytes per body:
// 0 free, 1 islanded, 2 static (never pushed).  Returns the island count.
template <int kMode>
__device__ __forceinline__ int find_islands(const DeviceState& d, const WorldCtx& x, int nc, const IslandScratch& S,
                                            int& maxBodies) {
    __shared__ int s_cnt[2];
    const int nent = S.off[x.nb];
    int* off = S.off;
    int2* csr = S.csr;
    std::uint8_t* islb;
    std::uint8_t* islc;
    std::uint8_t* islj;
    if constexpr (kMode == 2) {
        int* loff = reinterpret_cast<int*>(g_smem);
        int2* lcsr = reinterpret_cast<int2*>(g_smem) + ((x.nb + 2) >> 1);
        for (int b = x.lane; b <= x.nb; b += kWave) loff[b] = S.off[b];
        for (int e = x.lane; e < nent; e += kWave) lcsr[e] = S.csr[e];
        off = loff;
        csr = lcsr;
        islb = reinterpret_cast<std::uint8_t*>(lcsr + nent);
        islc = islb + x.nb;
        islj = islc + nc;
    } else if constexpr (kMode == 1) {
        islb = reinterpret_cast<std::uint8_t*>(g_smem);
        islc = islb + x.nb;
        islj = islc + nc;
    } else {
        islb = d.s_isl_b + x.bo;
        islc = d.s_isl_c + x.co;
        islj = d.s_isl_j + x.jo;
    }
    for (int b = x.lane; b < x.nb; b += kWave) {
        islb[b] = (d.bflags[x.bo + b] & prx::kSpeedable) ? 0 : 2;
        S.bisland[b] = -1;
    }
    for (int i = x.lane; i < nc; i += kWave) islc[i] = 0;
    for (int i = x.lane; i < x.nj; i += kWave) islj[i] = 0;
    __syncthreads();
    // One-body islands.
    int nsingle = 0;
    for (int base = 0; base < x.nb; base += kWave) {
        const int b = base + x.lane;
        bool single = false;
        if (b < x.nb) {
            const std::uint32_t f = d.bflags[x.bo + b];
            single = (f & prx::kAwake) && (f & prx::kEnabled) && (f & prx::kSpeedable) && off[b] == off[b + 1] &&
                     (x.nj == 0 || d.bjoint[x.bo + b].y == 0);
        }
        const unsigned long long m = __ballot(single);
        if (single) {
            const int r = nsingle + __popcll(m & lanes_below(x.lane));
            S.ibody[r] = b;
            S.istart[r] = int2{r, 0};
            S.jstart[r] = 0;
            S.bisland[b] = r;
            islb[b] = 1;
        }
        nsingle += __popcll(m);
    }
    __syncthreads();
    if (x.lane == 0) {
        int nisl = nsingle, nbe = nsingle, nce = 0, nje = 0;
        int* stack = d.s_stack + 2 * x.bo;
        // The reference's island body lists also hold the bodies without
        // velocity (static) that its contacts and joints reach, once per island
        // (AabbTreeWorld.cpp:1386-1398); they count in maxIslandBodies.  They
        // are marked 3 while counted and listed from the top of the stack
        // array (pushes plus listed bodies never exceed the world's bodies).
        int maxB = nsingle > 0 ? 1 : 0;
        for (int seed = 0; seed < x.nb; ++seed) {
            if (islb[seed]) continue;
            const std::uint32_t sf = d.bflags[x.bo + seed];
            if (!((sf & prx::kAwake) && (sf & prx::kEnabled))) continue;
            S.istart[nisl] = int2{nbe, nce};
            S.jstart[nisl] = nje;
            const int nbe0 = nbe;
            int nst = 0;
            int sp = 0;
            stack[sp++] = seed;
            islb[seed] = 1;
            while (sp > 0) {
                const int bb = stack[--sp];
                S.ibody[nbe++] = bb;
                S.bisland[bb] = nisl;
                const int e1 = off[bb + 1];
                for (int e = off[bb]; e < e1; ++e) {
                    const int2 en = csr[e];
                    if (islc[en.x]) continue;
                    S.icont[nce++] = en.x;
                    islc[en.x] = 1;
                    if (islb[en.y] == 0) {
                        islb[en.y] = 1;
                        stack[sp++] = en.y;
                    } else if (islb[en.y] == 2) {
                        islb[en.y] = 3;
                        stack[x.nb - 1 - nst++] = en.y;
                    }
                }
                // AddJointsToIsland (AabbTreeWorld.cpp:1446-1470).
                if (x.nj) {
                    const int2 bj = d.bjoint[x.bo + bb];
                    for (int k = 0; k < bj.y; ++k) {
                        const int2 je = d.bjoint_list[2 * x.jo + bj.x + k];
                        if (islj[je.x]) continue;
                        // A target joint has no other body (-1).
                        if (je.y >= 0 && !(d.bflags[x.bo + je.y] & prx::kEnabled)) continue;
                        islj[je.x] = 1;
                        S.ijoint[nje++] = je.x;
                        if (je.y < 0) continue;
                        if (islb[je.y] == 0) {
                            islb[je.y] = 1;
                            stack[sp++] = je.y;
                        } else if (islb[je.y] == 2) {
                            islb[je.y] = 3;
                            stack[x.nb - 1 - nst++] = je.y;
                        }
                    }
                }
            }
            for (int q = 0; q < nst; ++q) islb[stack[x.nb - 1 - q]] = 2;
            maxB = (nbe - nbe0 + nst > maxB) ? nbe - nbe0 + nst : maxB;
            ++nisl;
        }
        S.istart[nisl] = int2{nbe, nce};
        S.jstart[nisl] = nje;
        s_cnt[0] = nisl;
        s_cnt[1] = maxB;
    }
    __syncthreads();
    maxBodies = s_cnt[1];
    const int n = s_cnt[0];
    __syncthreads();
    return n;
}


// Islands with everything in LDS (ldsAvail bytes): the per-body adjacency
// (entries packed as contact index << 16 | other body, for worlds under 2^16
// contacts and bodies), the DFS stack and the flags.  Only the lists of
// speedable bodies are built: static bodies are never expanded (the DFS only
// marks them).  Each list is sorted by contact index (the body's contact
// list order).  Same islands, order and counts as find_islands; returns -1,
// with nothing written outside LDS, when the world does not fit.
__device__ __forceinline__ int find_islands_lds(const DeviceState& d, const WorldCtx& x, int nc, const IslandScratch& S,
                                                int& maxBodies, int ldsAvail) {
    __shared__ int s_cnt[2];
    if (x.nb >= 0xffff || nc >= 0xffff) return -1;
    const int nb = x.nb;
    // off: nb + 1 ints; stack: nb ints (first the scatter cursors); entries.
    if (4 * (2 * nb + 1) > ldsAvail) return -1;
    int* off = reinterpret_cast<int*>(g_smem);
    int* stack = off + nb + 1;
    std::uint32_t* ent = reinterpret_cast<std::uint32_t*>(stack + nb);
    auto listed = [&](int b) { return (d.bflags[x.bo + b] & prx::kSpeedable) != 0; };
    for (int b = x.lane; b <= nb; b += kWave) off[b] = 0;
    __syncthreads();
    for (int i = x.lane; i < nc; i += kWave) {
        const std::uint32_t f = d.c_flags[x.co + i];
        if ((f & prx::kCEnabled) && (f & prx::kTouching) && !(f & prx::kSensor)) {
            const int2 bd = d.c_body[x.co + i];
            if (listed(bd.x)) atomicAdd(&off[bd.x], 1);
            if (listed(bd.y)) atomicAdd(&off[bd.y], 1);
        }
    }
    __syncthreads();
    {
        int carry = 0;
        for (int b0 = 0; b0 <= nb; b0 += kWave) {
            const int b = b0 + x.lane;
            const int v = (b <= nb) ? off[b] : 0;
            const int incl = wave_incl_scan(v, x.lane);
            if (b <= nb) off[b] = carry + incl - v;
            carry += __shfl(incl, kWave - 1);
        }
    }
    __syncthreads();
    const int nent = off[nb];
    std::uint8_t* islb = reinterpret_cast<std::uint8_t*>(ent + nent);
    std::uint8_t* islc = islb + nb;
    std::uint8_t* islj = islc + nc;
    if (4 * (2 * nb + 1 + nent) + nb + nc + x.nj > ldsAvail) return -1;
    for (int b = x.lane; b < nb; b += kWave) stack[b] = off[b];
    __syncthreads();
    for (int i = x.lane; i < nc; i += kWave) {
        const std::uint32_t f = d.c_flags[x.co + i];
        if ((f & prx::kCEnabled) && (f & prx::kTouching) && !(f & prx::kSensor)) {
            const int2 bd = d.c_body[x.co + i];
            if (listed(bd.x)) ent[atomicAdd(&stack[bd.x], 1)] = (std::uint32_t(i) << 16) | std::uint32_t(bd.y);
            if (listed(bd.y)) ent[atomicAdd(&stack[bd.y], 1)] = (std::uint32_t(i) << 16) | std::uint32_t(bd.x);
        }
    }
    __syncthreads();
    for (int b = x.lane; b < nb; b += kWave) {
        const int s0 = off[b], e0 = off[b + 1];
        for (int i = s0 + 1; i < e0; ++i) {
            const std::uint32_t v = ent[i];
            int j = i - 1;
            while (j >= s0 && ent[j] > v) {
                ent[j + 1] = ent[j];
                --j;
            }
            ent[j + 1] = v;
        }
        islb[b] = listed(b) ? 0 : 2;
        S.bisland[b] = -1;
    }
    for (int i = x.lane; i < nc; i += kWave) islc[i] = 0;
    for (int i = x.lane; i < x.nj; i += kWave) islj[i] = 0;
    __syncthreads();
    // One-body islands.
    int nsingle = 0;
    for (int base = 0; base < nb; base += kWave) {
        const int b = base + x.lane;
        bool single = false;
        if (b < nb) {
            const std::uint32_t f = d.bflags[x.bo + b];
            single = (f & prx::kAwake) && (f & prx::kEnabled) && (f & prx::kSpeedable) && off[b] == off[b + 1] &&
                     (x.nj == 0 || d.bjoint[x.bo + b].y == 0);
        }
        const unsigned long long m = __ballot(single);
        if (single) {
            const int r = nsingle + __popcll(m & lanes_below(x.lane));
            S.ibody[r] = b;
            S.istart[r] = int2{r, 0};
            S.jstart[r] = 0;
            S.bisland[b] = r;
            islb[b] = 1;
        }
        nsingle += __popcll(m);
    }
    __syncthreads();
    if (x.lane == 0) {
        int nisl = nsingle, nbe = nsingle, nce = 0, nje = 0;
        // Reached static bodies are marked 3 while counted and listed from
        // the top of the stack array (AabbTreeWorld.cpp:1386-1398).
        int maxB = nsingle > 0 ? 1 : 0;
        for (int seed = 0; seed < nb; ++seed) {
            if (islb[seed]) continue;
            const std::uint32_t sf = d.bflags[x.bo + seed];
            if (!((sf & prx::kAwake) && (sf & prx::kEnabled))) continue;
            S.istart[nisl] = int2{nbe, nce};
            S.jstart[nisl] = nje;
            const int nbe0 = nbe;
            int nst = 0;
            int sp = 0;
            stack[sp++] = seed;
            islb[seed] = 1;
            while (sp > 0) {
                const int bb = stack[--sp];
                S.ibody[nbe++] = bb;
                S.bisland[bb] = nisl;
                const int e1 = off[bb + 1];
                for (int e = off[bb]; e < e1; ++e) {
                    const std::uint32_t en = ent[e];
                    const int ci = int(en >> 16), ob = int(en & 0xffffu);
                    if (islc[ci]) continue;
                    S.icont[nce++] = ci;
                    islc[ci] = 1;
                    if (islb[ob] == 0) {
                        islb[ob] = 1;
                        stack[sp++] = ob;
                    } else if (islb[ob] == 2) {
                        islb[ob] = 3;
                        stack[nb - 1 - nst++] = ob;
                    }
                }
                // AddJointsToIsland (AabbTreeWorld.cpp:1446-1470).
                if (x.nj) {
                    const int2 bj = d.bjoint[x.bo + bb];
                    for (int k = 0; k < bj.y; ++k) {
                        const int2 je = d.bjoint_list[2 * x.jo + bj.x + k];
                        if (islj[je.x]) continue;
                        // A target joint has no other body (-1).
                        if (je.y >= 0 && !(d.bflags[x.bo + je.y] & prx::kEnabled)) continue;
                        islj[je.x] = 1;
                        S.ijoint[nje++] = je.x;
                        if (je.y < 0) continue;
                        if (islb[je.y] == 0) {
                            islb[je.y] = 1;
                            stack[sp++] = je.y;
                        } else if (islb[je.y] == 2) {
                            islb[je.y] = 3;
                            stack[nb - 1 - nst++] = je.y;
                        }
                    }
                }
            }
            for (int q = 0; q < nst; ++q) islb[stack[nb - 1 - q]] = 2;
            maxB = (nbe - nbe0 + nst > maxB) ? nbe - nbe0 + nst : maxB;
            ++nisl;
        }
        S.istart[nisl] = int2{nbe, nce};
        S.jstart[nisl] = nje;
        s_cnt[0] = nisl;
        s_cnt[1] = maxB;
    }
    __syncthreads();
    maxBodies = s_cnt[1];
    const int n = s_cnt[0];
    __syncthreads();
    return n;
}

// Position-constraint record (see the layout above).
struct Pc {
    Manifold m;
    int bA, bB;
    float totalRadius;
    BodyC A, B;
};

__device__ __forceinline__ void store_pc(float4* u, const Manifold& m, int bA, int bB, float tr, const BodyC& A,
                                         const BodyC& B) {
    u[0] = float4{__uint_as_float(m.type | (m.n << 8)), __int_as_float(bA), __int_as_float(bB), tr};
    u[1] = float4{m.ln.x, m.ln.y, m.lpt.x, m.lpt.y};
    u[2] = float4{m.p[0].lp.x, m.p[0].lp.y, m.p[1].lp.x, m.p[1].lp.y};
    u[3] = float4{A.invMass, A.invRotI, A.lc.x, A.lc.y};
    u[4] = float4{B.invMass, B.invRotI, B.lc.x, B.lc.y};
}
__device__ __forceinline__ Pc load_pc(const float4* u) {
    const float4 a = u[0], g = u[1], q = u[2], ma = u[3], mb = u[4];
    Pc pc;
    const std::uint32_t h = __float_as_uint(a.x);
    pc.m.type = h & 0xffu;
    pc.m.n = (h >> 8) & 0xffu;
    pc.m.ln = V2{g.x, g.y};
    pc.m.lpt = V2{g.z, g.w};
    pc.m.p[0] = MPoint{V2{q.x, q.y}, 0u, 0.0f, 0.0f};
    pc.m.p[1] = MPoint{V2{q.z, q.w}, 0u, 0.0f, 0.0f};
    pc.bA = __float_as_int(a.y);
    pc.bB = __float_as_int(a.z);
    pc.totalRadius = a.w;
    pc.A = BodyC{ma.x, ma.y, V2{ma.z, ma.w}};
    pc.B = BodyC{mb.x, mb.y, V2{mb.z, mb.w}};
    return pc;
}

// Gauss-Seidel sweeps of one island.  Two forms with identical arithmetic:
//
// * lane form (island_velocity / island_position): one lane walks one
//   island's records, the next record loaded while the current one is
//   solved; many small islands proceed side by side on the wave's lanes;
// * wave form (island_velocity_wave / island_position_wave): the whole wave
//   serves one large island.  Records are staged through LDS in chunks of
//   kChunk; all lanes load the next chunk from global memory (coalesced) while
//   lane 0 solves the current chunk out of LDS, so the serial chain waits on
//   LDS, not on global memory.
//
// Velocity-record impulses are written back to global memory chunk by chunk
// (every later read of a record comes after its write-back in the wave's
// program order).

// Warm start of one velocity constraint (WarmStartVelocities,
// AabbTreeWorld.cpp:230-274).
template <bool kLds>
__device__ __forceinline__ void warm_start_one(const Vc& vc, const BodyC& A, const BodyC& B, const Bodies<kLds>& bs) {
    const V2 tangent = fwdperp(vc.normal);
    V2 dAl{0, 0}, dBl{0, 0};
    float dAa = 0.0f, dBa = 0.0f;
#pragma unroll
    for (std::uint32_t j = 0; j < 2; ++j) {
        if (j >= vc.n) continue;
        const VcPoint& p = vc.p[j];
        const V2 P = mul(p.ni, vc.normal) + mul(p.ti, tangent);
        const float LA = cross(p.relA, P) / 1.0f;
        const float LB = cross(p.relB, P) / 1.0f;
        dAl = dAl - mul(A.invMass, P);
        dAa = dAa - A.invRotI * LA;
        dBl = dBl + mul(B.invMass, P);
        dBa = dBa + B.invRotI * LB;
    }
    float4 va = bs.vel(vc.bA);
    va.x = va.x + dAl.x;
    va.y = va.y + dAl.y;
    va.z = va.z + dAa;
    bs.set_vel(vc.bA, va);
    float4 vb = bs.vel(vc.bB);
    vb.x = vb.x + dBl.x;
    vb.y = vb.y + dBl.y;
    vb.z = vb.z + dBa;
    bs.set_vel(vc.bB, vb);
}

// GaussSeidel::SolveVelocityConstraint (ContactSolver.cpp:481-492): tangent
// first, then normal (block solve for two points).  Returns the largest
// impulse increment.
template <bool kLds>
__device__ __forceinline__ float solve_velocity_one(Vc& vc, const BodyC& A, const BodyC& B, const Bodies<kLds>& bs) {
    float4 va = bs.vel(vc.bA);
    float4 vb = bs.vel(vc.bB);
    float mi = 0.0f;
    mi = fmax_std(mi, solve_dir(vc, true, A, B, va, vb));
    float nrm;
    if (vc.n == 1 || (vc.K00 == 0.0f && vc.K11 == 0.0f && vc.K01 == 0.0f)) nrm = solve_dir(vc, false, A, B, va, vb);
    else nrm = block_solve(vc, A, B, va, vb);
    mi = fmax_std(mi, nrm);
    bs.set_vel(vc.bA, va);
    bs.set_vel(vc.bB, vb);
    return mi;
}

// One pass over an island's joints in island order: the warm start of
// InitVelocity, or SolveVelocity (returns whether every joint left its
// bodies' velocities unchanged).
template <bool kLds, bool kKinds>
__device__ __forceinline__ bool joints_velocity_k(float4* jbase, int j0, int j1, const Bodies<kLds>& bs, bool warm) {
    bool ok = true;
    for (int e = j0; e < j1; ++e) {
        float4* u = jbase + std::size_t(e) * prx::kJcStride;
        const std::uint32_t kind = kKinds ? jc_kind(u) : prx::kJointRevolute;
        if (kind == prx::kJointGear) {
            // Four bodies, stored back in the order A, B, C, D (a body that is
            // two of them keeps the last store, as the reference's does).
            const int bA = __float_as_int(u[5].z), bB = __float_as_int(u[5].w);
            const int bC = __float_as_int(u[6].z), bD = __float_as_int(u[6].w);
            float4 va = bs.vel(bA), vb = bs.vel(bB), vc = bs.vel(bC), vd = bs.vel(bD);
            if (warm) gear_apply(u, u[3].x, va, vb, vc, vd);
            else ok &= gear_solve_velocity(u, va, vb, vc, vd);
            bs.set_vel(bA, va);
            bs.set_vel(bB, vb);
            bs.set_vel(bC, vc);
            bs.set_vel(bD, vd);
            continue;
        }
        if (kind != prx::kJointRevolute) {
            const float4 mo = u[5];
            const int bA = __float_as_int(mo.z), bB = __float_as_int(mo.w);
            float4 va = bs.vel(bA);
            float4 vb = bs.vel(bB);
            if (warm) spring_joint_warm_start(u, kind, va, vb);
            else if (kind == prx::kJointDistance) ok &= distance_solve_velocity(u, va, vb);
            else if (kind == prx::kJointWeld) ok &= weld_solve_velocity(u, va, vb);
            else if (kind == prx::kJointRope) ok &= rope_solve_velocity(u, va, vb);
            else if (kind == prx::kJointFriction) ok &= friction_solve_velocity(u, va, vb);
            else if (kind == prx::kJointWheel) ok &= wheel_solve_velocity(u, va, vb);
            else if (kind == prx::kJointPulley) ok &= pulley_solve_velocity(u, va, vb);
            else if (kind == prx::kJointPrismatic) ok &= prismatic_solve_velocity(u, va, vb);
            else if (kind == prx::kJointTarget) ok &= target_solve_velocity(u, vb);
            else ok &= motor_solve_velocity(u, va, vb);
            bs.set_vel(bA, va);
            bs.set_vel(bB, vb);
            continue;
        }
        Jc j = load_jc_velocity(u);
        float4 va = bs.vel(j.bA);
        float4 vb = bs.vel(j.bB);
        if (warm) {
            joint_warm_start(j, va, vb);
        } else {
            ok &= joint_solve_velocity(j, va, vb);
            u[3] = float4{j.ix, j.iy, j.iz, j.motorImp};
        }
        bs.set_vel(j.bA, va);
        bs.set_vel(j.bB, vb);
    }
    return ok;
}

// Without warm starting, InitVelocity still damps a target joint's body's spin
// (TargetJointConf.cpp:125-126): the joints' part of the warm-start pass that
// remains then.
template <bool kLds>
__device__ __forceinline__ void joints_target_damping(const float4* jbase, int j0, int j1, const Bodies<kLds>& bs) {
    for (int e = j0; e < j1; ++e) {
        const float4* u = jbase + std::size_t(e) * prx::kJcStride;
        if (jc_kind(u) != prx::kJointTarget) continue;
        const int b = __float_as_int(u[5].w);
        float4 v = bs.vel(b);
        v.z = v.z * 0.98f;
        bs.set_vel(b, v);
    }
}

// Batches whose joints are all revolute take the loop without the kind
// dispatch (StepParams::jointKinds 0).
template <bool kLds>
__device__ __forceinline__ bool joints_velocity(float4* jbase, int j0, int j1, const Bodies<kLds>& bs, bool warm,
                                                const StepParams& c) {
    return c.jointKinds ? joints_velocity_k<kLds, true>(jbase, j0, j1, bs, warm)
                        : joints_velocity_k<kLds, false>(jbase, j0, j1, bs, warm);
}

// The same pass recording each joint's inputs (velocities of both bodies
// and the impulses before the update) for the adjoint.
template <bool kLds>
__device__ __forceinline__ bool joints_velocity_tape(float4* jbase, int j0, int j1, const Bodies<kLds>& bs,
                                                     const prx::Tape& t, int jo, int it) {
    bool ok = true;
    for (int e = j0; e < j1; ++e) {
        float4* u = jbase + std::size_t(e) * prx::kJcStride;
        Jc j = load_jc_velocity(u);
        float4 va = bs.vel(j.bA);
        float4 vb = bs.vel(j.bB);
        float4* g = t.jgs + ((std::size_t(jo) + std::size_t(e)) * std::size_t(t.itV) + std::size_t(it)) * 3;
        g[0] = va;
        g[1] = vb;
        g[2] = float4{j.ix, j.iy, j.iz, j.motorImp};
        ok &= joint_solve_velocity(j, va, vb);
        u[3] = float4{j.ix, j.iy, j.iz, j.motorImp};
        bs.set_vel(j.bA, va);
        bs.set_vel(j.bB, vb);
    }
    return ok;
}

// One SolvePosition pass over an island's joints; returns whether all are
// within tolerance.
template <bool kLds, bool kKinds>
__device__ __forceinline__ bool joints_position_k(const float4* jbase, int j0, int j1, const Bodies<kLds>& bs,
                                                  const StepParams& c) {
    bool ok = true;
    for (int e = j0; e < j1; ++e) {
        const float4* u = jbase + std::size_t(e) * prx::kJcStride;
        const float4 mo = u[5];
        const int bA = __float_as_int(mo.z), bB = __float_as_int(mo.w);
        const std::uint32_t kind = kKinds ? jc_kind(u) : prx::kJointRevolute;
        if (kind == prx::kJointGear) {
            const int bC = __float_as_int(u[6].z), bD = __float_as_int(u[6].w);
            float4 pa = bs.pos(bA), pb = bs.pos(bB), pc = bs.pos(bC), pd = bs.pos(bD);
            ok &= gear_solve_position(u, pa, pb, pc, pd, c);
            bs.set_pos(bA, pa);
            bs.set_pos(bB, pb);
            bs.set_pos(bC, pc);
            bs.set_pos(bD, pd);
            continue;
        }
        float4 pa = bs.pos(bA);
        float4 pb = bs.pos(bB);
        if (kind == prx::kJointRevolute) ok &= joint_solve_position(u, pa, pb, c);
        else ok &= spring_joint_solve_position(u, kind, pa, pb, c);
        bs.set_pos(bA, pa);
        bs.set_pos(bB, pb);
    }
    return ok;
}
template <bool kLds>
__device__ __forceinline__ bool joints_position(const float4* jbase, int j0, int j1, const Bodies<kLds>& bs,
                                                const StepParams& c) {
    return c.jointKinds ? joints_position_k<kLds, true>(jbase, j0, j1, bs, c)
                        : joints_position_k<kLds, false>(jbase, j0, j1, bs, c);
}

// The same pass recording the positions before each update.
template <bool kLds>
__device__ __forceinline__ bool joints_position_tape(const float4* jbase, int j0, int j1, const Bodies<kLds>& bs,
                                                     const StepParams& c, const prx::Tape& t, int jo, int it) {
    bool ok = true;
    for (int e = j0; e < j1; ++e) {
        const float4* u = jbase + std::size_t(e) * prx::kJcStride;
        const float4 mo = u[5];
        const int bA = __float_as_int(mo.z), bB = __float_as_int(mo.w);
        float4 pa = bs.pos(bA);
        float4 pb = bs.pos(bB);
        float4* g = t.jngs + ((std::size_t(jo) + std::size_t(e)) * std::size_t(t.itP) + std::size_t(it)) * 2;
        g[0] = pa;
        g[1] = pb;
        ok &= joint_solve_position(u, pa, pb, c);
        bs.set_pos(bA, pa);
        bs.set_pos(bB, pb);
    }
    return ok;
}

// Lane form of the warm start and velocity sweeps of one island
// (AabbTreeWorld.cpp:1590-1625): contact warm start, joint warm start, then
// per iteration the joints followed by the contacts.
// What one island's sweeps report for the step statistics (IslandStats,
// StepStats.hpp:149-163): the largest incremental impulse of any velocity
// iteration and the iterations run; the smallest separation of any position
// iteration and the iteration that solved the island plus one (0: not solved).
struct VelResult {
    float maxInc;
    int iters;
};
struct PosResult {
    float minSep;
    int solvedAt;
};

template <bool kLds, bool kTape = false>
__device__ __forceinline__ VelResult island_velocity(const DeviceState& d, const WorldCtx& x, const Bodies<kLds>& bs,
                                                     int c0, int c1, int j0, int j1, const StepParams& c,
                                                     int islandSlot = 0) {
    VelResult res{0.0f, int(c.velIters)};
    if constexpr (kTape) d.tape.isl[islandSlot].z = 0;
    if (c0 >= c1 && j0 >= j1) {
        // Nothing to solve: the first iteration's increment is zero.
        if (c.velIters > 0 && 0.0f <= c.regMinMomentum) res.iters = 1;
        return res;
    }
    float4* base = d.s_vc + std::size_t(x.co) * prx::kVcStride;
    float4* jbase = d.s_jc + std::size_t(x.jo) * prx::kJcStride;
    if (c.warmStart && c0 < c1) {
        BodyC A, B;
        Vc vc = load_vc(base + std::size_t(c0) * prx::kVcStride, A, B);
        for (int e = c0; e < c1; ++e) {
            const int en = (e + 1 < c1) ? e + 1 : e;
            BodyC nA, nB;
            const Vc nvc = load_vc(base + std::size_t(en) * prx::kVcStride, nA, nB);
            warm_start_one(vc, A, B, bs);
            vc = nvc;
            A = nA;
            B = nB;
        }
    }
    if (c.warmStart) joints_velocity(jbase, j0, j1, bs, true, c);
    else if (c.jointKinds) joints_target_damping(jbase, j0, j1, bs);
    for (std::uint32_t it = 0; it < c.velIters; ++it) {
        if constexpr (kTape) d.tape.isl[islandSlot].z = int(it) + 1;
        bool jointsOkay;
        if constexpr (kTape) jointsOkay = joints_velocity_tape(jbase, j0, j1, bs, d.tape, x.jo, int(it));
        else jointsOkay = joints_velocity(jbase, j0, j1, bs, false, c);
        float newInc = 0.0f;
        if (c0 >= c1) {
            if (jointsOkay && newInc <= c.regMinMomentum) {
                res.iters = int(it) + 1;
                break;
            }
            continue;
        }
        BodyC A, B;
        Vc vc = load_vc(base + std::size_t(c0) * prx::kVcStride, A, B);
        for (int e = c0; e < c1; ++e) {
            const int en = (e + 1 < c1) ? e + 1 : e;
            BodyC nA, nB;
            const Vc nvc = load_vc(base + std::size_t(en) * prx::kVcStride, nA, nB);
            if constexpr (kTape) {
                const float4 va = bs.vel(vc.bA), vb = bs.vel(vc.bB);
                float4* g = d.tape.gs + ((std::size_t(x.co) + std::size_t(e)) * std::size_t(d.tape.itV) + it) * 3;
                g[0] = float4{va.x, va.y, va.z, vc.p[0].ti};
                g[1] = float4{vb.x, vb.y, vb.z, vc.p[0].ni};
                g[2] = float4{vc.p[0].ni, vc.p[0].ti, vc.p[1].ni, vc.p[1].ti};
            }
            const float mi = solve_velocity_one(vc, A, B, bs);
            float4* rec = base + std::size_t(e) * prx::kVcStride;
            rec[4] = float4{vc.p[0].ni, vc.p[0].ti, vc.p[0].nm, vc.p[0].tm};
            rec[7] = float4{vc.p[1].ni, vc.p[1].ti, vc.p[1].nm, vc.p[1].tm};
            newInc = fmax_std(newInc, mi);
            vc = nvc;
            A = nA;
            B = nB;
        }
        res.maxInc = fmax_std(res.maxInc, newInc);
        if (jointsOkay && newInc <= c.regMinMomentum) {
            res.iters = int(it) + 1;
            break;
        }
    }
    return res;
}

// Chunk staging of kStride-float4 records [c0, c1) of a global array through
// registers into the LDS stage: fetch issues the loads, commit stores them.
template <int kStride>
struct Stager {
    static constexpr int kFloats = kChunk * kStride;
    static constexpr int kPer = (kFloats + kWave - 1) / kWave;
    const float4* src;  // first record of the island
    float4* stage;
    int n;              // records in the island
    int lane;
    float4 reg[kPer];
    __device__ __forceinline__ int limit(int ch) const {
        const int e = (ch + 1) * kChunk;
        return ((e < n) ? e : n) * kStride - ch * kFloats;
    }
    __device__ __forceinline__ void fetch(int ch) {
        const int lim = limit(ch);
        const float4* s = src + std::size_t(ch) * kFloats;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int f = lane + k * kWave;
            reg[k] = (f < lim) ? s[f] : float4{0.0f, 0.0f, 0.0f, 0.0f};
        }
    }
    __device__ __forceinline__ void commit(int ch) {
        const int lim = limit(ch);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int f = lane + k * kWave;
            if (f < lim) stage[f] = reg[k];
        }
    }
};

template <bool kLds>
__device__ __forceinline__ VelResult island_velocity_wave(const DeviceState& d, const WorldCtx& x,
                                                          const Bodies<kLds>& bs, float4* stage, int c0, int c1,
                                                          int j0, int j1, const StepParams& c) {
    VelResult res{0.0f, int(c.velIters)};
    float4* jbase = d.s_jc + std::size_t(x.jo) * prx::kJcStride;
    float4* gbase = d.s_vc + (std::size_t(x.co) + std::size_t(c0)) * prx::kVcStride;
    const int n = c1 - c0;
    const int nch = (n + kChunk - 1) / kChunk;
    const bool resident = nch == 1;  // the whole island stays in the stage
    Stager<prx::kVcStride> sg{gbase, stage, n, x.lane, {}};
    sg.fetch(0);
    sg.commit(0);
    __syncthreads();
    const int npass = int(c.velIters) + (c.warmStart ? 1 : 0);
    if (!c.warmStart && c.jointKinds && x.lane == 0) joints_target_damping(jbase, j0, j1, bs);
    for (int pass = 0; pass < npass; ++pass) {
        const bool warm = c.warmStart && pass == 0;
        float newInc = 0.0f;
        bool jointsOkay = true;
        if (!warm && x.lane == 0) jointsOkay = joints_velocity(jbase, j0, j1, bs, false, c);
        for (int ch = 0; ch < nch; ++ch) {
            const int nx = (ch + 1 < nch) ? ch + 1 : 0;
            const bool more = !resident && (ch + 1 < nch || pass + 1 < npass);
            if (more) sg.fetch(nx);
            const int r1 = ((ch + 1) * kChunk < n ? kChunk : n - ch * kChunk);
            if (x.lane == 0) {
                BodyC A, B;
                Vc vc = load_vc(stage, A, B);
                for (int r = 0; r < r1; ++r) {
                    const int rn = (r + 1 < r1) ? r + 1 : r;
                    BodyC nA, nB;
                    const Vc nvc = load_vc(stage + rn * prx::kVcStride, nA, nB);
                    if (warm) {
                        warm_start_one(vc, A, B, bs);
                    } else {
                        const float mi = solve_velocity_one(vc, A, B, bs);
                        float4* rec = stage + r * prx::kVcStride;
                        rec[4] = float4{vc.p[0].ni, vc.p[0].ti, vc.p[0].nm, vc.p[0].tm};
                        rec[7] = float4{vc.p[1].ni, vc.p[1].ti, vc.p[1].nm, vc.p[1].tm};
                        newInc = fmax_std(newInc, mi);
                    }
                    vc = nvc;
                    A = nA;
                    B = nB;
                }
            }
            __syncthreads();
            if (!warm && !resident && x.lane < 2 * r1) {
                const int r = x.lane >> 1;
                const int f = r * prx::kVcStride + ((x.lane & 1) ? 7 : 4);
                gbase[std::size_t(ch) * kChunk * prx::kVcStride + f] = stage[f];
            }
            if (more) {
                __syncthreads();
                sg.commit(nx);
                __syncthreads();
            }
        }
        if (warm && x.lane == 0) joints_velocity(jbase, j0, j1, bs, true, c);
        newInc = __shfl(newInc, 0);
        jointsOkay = __shfl(int(jointsOkay), 0) != 0;
        if (!warm) {
            res.maxInc = fmax_std(res.maxInc, newInc);
            if (jointsOkay && newInc <= c.regMinMomentum) {
                res.iters = pass - (c.warmStart ? 1 : 0) + 1;
                break;
            }
        }
    }
    if (resident && x.lane < 2 * n) {
        const int r = x.lane >> 1;
        const int f = r * prx::kVcStride + ((x.lane & 1) ? 7 : 4);
        gbase[f] = stage[f];
    }
    __syncthreads();
    return res;
}

// Lane form of the position sweeps of one island (ContactSolver.cpp:494-643,
// AabbTreeWorld.cpp:1627-1644).  Returns whether the island was solved.
template <bool kLds, bool kTape = false>
__device__ __forceinline__ PosResult island_position(const DeviceState& d, const WorldCtx& x, const Bodies<kLds>& bs,
                                                int c0, int c1, int j0, int j1, const StepParams& c,
                                                int islandSlot = 0) {
    const float4* base = d.s_pc + std::size_t(x.co) * prx::kPcStride;
    const float4* jbase = d.s_jc + std::size_t(x.jo) * prx::kJcStride;
    PosResult res{__builtin_inff(), 0};
    if constexpr (kTape) d.tape.isl[islandSlot].w = 0;
    for (std::uint32_t it = 0; it < c.posIters; ++it) {
        if constexpr (kTape) d.tape.isl[islandSlot].w = int(it) + 1;
        float minSep = __builtin_inff();
        if (c0 < c1) {
            Pc pc = load_pc(base + std::size_t(c0) * prx::kPcStride);
            for (int e = c0; e < c1; ++e) {
                const int en = (e + 1 < c1) ? e + 1 : e;
                const Pc npc = load_pc(base + std::size_t(en) * prx::kPcStride);
                float4 pa = bs.pos(pc.bA);
                float4 pb = bs.pos(pc.bB);
                if constexpr (kTape) {
                    float4* g = d.tape.ngs + ((std::size_t(x.co) + std::size_t(e)) * std::size_t(d.tape.itP) + it) * 2;
                    g[0] = pa;
                    g[1] = pb;
                }
                const float sep = solve_pc(pc.m, pc.totalRadius, pc.A, pc.B, pa, pb, c);
                if constexpr (kTape) {
                    float4* g = d.tape.ngs + ((std::size_t(x.co) + std::size_t(e)) * std::size_t(d.tape.itP) + it) * 2;
                    g[0].w = sep;
                }
                bs.set_pos(pc.bA, pa);
                bs.set_pos(pc.bB, pb);
                minSep = fmin_std(minSep, sep);
                pc = npc;
            }
        }
        res.minSep = fmin_std(res.minSep, minSep);
        const bool contactsOkay = minSep >= c.regMinSeparation;
        bool jointsOkay;
        if constexpr (kTape) jointsOkay = joints_position_tape(jbase, j0, j1, bs, c, d.tape, x.jo, int(it));
        else jointsOkay = joints_position(jbase, j0, j1, bs, c);
        if (contactsOkay && jointsOkay) {
            res.solvedAt = int(it) + 1;
            return res;
        }
    }
    return res;
}

template <bool kLds>
__device__ __forceinline__ PosResult island_position_wave(const DeviceState& d, const WorldCtx& x, const Bodies<kLds>& bs,
                                                     float4* stage, int c0, int c1, int j0, int j1,
                                                     const StepParams& c) {
    const float4* jbase = d.s_jc + std::size_t(x.jo) * prx::kJcStride;
    const float4* gbase = d.s_pc + (std::size_t(x.co) + std::size_t(c0)) * prx::kPcStride;
    const int n = c1 - c0;
    const int nch = (n + kChunk - 1) / kChunk;
    const bool resident = nch == 1;
    Stager<prx::kPcStride> sg{gbase, stage, n, x.lane, {}};
    sg.fetch(0);
    sg.commit(0);
    __syncthreads();
    const int npass = int(c.posIters);
    PosResult res{__builtin_inff(), 0};
    for (int pass = 0; pass < npass; ++pass) {
        float minSep = __builtin_inff();
        for (int ch = 0; ch < nch; ++ch) {
            const int nx = (ch + 1 < nch) ? ch + 1 : 0;
            const bool more = !resident && (ch + 1 < nch || pass + 1 < npass);
            if (more) sg.fetch(nx);
            const int r1 = ((ch + 1) * kChunk < n ? kChunk : n - ch * kChunk);
            if (x.lane == 0) {
                Pc pc = load_pc(stage);
                for (int r = 0; r < r1; ++r) {
                    const int rn = (r + 1 < r1) ? r + 1 : r;
                    const Pc npc = load_pc(stage + rn * prx::kPcStride);
                    float4 pa = bs.pos(pc.bA);
                    float4 pb = bs.pos(pc.bB);
                    const float sep = solve_pc(pc.m, pc.totalRadius, pc.A, pc.B, pa, pb, c);
                    bs.set_pos(pc.bA, pa);
                    bs.set_pos(pc.bB, pb);
                    minSep = fmin_std(minSep, sep);
                    pc = npc;
                }
            }
            __syncthreads();
            if (more) {
                sg.commit(nx);
                __syncthreads();
            }
        }
        bool jointsOkay = true;
        if (x.lane == 0) jointsOkay = joints_position(jbase, j0, j1, bs, c);
        minSep = __shfl(minSep, 0);
        jointsOkay = __shfl(int(jointsOkay), 0) != 0;
        res.minSep = fmin_std(res.minSep, minSep);
        if (minSep >= c.regMinSeparation && jointsOkay) {
            res.solvedAt = pass + 1;
            break;
        }
    }
    __syncthreads();
    return res;
}

// ---------------------------------------------------------------------------
// Paired-lane forms of the contact updates for the level form: lane 2j
// (side 0) computes update j's body-A terms and lane 2j + 1 (side 1) its
// body-B terms, exchanging what the other body contributes through a DPP
// swap within the pair; terms of both bodies (relative velocity, impulses,
// block-solver cases, position-solver manifolds) are computed by both lanes
// from the same values.  Body A's changes are the negated body-B form
// (a - b*c = a + (-1)*(b*c) exactly), so both lanes run one instruction
// stream and every result equals the single-lane form bit for bit.
__device__ __forceinline__ float pair_swap(float v) {
    // quad_perm [1, 0, 3, 2]: each lane reads its pair partner.
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ V2 pair_swap(V2 v) { return V2{pair_swap(v.x), pair_swap(v.y)}; }

// GetContactRelVelocity (d2/Math.cpp:73-103) from the pair's two halves.
__device__ __forceinline__ V2 pair_rel_vel(V2 v, float w, V2 rel, int side) {
    const V2 u = v + mul(revperp(rel), w / 1.0f);
    const V2 o = pair_swap(u);
    const V2 uA = side ? o : u;
    const V2 uB = side ? u : o;
    return uB - uA;
}

struct PairBody {
    float m, i, sgn;  // inverse mass, inverse rotational inertia, -1 for body A
};

__device__ __forceinline__ void pair_apply(V2& v, float& w, const PairBody& b, V2 P, float L) {
    const V2 mp = mul(b.m, P);
    v = V2{v.x + b.sgn * mp.x, v.y + b.sgn * mp.y};
    w = w + b.sgn * (b.i * L);
}

__device__ __forceinline__ float solve_dir_pair(Vc& vc, bool tangent, const PairBody& me, int side, V2& v, float& w) {
    float maxInc = 0.0f;
    const V2 dir = tangent ? fwdperp(vc.normal) : vc.normal;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
        const int k = 1 - kk;  // point 1 first, then point 0
        if (k >= int(vc.n)) continue;
        VcPoint& p = vc.p[k];
        const V2 rel = side ? p.relB : p.relA;
        const V2 closing = pair_rel_vel(v, w, rel, side);
        float oldI, newI;
        if (tangent) {
            const float dirVel = vc.tangentSpeed - dot(closing, dir);
            const float lambda = p.tm * dirVel;
            const float maxI = vc.friction * p.ni;
            oldI = p.ti;
            newI = clamp_std(oldI + lambda, -maxI, maxI);
        } else {
            const float dirVel = dot(closing, dir);
            const float lambda = p.nm * (p.bias - dirVel);
            oldI = p.ni;
            newI = fmax_std(oldI + lambda, 0.0f);
        }
        const float inc = newI - oldI;
        const V2 P = mul(inc, dir);
        pair_apply(v, w, me, P, cross(rel, P) / 1.0f);
        maxInc = fmax_std(maxInc, __builtin_fabsf(inc));
        if (tangent) p.ti = oldI + inc;
        else p.ni = oldI + inc;
    }
    return maxInc;
}

__device__ __forceinline__ float block_solve_pair(Vc& vc, const PairBody& me, int side, V2& v, float& w) {
    const V2 normal = vc.normal;
    const V2 rel0 = side ? vc.p[0].relB : vc.p[0].relA;
    const V2 rel1 = side ? vc.p[1].relB : vc.p[1].relA;
    V2 bp;
    {
        const V2 dv0 = pair_rel_vel(v, w, rel0, side);
        const V2 dv1 = pair_rel_vel(v, w, rel1, side);
        const float vn0 = dot(dv0, normal);
        const float vn1 = dot(dv1, normal);
        const V2 b{vn0 - vc.p[0].bias, vn1 - vc.p[1].bias};
        float t0 = 0.0f;
        t0 += vc.K00 * vc.p[0].ni;
        t0 += vc.K01 * vc.p[1].ni;
        float t1 = 0.0f;
        t1 += vc.K01 * vc.p[0].ni;
        t1 += vc.K11 * vc.p[1].ni;
        bp = V2{b.x - t0, b.y - t1};
    }
    float x0 = 0.0f;
    x0 += vc.M00 * bp.x;
    x0 += vc.M01 * bp.y;
    float x1 = 0.0f;
    x1 += vc.M01 * bp.x;
    x1 += vc.M11 * bp.y;
    const float c1n0 = -x0, c1n1 = -x1;
    const bool ok1 = c1n0 >= 0.0f && c1n1 >= 0.0f;
    const float c2n0 = -vc.p[0].nm * bp.x;
    const float vn2 = vc.K01 * c2n0 + bp.y;
    const bool ok2 = c2n0 >= 0.0f && vn2 >= 0.0f;
    const float c3n1 = -vc.p[1].nm * bp.y;
    const float vn1 = vc.K01 * c3n1 + bp.x;
    const bool ok3 = c3n1 >= 0.0f && vn1 >= 0.0f;
    const bool ok4 = bp.x >= 0.0f && bp.y >= 0.0f;
    const float n0 = ok1 ? c1n0 : (ok2 ? c2n0 : 0.0f);
    const float n1 = ok1 ? c1n1 : (ok2 ? 0.0f : (ok3 ? c3n1 : 0.0f));
    const bool ok = ok1 || ok2 || ok3 || ok4;
    const float d0 = n0 - vc.p[0].ni;
    const float d1 = n1 - vc.p[1].ni;
    const V2 P0 = mul(d0, normal);
    const V2 P1 = mul(d1, normal);
    const V2 P = P0 + P1;
    const float L = (cross(rel0, P0) + cross(rel1, P1)) / 1.0f;
    V2 nv = v;
    float nw = w;
    pair_apply(nv, nw, me, P, L);
    v = ok ? nv : v;
    w = ok ? nw : w;
    vc.p[0].ni = ok ? n0 : vc.p[0].ni;
    vc.p[1].ni = ok ? n1 : vc.p[1].ni;
    return ok ? fmax_std(__builtin_fabsf(n0), __builtin_fabsf(n1)) : 0.0f;
}

template <bool kLds>
__device__ __forceinline__ float solve_velocity_pair(Vc& vc, const BodyC& A, const BodyC& B, const Bodies<kLds>& bs,
                                                     int side) {
    const int b = side ? vc.bB : vc.bA;
    const PairBody me{side ? B.invMass : A.invMass, side ? B.invRotI : A.invRotI, side ? 1.0f : -1.0f};
    const float4 vel = bs.vel(b);
    V2 v{vel.x, vel.y};
    float w = vel.z;
    float mi = 0.0f;
    mi = fmax_std(mi, solve_dir_pair(vc, true, me, side, v, w));
    float nrm;
    if (vc.n == 1 || (vc.K00 == 0.0f && vc.K11 == 0.0f && vc.K01 == 0.0f)) nrm = solve_dir_pair(vc, false, me, side, v, w);
    else nrm = block_solve_pair(vc, me, side, v, w);
    mi = fmax_std(mi, nrm);
    bs.set_vel(b, float4{v.x, v.y, w, vel.w});
    return mi;
}

template <bool kLds>
__device__ __forceinline__ void warm_start_pair(const Vc& vc, const BodyC& A, const BodyC& B, const Bodies<kLds>& bs,
                                                int side) {
    const int b = side ? vc.bB : vc.bA;
    const PairBody me{side ? B.invMass : A.invMass, side ? B.invRotI : A.invRotI, side ? 1.0f : -1.0f};
    const V2 tangent = fwdperp(vc.normal);
    V2 dl{0, 0};
    float da = 0.0f;
#pragma unroll
    for (std::uint32_t j = 0; j < 2; ++j) {
        if (j >= vc.n) continue;
        const VcPoint& p = vc.p[j];
        const V2 P = mul(p.ni, vc.normal) + mul(p.ti, tangent);
        pair_apply(dl, da, me, P, cross(side ? p.relB : p.relA, P) / 1.0f);
    }
    float4 vel = bs.vel(b);
    vel.x = vel.x + dl.x;
    vel.y = vel.y + dl.y;
    vel.z = vel.z + da;
    bs.set_vel(b, vel);
}

// GaussSeidel::SolvePositionConstraint (ContactSolver.cpp:494-643) in the
// paired form of solve_pc: each lane computes its own body's transformation
// (the sine and cosine of its angle) and position change.
__device__ __forceinline__ float solve_pc_pair(const Manifold& m, float totalRadius, const BodyC& A, const BodyC& B,
                                               float4& pos, int side, const StepParams& c) {
    const float invMassTotal = A.invMass + B.invMass;
    const BodyC& me = side ? B : A;
    const float sgn = side ? 1.0f : -1.0f;
    auto xfs = [&](V2 pl, float pa, Xf& xfA, Xf& xfB) {
        const Xf mine = get_xf(pl.x, pl.y, pa, me.lc);
        const Xf other{pair_swap(mine.p), pair_swap(mine.q)};
        xfA = side ? other : mine;
        xfB = side ? mine : other;
    };
    struct Sol {
        V2 dl;
        float da, sep;
    };
    auto solve = [&](const PSM& psm, V2 p) {
        const float separation = psm.sep - totalRadius;
        const V2 r = psm.point - p;
        const float cr = cross(r, psm.normal) / 1.0f;
        const float co = pair_swap(cr);
        const float crA = side ? co : cr;
        const float crB = side ? cr : co;
        const float K = invMassTotal + A.invRotI * (crA * crA) + B.invRotI * (crB * crB);
        const float C = -clamp_std(c.regResolutionRate * (separation + c.linearSlop), -c.maxLinearCorrection, 0.0f);
        const V2 P = divs(mul(psm.normal, C), K);
        const float L = cross(r, P) / 1.0f;
        const V2 mp = mul(me.invMass, P);
        return Sol{V2{sgn * mp.x, sgn * mp.y}, sgn * (me.invRotI * L), separation};
    };
    const V2 op{pos.x, pos.y};
    const float oa = pos.z;
    Xf xfA, xfB;
    xfs(op, oa, xfA, xfB);
    const PSM psm0 = get_psm(m, 0, xfA, xfB);
    const PSM psm1 = get_psm(m, 1, xfA, xfB);
    const bool two = m.n == 2;
    const bool eq = two && almost_equal(psm0.sep, psm1.sep);
    const bool lt = two && !eq && psm0.sep < psm1.sep;
    const bool gt = two && !eq && psm1.sep < psm0.sep;
    if (!(m.n == 1 || eq || lt || gt)) return __builtin_inff();
    const Sol s0 = solve(gt ? psm1 : psm0, op);
    V2 p = op + s0.dl;
    float a = oa + s0.da;
    const float minSep = two ? s0.sep : 0.0f + s0.sep;
    if (two) {
        Xf nA, nB;
        xfs(p, a, nA, nB);
        const PSM next = get_psm(m, gt ? 0u : 1u, nA, nB);
        const Sol s1 = solve(eq ? psm1 : next, eq ? op : p);
        p = p + s1.dl;
        a = a + s1.da;
    }
    pos.x = p.x;
    pos.y = p.y;
    pos.z = a;
    return minSep;
}

// ---------------------------------------------------------------------------
// Level form of the wave-wide sweeps, for islands without joints in worlds of
// at most kLevelMaxBodies bodies.  The Gauss-Seidel order is kept exactly: an
// update's level is one more than the last level that touched either of its
// bodies (bodies with zero inverse mass and inertia are never changed by an
// update, so they order nothing), and the updates of one level touch
// distinct bodies.  The wave runs each level's updates in parallel, level
// after level, two lanes to an update (each lane computes its own body's
// terms, the pair exchanging partial results), with the inputs the serial
// sweep gives them; warm-start
// deltas reach each body in the same order, and the early exits reduce the
// same maxima and minima.  The plan (levels, level order, rounds of at most
// one wave) is built once per step and serves the warm start, the velocity
// iterations and the position iterations; the records are copied into level
// order so that a round reads consecutive records.
constexpr int kLevelMaxBodies = prx::kStageFloat4 * 4;
constexpr int kLevelMinConstraints = 8;
// Updates per round: two lanes serve one update (one lane per body).
constexpr int kLevelRound = kWave / 2;

struct LevelPlan {
    int n, rounds;
    const int2* rb;    // per round: first and end position in level order
    const int* order;  // level position -> island position
};

__device__ __forceinline__ int* level_ints(const DeviceState& d, const WorldCtx& x, int c0) {
    return d.s_lvl + (std::size_t(x.co) + std::size_t(c0)) * prx::kLevelInts;
}
__device__ __forceinline__ int2* level_rounds(int* L, int n) { return reinterpret_cast<int2*>(L + ((3 * n + 3) & ~1)); }

// Builds the plan of island [c0, c1) (n >= kLevelMinConstraints); the
// record stage serves as the per-body last-level table.
__device__ __forceinline__ LevelPlan level_plan(const DeviceState& d, const WorldCtx& x, float4* stage, int c0,
                                                int c1) {
    const int n = c1 - c0;
    const int lane = x.lane;
    const float4* vbase = d.s_vc + (std::size_t(x.co) + std::size_t(c0)) * prx::kVcStride;
    const float4* pbase = d.s_pc + (std::size_t(x.co) + std::size_t(c0)) * prx::kPcStride;
    int* L = level_ints(d, x, c0);
    int* lvl = L;
    int* order = L + n;
    int* start = L + 2 * n;  // levels + 1 entries
    int2* rb = level_rounds(L, n);
    int* last = reinterpret_cast<int*>(stage);
    for (int b = lane; b < x.nb; b += kWave) last[b] = 0;
    __syncthreads();
    int levels = 0;
    for (int base = 0; base < n; base += kWave) {
        const int r = base + lane;
        int a = -1, bb = -1;
        if (r < n) {
            const float4* v = vbase + std::size_t(r) * prx::kVcStride;
            const float4 m = v[8];
            if (m.x != 0.0f || m.y != 0.0f) a = __float_as_int(v[2].w);
            if (m.z != 0.0f || m.w != 0.0f) bb = __float_as_int(v[5].z);
        }
        const int cnt = (n - base < kWave) ? n - base : kWave;
        for (int j = 0; j < cnt; ++j) {
            const int aj = __shfl(a, j), bj = __shfl(bb, j);
            if (lane == 0) {
                const int la = (aj >= 0) ? last[aj] : 0;
                const int lb = (bj >= 0) ? last[bj] : 0;
                const int l = ((la > lb) ? la : lb) + 1;
                if (aj >= 0) last[aj] = l;
                if (bj >= 0) last[bj] = l;
                lvl[base + j] = l - 1;
                levels = (l > levels) ? l : levels;
            }
        }
    }
    levels = __shfl(levels, 0);
    __syncthreads();
    for (int k = lane; k <= levels; k += kWave) start[k] = 0;
    __syncthreads();
    for (int r = lane; r < n; r += kWave) atomicAdd(&start[lvl[r]], 1);
    __syncthreads();
    {
        int carry = 0;
        for (int base = 0; base <= levels; base += kWave) {
            const int k = base + lane;
            const int v = (k <= levels) ? start[k] : 0;
            const int incl = wave_incl_scan(v, lane);
            if (k <= levels) start[k] = carry + incl - v;
            carry += __shfl(incl, kWave - 1);
        }
    }
    __syncthreads();
    int* cur = reinterpret_cast<int*>(rb);
    for (int k = lane; k < levels; k += kWave) cur[k] = start[k];
    __syncthreads();
    for (int r = lane; r < n; r += kWave) order[atomicAdd(&cur[lvl[r]], 1)] = r;
    __syncthreads();
    int rounds = 0;
    for (int base = 0; base < levels; base += kWave) {
        const int k = base + lane;
        const int cnt = (k < levels) ? start[k + 1] - start[k] : 0;
        const int nr = (cnt + kLevelRound - 1) / kLevelRound;
        const int incl = wave_incl_scan(nr, lane);
        for (int i = 0; i < nr; ++i) {
            const int b = start[k] + i * kLevelRound;
            const int e = (b + kLevelRound < start[k + 1]) ? b + kLevelRound : start[k + 1];
            rb[rounds + incl - nr + i] = int2{b, e};
        }
        rounds += __shfl(incl, kWave - 1);
    }
    float4* v2 = d.s_vc2 + (std::size_t(x.co) + std::size_t(c0)) * prx::kVcStride;
    float4* p2 = d.s_pc2 + (std::size_t(x.co) + std::size_t(c0)) * prx::kPcStride;
    for (int k = lane; k < n * prx::kVcStride; k += kWave) {
        const int pos = k / prx::kVcStride, f = k - pos * prx::kVcStride;
        v2[k] = vbase[std::size_t(order[pos]) * prx::kVcStride + f];
    }
    for (int k = lane; k < n * prx::kPcStride; k += kWave) {
        const int pos = k / prx::kPcStride, f = k - pos * prx::kPcStride;
        p2[k] = pbase[std::size_t(order[pos]) * prx::kPcStride + f];
    }
    __syncthreads();
    if (lane == 0) lvl[0] = rounds;  // levels are no longer needed
    __syncthreads();
    return LevelPlan{n, rounds, rb, order};
}

__device__ __forceinline__ LevelPlan level_plan_of(const DeviceState& d, const WorldCtx& x, int c0, int c1) {
    const int n = c1 - c0;
    int* L = level_ints(d, x, c0);
    return LevelPlan{n, L[0], level_rounds(L, n), L + n};
}

// Round bounds of a plan during its passes.  Round k covers level positions
// [end(k - 1), end(k)) (rounds are contiguous).  The ends are staged in LDS
// (the record stage is free while the passes run) when they fit in 16 bits
// and in the stage, so that no round waits on a global load for its bounds;
// otherwise they are read from the plan in global scratch.
struct RoundEnds {
    const unsigned short* lds;
    const int2* rb;
    int rounds;
    __device__ __forceinline__ int end(int k) const { return lds ? int(lds[k]) : rb[k].y; }
    __device__ __forceinline__ int2 round(int k) const {
        return int2{k > 0 ? end(k - 1) : 0, end(k)};
    }
};

__device__ __forceinline__ RoundEnds stage_round_ends(const LevelPlan& P, float4* stage, int lane) {
    constexpr int kCap = prx::kStageFloat4 * 8;  // 16-bit ends in the stage
    if (P.rounds > kCap || P.n > 0xffff) return RoundEnds{nullptr, P.rb, P.rounds};
    unsigned short* e = reinterpret_cast<unsigned short*>(stage);
    for (int k = lane; k < P.rounds; k += kWave) e[k] = static_cast<unsigned short>(P.rb[k].y);
    __syncthreads();
    return RoundEnds{e, P.rb, P.rounds};
}

// One pass of an island's level form.  Round k covers level positions
// [end(k - 1), end(k)); its records are loaded while round k - 1 is solved,
// and the end of round k + 1 is read while round k - 1 is solved, so that no
// round waits on memory it did not ask for a round earlier.  The records are
// read and the impulses written by buffer instructions that every lane
// issues in every round, lanes without an update at an offset past the
// buffer's end (no memory is touched for them): with no branch around them,
// the waits the compiler places count exactly the operations in flight, so
// a round's impulse stores retire while the next round runs.
constexpr int kOutOfRange = 0x40000000;  // a buffer offset past every island's records

template <int kStride, bool kLdsEnds>
struct LevelRounds {
    __amdgpu_buffer_rsrc_t rec;
    const unsigned short* e16;
    const int2* rb;
    int n, rounds, lane;
    __device__ __forceinline__ int end(int k) const {
        if (k >= rounds) return n;
        if constexpr (kLdsEnds) return int(e16[k]);
        else return rb[k].y;
    }
    __device__ __forceinline__ void fetch(int start, int stop, float4 (&buf)[kStride]) const {
        const int pos = start + (lane >> 1);
        const int off = (pos < stop) ? pos * (kStride * 16) : kOutOfRange;
#pragma unroll
        for (int f = 0; f < kStride; ++f)
            buf[f] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rec, off + 16 * f, 0, 0));
    }
    // Runs body(start, end, records) for every round, two rounds per trip
    // (the two record buffers alternate without copies).
    template <class F>
    __device__ __forceinline__ void run(F&& body) const {
        float4 a[kStride], b[kStride];
        int s = 0;
        int e = end(0);
        fetch(0, e, a);
        for (int k = 0; k < rounds; k += 2) {
            const int e1 = end(k + 1);
            fetch(e, e1, b);
            body(s, e, a);
            if (k + 1 >= rounds) break;
            const int e2 = end(k + 2);
            fetch(e1, e2, a);
            body(e, e1, b);
            s = e1;
            e = e2;
        }
    }
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t record_rsrc(const float4* base, int n, int stride) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(base), 0, n * stride * 16, 0x00020000);
}

using U4 = __attribute__((ext_vector_type(4))) unsigned;

template <bool kWarm, bool kLdsEnds, bool kLds>
__device__ __forceinline__ float velocity_level_pass(const LevelRounds<prx::kVcStride, kLdsEnds>& R,
                                                     const Bodies<kLds>& bs) {
    float newInc = 0.0f;
    const int side = R.lane & 1;
    R.run([&](int s, int e, float4 (&cur)[prx::kVcStride]) {
        const int pos = s + (R.lane >> 1);
        const bool act = pos < e;
        float4 o4 = cur[4], o7 = cur[7];
        if (act) {
            BodyC A, B;
            Vc vc = load_vc(cur, A, B);
            if constexpr (kWarm) {
                warm_start_pair(vc, A, B, bs, side);
            } else {
                const float mi = solve_velocity_pair(vc, A, B, bs, side);
                o4 = float4{vc.p[0].ni, vc.p[0].ti, vc.p[0].nm, vc.p[0].tm};
                o7 = float4{vc.p[1].ni, vc.p[1].ti, vc.p[1].nm, vc.p[1].tm};
                newInc = fmax_std(newInc, mi);
            }
        }
        if constexpr (!kWarm) {
            // The pair's lanes hold the same impulses: the body-A lane stores.
            const int off = (act && side == 0) ? pos * (prx::kVcStride * 16) : kOutOfRange;
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, o4), R.rec, off + 64, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, o7), R.rec, off + 112, 0, 0);
        }
        __syncthreads();
    });
    return newInc;
}

// Warm start and velocity sweeps of one island in level form.
template <bool kLdsEnds, bool kLds>
__device__ __forceinline__ VelResult island_velocity_level_run(const DeviceState& d, const WorldCtx& x,
                                                               const Bodies<kLds>& bs, const LevelPlan& P,
                                                               int rounds, const unsigned short* e16, int c0,
                                                               const StepParams& c) {
    VelResult res{0.0f, int(c.velIters)};
    float4* vbase = d.s_vc + (std::size_t(x.co) + std::size_t(c0)) * prx::kVcStride;
    float4* v2 = d.s_vc2 + (std::size_t(x.co) + std::size_t(c0)) * prx::kVcStride;
    const LevelRounds<prx::kVcStride, kLdsEnds> R{record_rsrc(v2, P.n, prx::kVcStride), e16, P.rb, P.n, rounds,
                                                  x.lane};
    if (c.warmStart) velocity_level_pass<true, kLdsEnds, kLds>(R, bs);
    for (int it = 0; it < int(c.velIters); ++it) {
        float newInc = velocity_level_pass<false, kLdsEnds, kLds>(R, bs);
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) newInc = fmax_std(newInc, __shfl_xor(newInc, off));
        res.maxInc = fmax_std(res.maxInc, newInc);
        if (newInc <= c.regMinMomentum) {
            res.iters = it + 1;
            break;
        }
    }
    // The passes' stores complete before other lanes read the records back:
    // s_waitcnt vmcnt(0) (expcnt and lgkmcnt at their maxima).
    __builtin_amdgcn_s_waitcnt(0x0F70);
    for (int pos = x.lane; pos < P.n; pos += kWave) {
        float4* dst = vbase + std::size_t(P.order[pos]) * prx::kVcStride;
        const float4* src = v2 + std::size_t(pos) * prx::kVcStride;
        dst[4] = src[4];
        dst[7] = src[7];
    }
    __syncthreads();
    return res;
}

template <bool kLds>
__device__ __forceinline__ VelResult island_velocity_level(const DeviceState& d, const WorldCtx& x,
                                                           const Bodies<kLds>& bs, const LevelPlan& P, int c0,
                                                           const StepParams& c, float4* stage) {
    const RoundEnds R = stage_round_ends(P, stage, x.lane);
    const int rounds = __builtin_amdgcn_readfirstlane(P.rounds);
    if (R.lds) return island_velocity_level_run<true, kLds>(d, x, bs, P, rounds, R.lds, c0, c);
    return island_velocity_level_run<false, kLds>(d, x, bs, P, rounds, nullptr, c0, c);
}

// Position sweeps of one island in level form; returns whether solved.
template <bool kLdsEnds, bool kLds>
__device__ __forceinline__ PosResult island_position_level_run(const DeviceState& d, const WorldCtx& x,
                                                               const Bodies<kLds>& bs, const LevelPlan& P,
                                                               int rounds, const unsigned short* e16, int c0,
                                                               const StepParams& c) {
    const float4* p2 = d.s_pc2 + (std::size_t(x.co) + std::size_t(c0)) * prx::kPcStride;
    const LevelRounds<prx::kPcStride, kLdsEnds> R{record_rsrc(p2, P.n, prx::kPcStride), e16, P.rb, P.n, rounds,
                                                  x.lane};
    PosResult res{__builtin_inff(), 0};
    for (std::uint32_t it = 0; it < c.posIters; ++it) {
        float minSep = __builtin_inff();
        const int side = x.lane & 1;
        R.run([&](int s, int e, float4 (&cur)[prx::kPcStride]) {
            if (s + (x.lane >> 1) < e) {
                const Pc pc = load_pc(cur);
                const int mine = side ? pc.bB : pc.bA;
                float4 p = bs.pos(mine);
                const float sep = solve_pc_pair(pc.m, pc.totalRadius, pc.A, pc.B, p, side, c);
                bs.set_pos(mine, p);
                minSep = fmin_std(minSep, sep);
            }
            __syncthreads();
        });
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) minSep = fmin_std(minSep, __shfl_xor(minSep, off));
        res.minSep = fmin_std(res.minSep, minSep);
        if (minSep >= c.regMinSeparation) {
            res.solvedAt = int(it) + 1;
            return res;
        }
    }
    return res;
}

template <bool kLds>
__device__ __forceinline__ PosResult island_position_level(const DeviceState& d, const WorldCtx& x, const Bodies<kLds>& bs,
                                                      const LevelPlan& P, int c0, const StepParams& c,
                                                      float4* stage) {
    const RoundEnds R = stage_round_ends(P, stage, x.lane);
    const int rounds = __builtin_amdgcn_readfirstlane(P.rounds);
    if (R.lds) return island_position_level_run<true, kLds>(d, x, bs, P, rounds, R.lds, c0, c);
    return island_position_level_run<false, kLds>(d, x, bs, P, rounds, nullptr, c0, c);
}

__device__ __forceinline__ bool level_form(const WorldCtx& x, int c0, int c1, int j0, int j1) {
    return j0 == j1 && x.nb <= kLevelMaxBodies && (c1 - c0) >= kLevelMinConstraints;
}

// What the island solves add to RegStepStats (Update(RegStepStats&,
// IslandStats), StepStats.cpp:29-38).
struct SolveSummary {
    std::uint32_t slept, solved, sumPos, sumVel;
    float minSep, maxInc;
};

template <bool kLds, bool kTape = false>
__device__ __forceinline__ SolveSummary solve_islands(const DeviceState& d, const WorldCtx& x, int nisl,
                                                      const IslandScratch& S, const StepParams& c, PhaseTimer& T) {
    // Per-lane sums and extrema of the islands each lane reports (lane-form
    // islands on their lane, wave-form islands on lane 0), reduced at the end.
    // The extrema do not depend on the order: incremental impulses are
    // magnitudes (never -0) and no separation is -0 (a - b of equal values
    // is +0 and the total radius is positive).
    float aMin = __builtin_inff(), aMax = 0.0f;
    std::uint32_t aVel = 0, aPos = 0, aSolved = 0;
    auto addVel = [&](const VelResult& r) {
        aMax = fmax_std(aMax, r.maxInc);
        aVel += std::uint32_t(r.iters);
    };
    auto addPos = [&](const PosResult& r) {
        aMin = fmin_std(aMin, r.minSep);
        aPos += r.solvedAt ? std::uint32_t(r.solvedAt) : c.posIters;
        aSolved += r.solvedAt ? 1u : 0u;
    };
    const int bo = x.bo;
    const Bodies<kLds> bs{d.s_bc_pos + bo, d.s_bc_vel + bo, x.nb};
    const int nbe = S.istart[nisl].x;
    const int nce = S.istart[nisl].y;
    // Body constraints (GetBodyConstraints, AabbTreeWorld.cpp:276-287) for
    // island bodies (which the island search also marks awake, :1410);
    // static bodies get their fixed pose and zero velocity.
    for (int b = x.lane; b < x.nb; b += kWave) {
        const int gb = bo + b;
        if (S.bisland[b] >= 0) {
            const float4 p1 = d.pos1[gb];
            d.pos0[gb] = p1;
            d.bflags[gb] |= prx::kAwake;
            bs.set_pos(b, p1);
            bs.set_vel(b, integrated_velocity(d, gb, c));
        } else if (!(d.bflags[gb] & prx::kSpeedable)) {
            bs.set_pos(b, d.pos1[gb]);
            bs.set_vel(b, float4{0.0f, 0.0f, 0.0f, 0.0f});
        }
    }
    for (int i = x.lane; i < nisl; i += kWave) {
        S.isl_word[i] = 0u;
        S.isl_under[i] = 0x7f800000u;
    }
    __syncthreads();
    // Velocity and position constraints of all island contacts, in island
    // order (independent per contact).
    for (int e = x.lane; e < nce; e += kWave) {
        const int gi = x.co + S.icont[e];
        const int2 bd = d.c_body[gi];
        const int2 ss = d.c_shape[gi];
        const float rA = d.shape_f[ss.x].x, rB = d.shape_f[ss.y].x;
        const BodyC A = load_bodyc(d, bo + bd.x), B = load_bodyc(d, bo + bd.y);
        const float4 pA = bs.pos(bd.x), pB = bs.pos(bd.y);
        const Xf xfA = get_xf(pA.x, pA.y, pA.z, A.lc);
        const Xf xfB = get_xf(pB.x, pB.y, pB.z, B.lc);
        const Manifold m = load_manifold(d, gi);
        const WM wm = world_manifold(m, xfA, rA, xfB, rB);
        const float4 mt = d.c_mat[gi];
        const Vc vc = make_vc(wm, mt.x, mt.y, mt.z, bd.x, bd.y, A, pA, bs.vel(bd.x), B, pB, bs.vel(bd.y), c);
        const std::size_t slot = std::size_t(x.co) + std::size_t(e);
        store_vc(d.s_vc + slot * prx::kVcStride, vc, A, B);
        store_pc(d.s_pc + slot * prx::kPcStride, m, bd.x, bd.y, rA + rB, A, B);
        if constexpr (kTape) {
            float4* t = d.tape.vc + slot * prx::kTapeVc;
            store_vc(t, vc, A, B);
            const float4 vA = bs.vel(bd.x), vB = bs.vel(bd.y);
            const float vn0 =
                dot(rel_vel(V2{vA.x, vA.y}, vA.z, vc.p[0].relA, V2{vB.x, vB.y}, vB.z, vc.p[0].relB), vc.normal);
            // Manifold points in world coordinates (for circles: the two
            // circle centres; for faces: body A's position and the clip point).
            const V2 wA = (m.type == kCircles) ? transform(m.lpt, xfA) : V2{pA.x, pA.y};
            const V2 wB = transform(m.p[0].lp, xfB);
            t[9] = float4{wA.x, wA.y, wB.x, wB.y};
            t[10] = float4{vn0, __uint_as_float(d.c_id[gi]), __uint_as_float(m.type | (m.n << 8)), rA};
            t[11] = float4{rB, __int_as_float(ss.x), __int_as_float(ss.y), 0.0f};
            t[12] = vA;
            t[13] = vB;
            t[15] = pA;
            t[16] = pB;
            store_pc(t + 17, m, bd.x, bd.y, rA + rB, A, B);
            t[22] = float4{m.p[0].ni, m.p[0].ti, m.p[1].ni, m.p[1].ti};
            t[23] = float4{__uint_as_float(m.p[0].cf), __uint_as_float(m.p[1].cf), 0.0f, 0.0f};
        }
    }
    if constexpr (kTape) {
        if (x.lane == 0) d.tape.nisl[x.w] = nisl;
        for (int is = x.lane; is < nisl; is += kWave) {
            d.tape.isl[x.bo + x.w + is] = int4{S.istart[is].y, S.istart[is + 1].y, 0, 0};
            d.tape.islj[x.bo + x.w + is] = int2{S.jstart[is], S.jstart[is + 1]};
        }
        for (int b = x.lane; b < x.nb; b += kWave) d.tape.inisl[x.bo + b] = S.bisland[b] >= 0 ? 1 : 0;
    }
    // Joint set-up (the non-warm-start half of InitVelocity), island order.
    const int nje = S.jstart[nisl];
    for (int e = x.lane; e < nje; e += kWave) {
        const int jg = x.jo + S.ijoint[e];
        const int2 bd = d.j_body[jg];
        const BodyC A = load_bodyc(d, bo + bd.x), B = load_bodyc(d, bo + bd.y);
        if constexpr (kTape) {
            float4* t = d.tape.jc + (std::size_t(x.jo) + std::size_t(e)) * prx::kTapeJc;
            t[prx::kJcStride] = float4{bs.pos(bd.x).z, bs.pos(bd.y).z, d.j_cfg2[jg].x, __uint_as_float(d.j_limit[jg])};
            t[prx::kJcStride + 1] = d.j_imp[jg];
        }
        const std::uint32_t jflags = __float_as_uint(d.j_cfg2[jg].y);
        if ((jflags >> 8) == prx::kJointGear) {
            const float4 g0 = d.j_gear[2 * std::size_t(jg)];
            const int bC = __float_as_int(g0.z), bD = __float_as_int(g0.w);
            const BodyC C = load_bodyc(d, bo + bC), D = load_bodyc(d, bo + bD);
            build_jc_gear(d, jg, d.s_jc + (std::size_t(x.jo) + std::size_t(e)) * prx::kJcStride, A, bs.pos(bd.x), B,
                          bs.pos(bd.y), C, bs.pos(bC), D, bs.pos(bD), bd.x, bd.y, bC, bD, jflags, c);
        } else {
            build_jc(d, x, jg, d.s_jc + (std::size_t(x.jo) + std::size_t(e)) * prx::kJcStride, A, bs.pos(bd.x), B,
                     bs.pos(bd.y), bd.x, bd.y, c);
        }
        if constexpr (kTape) {
            const float4* u = d.s_jc + (std::size_t(x.jo) + std::size_t(e)) * prx::kJcStride;
            float4* t = d.tape.jc + (std::size_t(x.jo) + std::size_t(e)) * prx::kTapeJc;
            for (int k = 0; k < prx::kJcStride; ++k) t[k] = u[k];
            t[7].w = __int_as_float(jg);  // the record leaves .w unused
        }
    }
    __syncthreads();
    T.mark(prx::kPhConstraints);
    // Warm start and velocity sweeps: islands below coopMin contacts one lane
    // each, larger ones by the whole wave, one after another.
    float4* stage = g_smem + (int(c.ldsBytes) / 16 - prx::kStageFloat4);
    const int coopMin = (c.coopMin == 0u) ? 1 : (c.coopMin > 0x7fffffffu ? 0x7fffffff : int(c.coopMin));
    auto constraints = [&](int is) {
        return (S.istart[is + 1].y - S.istart[is].y) + (S.jstart[is + 1] - S.jstart[is]);
    };
    for (int is = x.lane; is < nisl; is += kWave) {
        if (constraints(is) < coopMin)
            addVel(island_velocity<kLds, kTape>(d, x, bs, S.istart[is].y, S.istart[is + 1].y, S.jstart[is],
                                                S.jstart[is + 1], c, x.bo + x.w + is));
    }
    __syncthreads();
    for (int base = 0; base < nisl; base += kWave) {
        const int is = base + x.lane;
        const bool big = is < nisl && constraints(is) >= coopMin;
        for (unsigned long long m = __ballot(big); m; m &= m - 1) {
            const int isl = base + __ffsll(m) - 1;
            const int c0 = S.istart[isl].y, c1 = S.istart[isl + 1].y;
            VelResult r;
            if (level_form(x, c0, c1, S.jstart[isl], S.jstart[isl + 1])) {
                const LevelPlan P = level_plan(d, x, stage, c0, c1);
                r = island_velocity_level<kLds>(d, x, bs, P, c0, c, stage);
            } else {
                r = island_velocity_wave<kLds>(d, x, bs, stage, c0, c1, S.jstart[isl], S.jstart[isl + 1], c);
            }
            if (x.lane == 0) addVel(r);
        }
    }
    T.mark(prx::kPhVelocity);
    // IntegratePositions (AabbTreeWorld.cpp:176-188).
    for (int e = x.lane; e < nbe; e += kWave) {
        const int b = S.ibody[e];
        float4 p = bs.pos(b);
        const float4 v = bs.vel(b);
        const float tx = c.dt * v.x, ty = c.dt * v.y, r = c.dt * v.z;
        p.x = p.x + tx;
        p.y = p.y + ty;
        p.z = p.z + r;
        bs.set_pos(b, p);
    }
    __syncthreads();
    // Position sweeps, split like the velocity sweeps.
    for (int is = x.lane; is < nisl; is += kWave) {
        if (constraints(is) < coopMin) {
            const PosResult r = island_position<kLds, kTape>(d, x, bs, S.istart[is].y, S.istart[is + 1].y,
                                                             S.jstart[is], S.jstart[is + 1], c, x.bo + x.w + is);
            S.isl_word[is] = std::uint32_t(r.solvedAt);
            addPos(r);
        }
    }
    __syncthreads();
    for (int base = 0; base < nisl; base += kWave) {
        const int is = base + x.lane;
        const bool big = is < nisl && constraints(is) >= coopMin;
        for (unsigned long long m = __ballot(big); m; m &= m - 1) {
            const int isl = base + __ffsll(m) - 1;
            const int c0 = S.istart[isl].y, c1 = S.istart[isl + 1].y;
            const PosResult r = level_form(x, c0, c1, S.jstart[isl], S.jstart[isl + 1])
                                    ? island_position_level<kLds>(d, x, bs, level_plan_of(d, x, c0, c1), c0, c,
                                                                  stage)
                                    : island_position_wave<kLds>(d, x, bs, stage, c0, c1, S.jstart[isl],
                                                                 S.jstart[isl + 1], c);
            if (x.lane == 0) {
                S.isl_word[isl] = std::uint32_t(r.solvedAt);
                addPos(r);
            }
        }
    }
    __syncthreads();
    T.mark(prx::kPhPosition);
    // AssignImpulses (AabbTreeWorld.cpp:208-227).
    for (int e = x.lane; e < nce; e += kWave) {
        const int gi = x.co + S.icont[e];
        const float4* rec = d.s_vc + (std::size_t(x.co) + std::size_t(e)) * prx::kVcStride;
        if constexpr (kTape)
            d.tape.vc[(std::size_t(x.co) + std::size_t(e)) * prx::kTapeVc + 14] = float4{rec[4].x, rec[4].y, rec[7].x, rec[7].y};
        const std::uint32_t n = __float_as_uint(rec[1].w);
        const float4 q0 = rec[4];
        float4 m0 = d.m_p0[gi];
        m0.z = q0.x;
        m0.w = q0.y;
        d.m_p0[gi] = m0;
        if (n > 1) {
            const float4 q1 = rec[7];
            float4 m1 = d.m_p1[gi];
            m1.z = q1.x;
            m1.w = q1.y;
            d.m_p1[gi] = m1;
        }
    }
    // Joint state carried to the next step: impulses and limit state.
    for (int e = x.lane; e < nje; e += kWave) {
        const int jg = x.jo + S.ijoint[e];
        const float4* u = d.s_jc + (std::size_t(x.jo) + std::size_t(e)) * prx::kJcStride;
        d.j_imp[jg] = u[3];
        d.j_limit[jg] = (__float_as_uint(u[2].w) >> 8) & 0xffu;
    }
    // Body writeback, under-active times (AabbTreeWorld.cpp:1652-1664, 403-417).
    for (int e = x.lane; e < nbe; e += kWave) {
        const int b = S.ibody[e];
        const int gb = bo + b;
        const float4 bv = bs.vel(b);
        const float4 bp = bs.pos(b);
        const float4 p1 = d.pos1[gb];
        std::uint32_t f = d.bflags[gb];
        if (!(p1.x == bp.x && p1.y == bp.y) || p1.z != bp.z) {
            d.pos1[gb] = float4{bp.x, bp.y, bp.z, 0.0f};
            const float4 l = d.lcinv[gb];
            const Xf xf = get_xf(bp.x, bp.y, bp.z, V2{l.x, l.y});
            d.xf[gb] = float4{xf.p.x, xf.p.y, xf.q.x, xf.q.y};
            f |= 0x100u;  // position changed this step
            d.bflags[gb] = f;
        }
        float4 v = d.vel[gb];
        v.x = bv.x;
        v.y = bv.y;
        v.z = bv.z;
        const float linSq = magsq(V2{v.x, v.y});
        const float angSq = v.z * v.z;
        const bool under = (angSq <= c.angularSleepTolerance * c.angularSleepTolerance) &&
                           (linSq <= c.linearSleepTolerance * c.linearSleepTolerance);
        const float t = ((f & prx::kAutoSleep) && under) ? v.w + c.dt : 0.0f;
        if (t == 0.0f || (f & prx::kAccelerable)) v.w = t;
        d.vel[gb] = v;
        // Under-active times are non-negative: their bit patterns order as floats.
        atomicMin(&S.isl_under[S.bisland[b]], __float_as_uint(t));
    }
    __syncthreads();
    // Sleepem (AabbTreeWorld.cpp:419-429) for islands still long enough and solved.
    std::uint32_t slept = 0;
    for (int e = x.lane; e < nbe; e += kWave) {
        const int b = S.ibody[e];
        const int is = S.bisland[b];
        if (__uint_as_float(S.isl_under[is]) >= c.minStillTimeToSleep && S.isl_word[is]) {
            const int gb = bo + b;
            const std::uint32_t f = d.bflags[gb];
            if ((f & prx::kAwake) && (f & prx::kAutoSleep)) {
                d.bflags[gb] = f & ~prx::kAwake;
                d.vel[gb] = float4{0.0f, 0.0f, 0.0f, 0.0f};
                ++slept;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        slept += __shfl_xor(slept, o);
        aSolved += __shfl_xor(aSolved, o);
        aPos += __shfl_xor(aPos, o);
        aVel += __shfl_xor(aVel, o);
        aMin = fmin_std(aMin, __shfl_xor(aMin, o));
        aMax = fmax_std(aMax, __shfl_xor(aMax, o));
    }
    __syncthreads();
    T.mark(prx::kPhWriteback);
    return SolveSummary{slept, aSolved, aPos, aVel, aMin, aMax};
}

// Synchronize (AabbTreeWorld.cpp:2564-2585) for one body.
__device__ __forceinline__ std::uint32_t synchronize_body(const DeviceState& d, const WorldCtx& x, int b, const StepParams& c) {
    const int gb = x.bo + b;
    const Xf xf0 = body_xf0(d, gb);
    const Xf xf1 = body_xf1(d, gb);
    const int2 bp = d.bproxy[gb];
    std::uint32_t moved = 0;
    for (int k = 0; k < bp.y; ++k) {
        const int p = d.bproxy_list[x.po + bp.x + k];
        const int gp = x.po + p;
        const Proxy px = shape_proxy(d, d.pshape[gp]);
        const float4 tight = proxy_aabb(px, xf0, xf1, c, false);
        const float4 fat = d.paabb[gp];
        const bool contains = fat.x <= tight.x && fat.z >= tight.z && fat.y <= tight.y && fat.w >= tight.w;
        if (!contains) {
            d.paabb[gp] = proxy_aabb(px, xf0, xf1, c, true);
            d.pflags[gp] |= prx::kProxyMoved;
            ++moved;
        }
    }
    return moved;
}

}  // namespace prxd

#include "toi_device.hpp"

namespace prxd {

// Post-solve listener calls (AabbTreeWorld.cpp:1671-1674), island after
// island in island contact order: the velocity constraint's points and
// impulses (a constraint can keep one point of a two-point manifold) and the
// iteration that solved the island's positions (255, the 8-bit
// InvalidIteration, if none did).
__device__ __forceinline__ void post_solve_calls(const DeviceState& d, const WorldCtx& x, Counters& k,
                                                 const IslandScratch& S, int nisl) {
    for (int is = 0; is < nisl; ++is) {
        const int c0 = S.istart[is].y, c1 = S.istart[is + 1].y;
        const std::uint32_t sw = S.isl_word[is];
        const std::uint32_t iter = sw ? ((sw - 1u) & 0xffu) : 255u;
        for (int base = c0; base < c1; base += kWave) {
            const int e = base + x.lane;
            const bool v = e < c1;
            const int at = ev_claim(x, k, v ? prx::kEvPostWords : 0, v ? 1 : 0);
            if (at < 0) continue;
            const int gi = x.co + S.icont[e];
            const float4* rec = d.s_vc + (std::size_t(x.co) + std::size_t(e)) * prx::kVcStride;
            const std::uint32_t n = __float_as_uint(rec[1].w);
            const float4 q0 = rec[4], q1 = rec[7];
            std::uint32_t* o = d.ev + at;
            o[0] = prx::kEvPostSolve;
            o[1] = d.c_id[gi];
            o[2] = n;
            o[3] = iter;
            o[4] = n > 0 ? __float_as_uint(q0.x) : 0u;
            o[5] = n > 1 ? __float_as_uint(q1.x) : 0u;
            o[6] = n > 0 ? __float_as_uint(q0.y) : 0u;
            o[7] = n > 1 ? __float_as_uint(q1.y) : 0u;
        }
    }
    __syncthreads();
}

// The TOI phase of one world after its regular phase (a kernel of its own:
// the serial event solves would otherwise set the register budget of the
// regular step).  Counters and listener calls continue from the header.
// Instruction-issue priority of a world's wavefront on its SIMD: the step
// time of a batch is the time of its longest worlds, so larger worlds issue
// ahead of the smaller ones that share their SIMD (these have slack).
__device__ __forceinline__ void world_priority(int nbodies) {
    if (nbodies >= 384) __builtin_amdgcn_s_setprio(3);
    else if (nbodies >= 256) __builtin_amdgcn_s_setprio(2);
    else if (nbodies >= 128) __builtin_amdgcn_s_setprio(1);
}

__device__ __forceinline__ void toi_world(const DeviceState& d, int w, const StepParams& c) {
    WorldHeader& H = d.hdr[w];
    WorldCtx x;
    x.w = w;
    x.bo = H.body_off;
    x.nb = H.nbodies;
    x.po = H.proxy_off;
    x.np = H.nproxies;
    x.co = H.contact_off;
    x.cap = H.contact_cap;
    x.jo = H.joint_off;
    x.nj = H.njoints;
    x.eo = H.ev_off;
    x.ecap = H.ev_cap;
    x.lane = int(threadIdx.x);
    Counters k;
    k.nc = H.ncontacts;
    k.nfree = H.nfree;
    k.pool = H.pool_size;
    k.status = std::uint32_t(H.status);
    k.evw = H.ev_words;
    k.evn = H.ev_count;
    world_priority(x.nb);
    IslandScratch S;
    S.off = d.s_csr_off + x.bo + w;
    S.csr = reinterpret_cast<int2*>(d.s_csr) + 2 * x.co;
    S.ibody = d.s_isl_body + x.bo + x.co;
    S.icont = d.s_isl_contact + x.co;
    S.istart = d.s_isl_start + x.bo + w;
    S.ijoint = d.s_isl_joint + x.jo;
    S.jstart = d.s_isl_jstart + x.bo + w;
    S.bisland = d.s_stack + 2 * x.bo + x.nb;
    S.isl_word = d.s_keys + x.co;
    S.isl_under = d.s_keys2 + x.co;
    __syncthreads();
    toi_phase(d, x, k, S, c, H);
    if (x.lane == 0) {
        H.ncontacts = k.nc;
        H.nfree = k.nfree;
        H.pool_size = k.pool;
        H.status = int(k.status);
        H.ev_words = k.evw;
        H.ev_count = k.evn;
    }
}

template <bool kTape = false>
__device__ __forceinline__ void step_world(const DeviceState& d, int w, const StepParams& c) {
    WorldHeader& H = d.hdr[w];
    WorldCtx x;
    x.w = w;
    x.bo = H.body_off;
    x.nb = H.nbodies;
    x.po = H.proxy_off;
    x.np = H.nproxies;
    x.co = H.contact_off;
    x.cap = H.contact_cap;
    x.jo = H.joint_off;
    x.nj = H.njoints;
    x.eo = H.ev_off;
    x.ecap = H.ev_cap;
    x.lane = int(threadIdx.x);
    Counters k;
    k.nc = H.ncontacts;
    k.nfree = H.nfree;
    k.pool = H.pool_size;
    k.status = std::uint32_t(H.status);
    k.evw = 0;
    k.evn = 0;
    // Step statistics go to the header as they are produced (lane 0).
    auto stat = [&](int i, std::uint32_t v) {
        if (x.lane == 0) H.stats[i] = v;
    };
    if (x.lane == 0) {
        for (int i = 0; i < prx::kStats; ++i) H.stats[i] = 0u;
        H.stats[prx::kStRegMinSeparation] = 0x7f800000u;  // +infinity: no position iteration ran
        H.stats[prx::kStToiMinSeparation] = 0x7f800000u;
    }
    world_priority(x.nb);
    PhaseTimer T;
    T.begin(d.phase_cycles ? d.phase_cycles + std::size_t(w) * prx::kPhases : nullptr);
    // Carry map of the recorded step: contacts existing before the step owe
    // their impulses to the step before (kCarryFresh clear).
    std::uint32_t* const cmap = kTape ? d.tape.cmap + x.co : nullptr;
    if constexpr (kTape) {
        for (int i = x.lane; i < k.nc; i += kWave) cmap[d.c_id[x.co + i]] = 0u;
    }

    // A: proxies of newly attached shapes (CreateProxies, AabbTreeWorld.cpp:557-579).
    std::uint32_t created = 0;
    for (int p = x.lane; p < x.np; p += kWave) {
        const int gp = x.po + p;
        if (d.pflags[gp] & prx::kProxyPending) {
            const int gb = x.bo + d.pbody[gp];
            d.paabb[gp] = proxy_aabb(shape_proxy(d, d.pshape[gp]), body_xf0(d, gb), body_xf1(d, gb), c, true);
            d.pflags[gp] = prx::kProxyMoved;
            ++created;
        }
    }
    for (int o = 32; o > 0; o >>= 1) created += __shfl_xor(created, o);
    stat(prx::kStPreProxiesCreated, created);
    __syncthreads();
    // Bodies edited since the last step (SetTransformation and its kin):
    // Synchronize (AabbTreeWorld.cpp:2162-2173).
    std::uint32_t premoved = 0;
    for (int b = x.lane; b < x.nb; b += kWave) {
        const std::uint32_t f = d.bflags[x.bo + b];
        if (f & prx::kNeedsSync) {
            premoved += synchronize_body(d, x, b, c);
            d.bflags[x.bo + b] = f & ~prx::kNeedsSync;
        }
    }
    for (int o = 32; o > 0; o >>= 1) premoved += __shfl_xor(premoved, o);
    stat(prx::kStPreProxiesMoved, premoved);
    __syncthreads();
    T.mark(prx::kPhProxies);
    // B, C, D.
    stat(prx::kStPreDestroyed, phase_destroy(d, x, k));
    __syncthreads();
    T.mark(prx::kPhDestroy);
    {
        const std::uint32_t updated = phase_update(d, x, k, c, int(c.ldsBytes));
        stat(prx::kStPreUpdated, updated);
        stat(prx::kStPreSkipped, std::uint32_t(k.nc) - updated);
    }
    __syncthreads();
    T.mark(prx::kPhUpdate);
    const int ncBeforeAdd = k.nc;
    stat(prx::kStPreAdded, phase_find_add(d, x, k, c, int(c.ldsBytes)));
    __syncthreads();
    if constexpr (kTape) {
        // Contacts created before the solve start from impulses of their own.
        for (int i = ncBeforeAdd + x.lane; i < k.nc; i += kWave) cmap[d.c_id[x.co + i]] = prx::kCarryFresh;
        __syncthreads();
    }
    T.mark(prx::kPhFindAdd);

    if (c.dt != 0.0f) {
        IslandScratch S;
        S.off = d.s_csr_off + x.bo + w;
        S.csr = reinterpret_cast<int2*>(d.s_csr) + 2 * x.co;
        S.ibody = d.s_isl_body + x.bo + x.co;
        S.icont = d.s_isl_contact + x.co;
        S.istart = d.s_isl_start + x.bo + w;
        S.ijoint = d.s_isl_joint + x.jo;
        S.jstart = d.s_isl_jstart + x.bo + w;
        S.bisland = d.s_stack + 2 * x.bo + x.nb;
        S.isl_word = d.s_keys + x.co;
        S.isl_under = d.s_keys2 + x.co;
        int maxBodies = 0;
        int nisl = find_islands_lds(d, x, k.nc, S, maxBodies, int(c.ldsBytes));
        T.mark(prx::kPhAdjacency);
        if (nisl < 0) {
            build_adjacency(d, x, k.nc, S);
            const int nent = S.off[x.nb];
            const int ldsAdj = 4 * (((x.nb + 2) >> 1) * 2) + 8 * nent + x.nb + k.nc + x.nj;
            if (ldsAdj <= int(c.ldsBytes)) nisl = find_islands<2>(d, x, k.nc, S, maxBodies);
            else if (x.nb + k.nc + x.nj <= int(c.ldsBytes)) nisl = find_islands<1>(d, x, k.nc, S, maxBodies);
            else nisl = find_islands<0>(d, x, k.nc, S, maxBodies);
        }
        stat(prx::kStRegIslandsFound, std::uint32_t(nisl));
        stat(prx::kStRegMaxIslandBodies, std::uint32_t(maxBodies));
        T.mark(prx::kPhIslands);
        const bool ldsBodies = x.nb * 32 + prx::kStageFloat4 * 16 <= int(c.ldsBytes);
        const SolveSummary sv = ldsBodies ? solve_islands<true, kTape>(d, x, nisl, S, c, T)
                                          : solve_islands<false, kTape>(d, x, nisl, S, c, T);
        stat(prx::kStRegBodiesSlept, sv.slept);
        stat(prx::kStRegIslandsSolved, sv.solved);
        stat(prx::kStRegSumPosIters, sv.sumPos);
        stat(prx::kStRegSumVelIters, sv.sumVel);
        stat(prx::kStRegMinSeparation, __float_as_uint(sv.minSep));
        stat(prx::kStRegMaxIncImpulse, __float_as_uint(sv.maxInc));
        if (x.ecap > 0) post_solve_calls(d, x, k, S, nisl);
        // Synchronise proxies of islanded speedable bodies (AabbTreeWorld.cpp:1532-1544).
        std::uint32_t moved = 0;
        for (int b = x.lane; b < x.nb; b += kWave) {
            if (S.bisland[b] >= 0) moved += synchronize_body(d, x, b, c);
        }
        for (int o = 32; o > 0; o >>= 1) moved += __shfl_xor(moved, o);
        stat(prx::kStRegProxiesMoved, moved);
        __syncthreads();
        T.mark(prx::kPhSync);
        // Contacts of bodies whose position changed need updating.
        for (int i = x.lane; i < k.nc; i += kWave) {
            const int2 bd = d.c_body[x.co + i];
            if ((d.bflags[x.bo + bd.x] | d.bflags[x.bo + bd.y]) & 0x100u) d.c_flags[x.co + i] |= prx::kNeedsUpdating;
        }
        __syncthreads();
        for (int b = x.lane; b < x.nb; b += kWave) d.bflags[x.bo + b] &= ~0x100u;
        if constexpr (kTape) {
            // Contacts the update leaves alone keep their points and impulses.
            for (int i = x.lane; i < k.nc; i += kWave) {
                std::uint32_t& e = cmap[d.c_id[x.co + i]];
                e = (e & prx::kCarryFresh) | 0x10u | (((d.m_hdr[x.co + i].x >> 8) & 0xffu) << 8);
            }
        }
        __syncthreads();
        {
            const std::uint32_t updated = phase_update(d, x, k, c, int(c.ldsBytes), cmap);
            stat(prx::kStRegUpdated, updated);
            stat(prx::kStRegSkipped, std::uint32_t(k.nc) - updated);
        }
        __syncthreads();
        T.mark(prx::kPhUpdate2);
        stat(prx::kStRegAdded, phase_find_add(d, x, k, c, int(c.ldsBytes)));
        __syncthreads();
        T.mark(prx::kPhFindAdd2);
    }
    T.end();
    if (x.lane == 0) {
        H.ncontacts = k.nc;
        H.nfree = k.nfree;
        H.pool_size = k.pool;
        H.status = int(k.status);
        H.ev_words = k.evw;
        H.ev_count = k.evn;
    }
}

}  // namespace prxd
