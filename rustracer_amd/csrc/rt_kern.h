// Shared by the kernel translation units and the host code of librtamd:
// launch geometry, the per-lane LDS traversal stack and k-NN heap, and the
// device-resident layouts of the wavefront renderer (path records, queues,
// shadow requests, batch views, adjoint replay records).
#pragma once
#include <type_traits>
#include "dev_shade.h"
#include "dev_adjoint.h"
#include "dev_accum.h"
#include "../../include/rt_api.h"

using namespace rt;

#define BLOCK 128

struct LdsStack {
    __device__ static lds_i32* column() {
        __shared__ int32_t s[RT_LDS_STACK * BLOCK];
        return (lds_i32*)(s + threadIdx.x);
    }
};
// The k-NN photon gather's max-heap: RT_KNN_MAX floats per lane (32 KB per
// block), allocated only by the kernel variants that search k-NN radii.
struct LdsHeap {
    __device__ static lds_f32* column() {
        __shared__ float h[RT_KNN_MAX * BLOCK];
        return (lds_f32*)(h + threadIdx.x);
    }
};

// A traversal stack: the lane's LDS column plus its private overflow array,
// declared beside the stack so the Stack itself stays in registers.
#define RT_STACK(st) int32_t st##_priv[RT_PRIV_STACK]; Stack st; init_stack(st, st##_priv)
#define RT_STACK_BS(st, BS) int32_t st##_priv[RT_PRIV_STACK]; Stack st; init_stack_bs<BS>(st, st##_priv)

__device__ __forceinline__ void init_stack(Stack& st, int32_t* priv) {
    st.lds = LdsStack::column(); st.stride = BLOCK; st.priv = (priv_i32*)priv; st.sp = 0; st.heap = nullptr; st.has_heap = false;
    st.n_nodes = 0; st.n_prims = 0; st.n_rays = 0;
    st.s_nodes = 0; st.s_prims = 0; st.s_rays = 0;
    st.n_pbytes = 0;
}
// The traversal stages of the wavefront are compiled for blocks of 64, 128
// and 256 lanes (the autotuner picks one); their LDS stack is sized to it.
template <int BS>
struct LdsStackT {
    __device__ static lds_i32* column() {
        __shared__ int32_t s[RT_LDS_STACK * BS];
        return (lds_i32*)(s + threadIdx.x);
    }
};
template <int BS>
__device__ __forceinline__ void init_stack_bs(Stack& st, int32_t* priv) {
    st.lds = LdsStackT<BS>::column(); st.stride = BS; st.priv = (priv_i32*)priv; st.sp = 0; st.heap = nullptr; st.has_heap = false;
    st.n_nodes = 0; st.n_prims = 0; st.n_rays = 0;
    st.s_nodes = 0; st.s_prims = 0; st.s_rays = 0;
    st.n_pbytes = 0;
}

__device__ __forceinline__ Ray ray_from_c(const rt_ray& q) {
    return make_ray(mk(q.origin[0], q.origin[1], q.origin[2]), mk(q.dir[0], q.dir[1], q.dir[2]), q.time);
}

// Add a per-lane count into *count with one atomic per wave (every lane of
// the wave calls it).
__device__ __forceinline__ void wave_count(uint32_t local, uint32_t* count) {
    for (int off = 32; off > 0; off >>= 1) local += __shfl_xor(local, off);
    if (__lane_id() == 0 && local) atomicAdd(count, local);
}
#define RT_COUNT_GRID 1024u   // blocks of the grid-stride counting kernels: few atomics on one counter

#define RT_MAX_PARAMS 64

// Photon emission + tracing (PhotonMap::build_from_quad / build, src/photon.rs:41-135, 232-277).
struct DEmitter {
    uint32_t kind;
    float o[3], u[3], v[3], n[3], ta[3], ba[3];
    float power[3];
    float lum_clamp;
    float disk_c[3], disk_r, t[3], b[3], sun_down[3];
    uint32_t wide;             // photon wavelengths over [380, 780) (Ray::wl 1), else [380, 700)
};

// ============================================================================ wavefront
// Path records live in HBM (128 B each, one cache line); queues hold path ids.
// Per bounce: k_wf_extend traces every active path and files its id under the
// material key of the hit (the sort), k_wf_shade runs one bounce in key order
// and appends survivors and shadow requests, k_wf_shadow resolves the shadow
// requests.  All three are persistent grid-stride kernels that read their
// queue length from device memory, so the host never synchronises inside a
// frame.  Contributions go straight into the frame accumulator with float
// atomics.
#define WF_NB (M_KINDS + 2)   // sort keys (sort 1): 0 miss, 1..M_KINDS material kind, M_KINDS + 1 medium event;
                          // (sort 2): 0 miss / specular / medium, 1 every other material
#define WF_CNT_ACTIVE 0   // counters[0]: paths to extend this bounce
#define WF_CNT_DENSE 1    // counters[1]: k-NN queries holding at least k photons (second k-NN pass)
#define WF_CNT_NEXT 2     // counters[2]: survivors written by shade
#define WF_CNT_SHADOW 3   // counters[3]: shadow requests; with NEXT one 8-byte-aligned pair, so the
                          // shade stage reserves both with one 64-bit atomic per wave
static_assert(WF_CNT_NEXT % 2 == 0 && WF_CNT_SHADOW == WF_CNT_NEXT + 1, "NEXT / SHADOW pair");
// bucket sizes: bucket b's counter at WF_CNT_BKT(b), 256 B apart, so the
// atomics of one extend wave's buckets go to different L2 channels
#define WF_BKT_STRIDE 64
#define WF_CNT_BKT(b) (64 + (b) * WF_BKT_STRIDE)
#define WF_CNT_WORDS 1024 // the counter block: 4 KB
static_assert(WF_CNT_BKT(WF_NB) <= WF_CNT_WORDS, "bucket counters fit the block");
#define WF_CNT_FIRST 32   // counters[32..33]: 64-bit frame position of a graph-replayed chunk
#define WF_CNT_GEN 34     // counters[34]: paths generated when flagged pixels are skipped

struct __align__(16) WfPath {
    float4 o_time;       // ray origin, time
    float4 d_wl;         // ray direction, wavelength
    float4 thr_mis;      // throughput, prev_mis
    float sun_w; uint32_t flags; uint32_t pixel; uint32_t pad0;   // flags: 1 spectral, 2 prev_specular, bits 2-3 Ray::wl, 4-5 Ray::tint
    uint64_t rng[4];
    float hit_t; int32_t hit_slot; int32_t hit_prim; uint32_t key;
    uint32_t sid;        // sample index within the chunk (per-sample radiance of the adjoint's primal pass)
    uint32_t pad1[3];
};
static_assert(sizeof(WfPath) == 128, "path record is one 128-B line");

struct __align__(16) WfShadow { float4 o_tmax; float4 d_time; float4 c_pixel; uint4 meta; };   // meta.x = view

// Every field has a default, so a WfQueues is never handed to a kernel with
// an unset pointer (a forward stage once wrote through an uninitialised one).
struct WfQueues {
    WfPath* paths = nullptr;
    uint32_t cap = 0;
    uint32_t* active = nullptr;     // cap
    uint32_t* next = nullptr;       // cap
    uint32_t* buckets = nullptr;    // WF_NB * cap
    // 3 * cap: the forward shade stage writes the k-th shadow request of the
    // bounce's i-th path at k * cap + i (at most three: area light, sun,
    // environment) and lists the filled slots in shadow_idx, which k_wf_shadow
    // walks; the replay kernels (k_wfb_*) fill it densely; between extend and
    // shade the k-NN stage keeps its dense queries here
    WfShadow* shadow = nullptr;
    uint32_t* shadow_idx = nullptr; // 3 * cap
    uint32_t* counters = nullptr;   // WF_CNT_WORDS: 32 per-bounce counters, the chunk position, the bucket sizes
    float* fb = nullptr;
    float* lsamp = nullptr;         // non-null: contributions go to the path's sample (3 floats per sid), not to fb
    uint32_t sort = 1;              // 0: one bucket; 1: a bucket per material kind; 2: two shade classes
    uint32_t* dense = nullptr;      // cap: paths whose k-NN query needs the heap search
    const uint32_t* stop = nullptr; // non-null and nonzero at chunk start: the chunk renders no path (a stopped render loop)
    const uint8_t* conv = nullptr;  // non-null: flagged pixels get no path (the render loop's converged pixels)
    // dense k-NN queries at most this many: one query per wave, else one per
    // lane (on the 2M-photon gem at 1280 x 720 the crossover lies between
    // 64k and 256k queries; camera-ray bounces exceed it, later ones do not)
    uint32_t knn_wave_max = 65536;
    // four hero wavelengths: 2 per path, {q0, q1, q2, r0} {r1, r2, on, 0} (Spec4),
    // read and written by the shade stage's four-wavelength kernels (h4 != 0)
    float4* spec = nullptr;
    uint32_t h4 = 0;
};
static_assert(std::is_trivially_copyable<WfQueues>::value, "kernel argument");
// A default WfQueues has every pointer null and every count zero, checked at
// compile time in every unit that launches the stages.
constexpr WfQueues kWfQueuesDefault{};
static_assert(kWfQueuesDefault.paths == nullptr && kWfQueuesDefault.active == nullptr &&
              kWfQueuesDefault.next == nullptr && kWfQueuesDefault.buckets == nullptr &&
              kWfQueuesDefault.shadow == nullptr && kWfQueuesDefault.shadow_idx == nullptr &&
              kWfQueuesDefault.counters == nullptr &&
              kWfQueuesDefault.fb == nullptr && kWfQueuesDefault.lsamp == nullptr &&
              kWfQueuesDefault.dense == nullptr && kWfQueuesDefault.stop == nullptr &&
              kWfQueuesDefault.conv == nullptr && kWfQueuesDefault.spec == nullptr &&
              kWfQueuesDefault.cap == 0 && kWfQueuesDefault.h4 == 0,
              "every WfQueues field has a null / zero default");

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Wave-aggregated append: one atomic per distinct key in the wave, all of
// them issued together (the lowest lane of each key's group adds the group's
// count to that key's counter in one vector atomic), so the wave waits for
// one round trip however many keys it holds.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counters, uint32_t key, bool valid, uint32_t stride = 1) {
    const uint32_t lane = lane_id();
    unsigned long long same = 0;                 // the lanes holding this lane's key
    unsigned long long pending = __ballot(valid);
    while (pending) {
        const int leader = __ffsll(pending) - 1;
        const uint32_t k = __shfl(key, leader);
        const unsigned long long m = __ballot(valid && key == k);
        if (valid && key == k) same = m;
        pending &= ~m;
    }
    const int first = valid ? __ffsll(same) - 1 : (int)lane;
    uint32_t base = 0;
    if (valid && (int)lane == first) base = atomicAdd(&counters[key * stride], (uint32_t)__popcll(same));
    base = __shfl(base, first);
    return valid ? base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull)) : 0u;
}

// Reserve `a` (0 or 1) slots in the count at pair[0] and `b` (0 to 3) in the
// count at pair[1] for every lane of the wave with one 64-bit atomic (pair
// 8-byte aligned; the first count never reaches 2^32, so no carry crosses).
// Returns each lane's first slot in the two lists.  Every lane of the wave
// must call it (the prefix sums are over the wave).
__device__ __forceinline__ uint2 wave_reserve2(uint32_t* pair, uint32_t a, uint32_t b) {
    const uint32_t lane = lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const unsigned long long ma = __ballot(a != 0u), m1 = __ballot((b & 1u) != 0u), m2 = __ballot((b & 2u) != 0u);
    const uint32_t ta = (uint32_t)__popcll(ma), tb = (uint32_t)__popcll(m1) + 2u * (uint32_t)__popcll(m2);
    const int leader = __ffsll(__ballot(true)) - 1;
    unsigned long long base = 0;
    if ((int)lane == leader && (ta | tb) != 0u)
        base = atomicAdd(reinterpret_cast<unsigned long long*>(pair), ((unsigned long long)tb << 32) | ta);
    base = __shfl(base, leader);
    return make_uint2((uint32_t)base + (uint32_t)__popcll(ma & lt),
                      (uint32_t)(base >> 32) + (uint32_t)__popcll(m1 & lt) + 2u * (uint32_t)__popcll(m2 & lt));
}

__device__ __forceinline__ void wf_load(const WfPath& w, PathState& ps, Rng& rng) {
    ps.ray = make_ray(mk(w.o_time.x, w.o_time.y, w.o_time.z), mk(w.d_wl.x, w.d_wl.y, w.d_wl.z), w.o_time.w);
    ps.ray.wavelength = w.d_wl.w;
    ps.ray.spectral = (w.flags & 1u) != 0;
    ps.ray.wl = (uint8_t)((w.flags >> 2) & 3u);
    ps.ray.tint = (uint8_t)((w.flags >> 4) & 3u);
    ps.thr = mk(w.thr_mis.x, w.thr_mis.y, w.thr_mis.z);
    ps.prev_mis = w.thr_mis.w;
    ps.prev_sun_w = w.sun_w;
    ps.prev_specular = (w.flags & 2u) != 0;
    ps.knn_irr = mk(__uint_as_float(w.pad1[0]), __uint_as_float(w.pad1[1]), __uint_as_float(w.pad1[2]));
    rng.s0 = w.rng[0]; rng.s1 = w.rng[1]; rng.s2 = w.rng[2]; rng.s3 = w.rng[3];
}
__device__ __forceinline__ void spec4_load(const WfQueues& Q, uint32_t pid, Spec4& s) {
    float4 a = Q.spec[2 * (size_t)pid], b = Q.spec[2 * (size_t)pid + 1];
    s.q[0] = a.x; s.q[1] = a.y; s.q[2] = a.z; s.r[0] = a.w; s.r[1] = b.x; s.r[2] = b.y; s.on = b.z != 0.0f;
}
__device__ __forceinline__ void spec4_store(const WfQueues& Q, uint32_t pid, const Spec4& s) {
    Q.spec[2 * (size_t)pid] = make_float4(s.q[0], s.q[1], s.q[2], s.r[0]);
    Q.spec[2 * (size_t)pid + 1] = make_float4(s.r[1], s.r[2], s.on ? 1.0f : 0.0f, 0.0f);
}
__device__ __forceinline__ void wf_store(WfPath& w, const PathState& ps, const Rng& rng) {
    w.o_time = make_float4(ps.ray.o.x, ps.ray.o.y, ps.ray.o.z, ps.ray.time);
    w.d_wl = make_float4(ps.ray.d.x, ps.ray.d.y, ps.ray.d.z, ps.ray.wavelength);
    w.thr_mis = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, ps.prev_mis);
    w.sun_w = ps.prev_sun_w;
    w.flags = (ps.ray.spectral ? 1u : 0u) | (ps.prev_specular ? 2u : 0u) | ((uint32_t)ps.ray.wl << 2) | ((uint32_t)ps.ray.tint << 4);
    w.rng[0] = rng.s0; w.rng[1] = rng.s1; w.rng[2] = rng.s2; w.rng[3] = rng.s3;
}


// One render of a batch: its camera and pass parameters, its slice of the
// global path index space and of the concatenated frame accumulator.
struct DView {
    DCamera cam;
    DParams P;
    uint64_t first, count;    // global path range [first, first + count)
    uint32_t fb_off, npix;
};
struct WfBatch {
    const DScene* scenes;     // per view
    const DView* views;
    uint32_t nviews;
};

// A view's scene record with every pointer re-typed as global memory.  The
// records sit in an array in device memory, so a pointer loaded from one is
// generic and every access through it would be a flat access (which also
// waits on the LDS counter); rebuilt from its integer value as an
// address-space-1 pointer, the accesses compile to global loads.
template <class T>
__device__ __forceinline__ const T* as_global(const T* p) {
    typedef const __attribute__((address_space(1))) T gT;
    return (const T*)(gT*)(uintptr_t)p;
}
__device__ __forceinline__ DScene view_scene(const WfBatch& B, uint32_t view) {
    DScene S = B.scenes[view];
    S.nodes = as_global(S.nodes); S.objs = as_global(S.objs); S.lights = as_global(S.lights);
    S.xforms = as_global(S.xforms); S.tris = as_global(S.tris); S.tri_tag = as_global(S.tri_tag);
    S.tri_n = as_global(S.tri_n); S.tri_t = as_global(S.tri_t); S.tri_uv = as_global(S.tri_uv);
    S.quads = as_global(S.quads); S.planes = as_global(S.planes); S.mats = as_global(S.mats);
    S.media = as_global(S.media); S.grids = as_global(S.grids); S.tex_pix = as_global(S.tex_pix);
    S.srgb_lut = as_global(S.srgb_lut); S.nacre = as_global(S.nacre);
    S.env_aprob = as_global(S.env_aprob); S.env_alias = as_global(S.env_alias); S.env_rgb = as_global(S.env_rgb);
    S.env_marg = as_global(S.env_marg); S.env_cond = as_global(S.env_cond); S.elut = as_global(S.elut);
    S.ph_pos = as_global(S.ph_pos); S.ph_pow = as_global(S.ph_pow); S.ph_nrm = as_global(S.ph_nrm);
    S.pk_rec = as_global(S.pk_rec); S.pk_pow = as_global(S.pk_pow); S.pk_split = as_global(S.pk_split);
    S.pk_leaf = as_global(S.pk_leaf);
    return S;
}

// Path j of a sample pass covers pixel tile_order_pixel(j): bands of 8 rows,
// each walked in 8-column groups, row-major inside a group, so a wave's 64
// camera rays leave an 8 x 8 block of the image (coherent traversal and
// photon lookups) instead of a 64-pixel row.  A bijection of [0, W * H); the
// pixel's random stream depends on the pixel, not on j.
__device__ __forceinline__ uint32_t tile_order_pixel(uint32_t j, uint32_t W, uint32_t H) {
    uint32_t band = j / (8u * W);
    uint32_t hb = min(8u, H - 8u * band);
    uint32_t k = j - band * 8u * W;
    uint32_t grp = k / (8u * hb);
    uint32_t gw = min(8u, W - 8u * grp);
    uint32_t local = k - grp * 8u * hb;
    uint32_t row = local / gw;
    return (band * 8u + row) * W + grp * 8u + (local - row * gw);
}

// Whether the path's new hit takes a k-NN photon estimate in shade: a surface
// (not a medium event) whose material can take the diffuse branch.
__device__ __forceinline__ bool knn_gathers(const DScene& S, const WfPath& w) {
    if (S.ph_knn == 0 || S.num_photons == 0 || w.hit_slot < 0 || w.hit_prim == -2) return false;
    uint32_t kind = S.mats[S.objs[w.hit_slot].material].kind;
    return kind == M_LAMBERTIAN || kind == M_PBR || kind == M_PEARL;
}

// ---- wavefront path replay (adjoint) ---------------------------------------------------------
// Per path slot, beside its WfPath: the radiance still to be collected, the
// loss gradient of its sample, the pending score of the last scattered vertex
// (applied when the next vertex is shaded, after that vertex's shadow rays
// have been subtracted) and the vertex score of the current vertex's NEE.
struct __align__(16) WfAdj {
    float4 rem;                       // xyz
    float4 delta;                     // xyz
    int32_t pn, vn;
    int8_t pp[RT_SCORE_MAX], vp[3];
    int8_t pad0[7];
    float pv[3 * RT_SCORE_MAX];
    float vv[9];
    float pad1[3];
};
static_assert(sizeof(WfAdj) == 176, "WfAdj layout");

struct WfAdjQ {
    WfAdj* adj;
    const float* lsamp;               // primal radiance per sample of the chunk
    const float* dimg;                // dLoss/dI per pixel
    const int8_t* fmap;
    const MedPar* mp;                 // parameterised media
    int nmp;
    float inv_spp;
    uint32_t nparams;
    float* grad;
};
