// librtamd: the C ABI of include/rt_api.h on MI355X (gfx950).
//
// Host code: contexts, scene upload (flattening, BVH build, tables), photon
// map construction, and the launches of the kernels, which live in the k_*.hip
// translation units (declared in rt_kdecl.h) so they compile in parallel.
#include "rt_kern.h"
#include "rt_kdecl.h"
#include "bvh_build.h"

#include <string>
#include <vector>
#include <cstring>
#include <cstdio>
#include <memory>
#include <algorithm>
#include <cmath>
#include <mutex>

using namespace rt;


// ============================================================================ host side
namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    float* elut = nullptr;
    float norms[3], fire_norms[3];
    float norms_w[3], fire_norms_w[3];     // over the wide table (380-780 nm)
    // measurement
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double kernel_ms = 0.0;
    uint64_t launches = 0;
    unsigned long long* counters = nullptr;   // non-null: counting pass
    // wavefront buffers (grown on demand, kept across frames)
    void* wf_mem = nullptr;
    size_t wf_bytes = 0;
    uint32_t wf_cap = 0;
    // per-launch events around k_wf_extend (timing on, graphs off)
    std::vector<hipEvent_t> ext_ev;           // five per timed bounce (wf_chunk)
    size_t ext_used = 0;
    std::vector<bool> ext_knn;                // per timed bounce: the k-NN stage ran
    double stage_ms[4] = {0.0, 0.0, 0.0, 0.0};   // wavefront stages: extend, k-NN, shade, shadow
    uint64_t stage_n[4] = {0, 0, 0, 0};
    uint32_t* prog_ctl = nullptr;             // rt_render_progressive: stopped, converged count, passes run
    // extra wavefront streams and their queues (rt_wavefront_opts.streams > 1)
    hipStream_t xstream[RT_WF_MAX_STREAMS - 1] = {};
    void* xwf_mem[RT_WF_MAX_STREAMS - 1] = {};
    size_t xwf_bytes[RT_WF_MAX_STREAMS - 1] = {};
    hipEvent_t ev_fork = nullptr, ev_join[RT_WF_MAX_STREAMS - 1] = {};
    // this device's copy of the file-scope `scratch` array (rt_render_wavefront_noalloc)
    void* scratch = nullptr;
    bool no_alloc = false;                    // inside rt_render_wavefront_noalloc: no event creation
    int mk_occ = 0;                           // register budget of the plain megakernel (rt_ctx_set_megakernel_occupancy)
};

struct rt_scene {
    rt_ctx* ctx = nullptr;
    DScene S;
    std::vector<void*> allocs;
    uint64_t stats[8] = {0};
    std::vector<float4> photons_host;
    std::vector<DMaterial> mats_host;     // the materials as on the device (rt_scene_set_params edits both)
    std::vector<DMedium> media_host;      // the media as on the device (rt_scene_set_params edits densities)
    std::vector<uint32_t> media_src;      // per device medium: descriptor medium index
    std::vector<uint8_t> media_xf;        // per device medium: boundary under transforms
    std::vector<int32_t> media_slot;      // per device medium: leaf slot of its object
    uint32_t desc_media = 0;
    void* ph_mem = nullptr;                // photon map (positions, powers, normals), replaced on rebuild
    float world_min[3] = {0, 0, 0}, world_max[3] = {0, 0, 0};
    // photon_bounds of the world (src/bvh.rs:206-208): the union of the
    // objects' boxes without the infinite planes' proxies
    bool have_pb = false;
    float pb_min[3] = {0, 0, 0}, pb_max[3] = {0, 0, 0};
    ~rt_scene() { for (void* p : allocs) (void)hipFree(p); if (ph_mem) (void)hipFree(ph_mem); }
};

static int fail(rt_ctx* c, int code, const std::string& msg) { if (c) c->err = msg; return code; }
#define HIPCHECK(ctx, call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return fail(ctx, RT_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); } while (0)

// Alias table over env texels with weights lum * sin(theta_row) (the density
// of EnvMapData, src/renderer.rs:32-74): Vose's O(n) construction in double,
// small and large worklists taken in index order, so the table is a pure
// function of the image.
static bool env_alias_table(const float* rgb, uint32_t w, uint32_t h, std::vector<float>& prob, std::vector<uint32_t>& alias) {
    size_t n = (size_t)w * h;
    std::vector<double> q(n);
    double total = 0.0;
    for (uint32_t j = 0; j < h; j++) {
        float sj = std::sin(PI * ((float)j + 0.5f) / (float)h);
        for (uint32_t i = 0; i < w; i++) {
            const float* p = rgb + ((size_t)j * w + i) * 3;
            float wt = lum(mk(p[0], p[1], p[2])) * sj;
            q[(size_t)j * w + i] = wt > 0.0f ? (double)wt : 0.0;
            total += q[(size_t)j * w + i];
        }
    }
    if (!(total > 0.0)) return false;
    prob.assign(n, 1.0f);
    alias.resize(n);
    std::vector<uint32_t> small, large;
    for (size_t t = 0; t < n; t++) {
        q[t] = q[t] * (double)n / total;
        alias[t] = (uint32_t)t;
        (q[t] < 1.0 ? small : large).push_back((uint32_t)t);
    }
    size_t si = 0, li = 0;
    while (si < small.size() && li < large.size()) {
        uint32_t a = small[si++], g = large[li];
        prob[a] = (float)q[a];
        alias[a] = g;
        q[g] = (q[g] + q[a]) - 1.0;
        if (q[g] < 1.0) { li++; small.push_back(g); }
    }
    return true;
}

static bool is_device_ptr(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) { (void)hipGetLastError(); return false; }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

template <class T>
static int upload(rt_scene* s, const T* host, size_t count, const T** out) {
    *out = nullptr;
    if (count == 0) return 0;
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, count * sizeof(T));
    if (e != hipSuccess) return fail(s->ctx, RT_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    s->allocs.push_back(d);
    e = hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(s->ctx, RT_ERR_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
    *out = (const T*)d;
    return 0;
}

// Staging for caller buffers that may be host or device memory.
struct Staged {
    void* dev = nullptr;
    bool owned = false;
    ~Staged() { if (owned && dev) (void)hipFree(dev); }
};
static int stage_in(rt_ctx* c, const void* p, size_t bytes, Staged& s) {
    if (!p || bytes == 0) { s.dev = nullptr; return 0; }
    if (is_device_ptr(p)) { s.dev = const_cast<void*>(p); return 0; }
    HIPCHECK(c, hipMalloc(&s.dev, bytes));
    s.owned = true;
    HIPCHECK(c, hipMemcpy(s.dev, p, bytes, hipMemcpyHostToDevice));
    return 0;
}
static int stage_out(rt_ctx* c, void* p, size_t bytes, Staged& s, bool copy_in) {
    if (is_device_ptr(p)) { s.dev = p; return 0; }
    HIPCHECK(c, hipMalloc(&s.dev, bytes));
    s.owned = true;
    if (copy_in) HIPCHECK(c, hipMemcpy(s.dev, p, bytes, hipMemcpyHostToDevice));
    return 0;
}
static int finish_out(rt_ctx* c, void* p, size_t bytes, Staged& s) {
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (s.owned) HIPCHECK(c, hipMemcpy(p, s.dev, bytes, hipMemcpyDeviceToHost));
    return 0;
}

// ---- spectral normalisation (src/spectrum.rs:80-95, 162-178), host f32 ----------
// n = RT_CIE_N: the reference's table (380-700 nm); n = RT_CIE_WIDE_N: the
// same normalisation over 380-780 nm for the wide wavelength modes.
static void spectral_norms(float norms[3], float fire[3], int n = RT_CIE_N) {
    float s[3] = {0, 0, 0}, f[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) {
        float rgb[3]; xyz_srgb(RT_CIE_WX[i], RT_CIE_WY[i], RT_CIE_WZ[i], rgb);
        s[0] += rgb[0]; s[1] += rgb[1]; s[2] += rgb[2];
        float c[3]; fire_color(RT_CIE_WX[i], RT_CIE_WY[i], RT_CIE_WZ[i], c);
        f[0] += c[0]; f[1] += c[1]; f[2] += c[2];
    }
    for (int k = 0; k < 3; k++) { norms[k] = (float)n / std::fmax(s[k], 1e-6f); fire[k] = (float)n / std::fmax(f[k], 1e-6f); }
}

// ---- wavefront buffer layout ------------------------------------------------------------------
// One block per lane: path records, the active / next / dense queues, the
// per-key buckets, the shadow requests and their index list, then 4 KB of
// counters.
constexpr size_t wf_al(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t wf_layout_bytes(uint32_t cap) {
    return wf_al((size_t)cap * sizeof(WfPath)) + 3 * wf_al((size_t)cap * 4) + wf_al((size_t)WF_NB * cap * 4) +
           wf_al((size_t)3 * cap * sizeof(WfShadow)) + wf_al((size_t)3 * cap * 4) + wf_al((size_t)2 * cap * sizeof(float4)) + 4 * WF_CNT_WORDS;
}
static void wf_layout(char* m, uint32_t cap, WfQueues& Q) {
    Q.paths = (WfPath*)m; m += wf_al((size_t)cap * sizeof(WfPath));
    Q.active = (uint32_t*)m; m += wf_al((size_t)cap * 4);
    Q.next = (uint32_t*)m; m += wf_al((size_t)cap * 4);
    Q.dense = (uint32_t*)m; m += wf_al((size_t)cap * 4);
    Q.buckets = (uint32_t*)m; m += wf_al((size_t)WF_NB * cap * 4);
    Q.shadow = (WfShadow*)m; m += wf_al((size_t)3 * cap * sizeof(WfShadow));
    Q.shadow_idx = (uint32_t*)m; m += wf_al((size_t)3 * cap * 4);
    Q.spec = (float4*)m; m += wf_al((size_t)2 * cap * sizeof(float4));
    Q.counters = (uint32_t*)m;
    Q.cap = cap;
    Q.fb = nullptr;
    Q.lsamp = nullptr;        // per-sample contributions: the adjoint's primal pass, the wavefront render loop
    Q.stop = nullptr;
    Q.conv = nullptr;
    Q.sort = 1;
}

// The no-allocation wavefront's queues: a file-scope device array of one
// lane's buffers at RT_NOALLOC_CAPACITY paths (53.5 MiB), held by every GPU the
// library runs on until the process ends.  Host code never uses the name
// itself as a device address: rt_ctx_create looks up this device's copy once
// (hipGetSymbolAddress, with the device current) and keeps it in ctx->scratch.
static_assert(RT_NOALLOC_CAPACITY % BLOCK == 0, "capacity is whole blocks");
static __device__ __align__(256) char scratch[wf_layout_bytes(RT_NOALLOC_CAPACITY)];

static int scratch_lookup(rt_ctx* c) {
    static std::mutex mu;
    static void* per_device[256] = {};
    std::lock_guard<std::mutex> lock(mu);
    if (c->device >= 256) return RT_OK;                 // no copy: the no-allocation entry point refuses
    if (!per_device[c->device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return fail(c, RT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        per_device[c->device] = ptr;
    }
    c->scratch = per_device[c->device];
    return RT_OK;
}

extern "C" {

int rt_api_version(void) { return RT_API_VERSION; }

const char* rt_last_error(rt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rt_ctx_create(int device, rt_ctx** out) {
    if (!out) return RT_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { (void)hipGetLastError(); return RT_ERR_DEVICE; }
    if (device < 0 || device >= n) return RT_ERR_INVALID;
    std::unique_ptr<rt_ctx> c(new rt_ctx());
    c->device = device;
    HIPCHECK(c.get(), hipSetDevice(device));
    HIPCHECK(c.get(), hipStreamCreate(&c->stream));
    if (int rc = scratch_lookup(c.get())) return rc;
    HIPCHECK(c.get(), hipMalloc(&c->elut, 64 * 64 * sizeof(float)));
    hipLaunchKernelGGL(k_elut, dim3(64 * 64 / 64), dim3(64), 0, c->stream, c->elut);
    HIPCHECK(c.get(), hipGetLastError());
    HIPCHECK(c.get(), hipStreamSynchronize(c->stream));
    spectral_norms(c->norms, c->fire_norms);
    spectral_norms(c->norms_w, c->fire_norms_w, RT_CIE_WIDE_N);
    *out = c.release();
    return RT_OK;
}

void rt_ctx_destroy(rt_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->elut) (void)hipFree(ctx->elut);
    if (ctx->counters) (void)hipFree(ctx->counters);
    if (ctx->wf_mem) (void)hipFree(ctx->wf_mem);
    if (ctx->prog_ctl) (void)hipFree(ctx->prog_ctl);
    for (int k = 0; k < RT_WF_MAX_STREAMS - 1; k++) {
        if (ctx->xwf_mem[k]) (void)hipFree(ctx->xwf_mem[k]);
        if (ctx->ev_join[k]) (void)hipEventDestroy(ctx->ev_join[k]);
        if (ctx->xstream[k]) (void)hipStreamDestroy(ctx->xstream[k]);
    }
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    for (hipEvent_t e : ctx->ext_ev) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // extern "C"

// ---- scene flattening ------------------------------------------------------------------
namespace {

using rtb::Box;

Box sphere_box(const float* p) {
    Box b;
    for (int k = 0; k < 3; k++) { b.mn[k] = p[k] - p[3]; b.mx[k] = p[k] + p[3]; }
    if (p[7] != 0.0f) {
        Box b1;
        for (int k = 0; k < 3; k++) { b1.mn[k] = p[4 + k] - p[3]; b1.mx[k] = p[4 + k] + p[3]; }
        b = rtb::box_union(b, b1);
    }
    return b;
}

DQuad make_quad(const float q[3], const float u[3], const float v[3]) {
    DQuad Q;
    memset(&Q, 0, sizeof Q);
    for (int k = 0; k < 3; k++) { Q.q[k] = q[k]; Q.u[k] = u[k]; Q.v[k] = v[k]; }
    V3 U = mk(u[0], u[1], u[2]), Vv = mk(v[0], v[1], v[2]), Qp = mk(q[0], q[1], q[2]);
    V3 n = cross(U, Vv);
    float area = std::sqrt(len2(n));
    V3 nn = n / area;
    float d = dot(nn, Qp);
    V3 w = n / dot(n, n);
    Q.n[0] = nn.x; Q.n[1] = nn.y; Q.n[2] = nn.z;
    Q.w[0] = w.x; Q.w[1] = w.y; Q.w[2] = w.z;
    Q.d = d; Q.area = area;
    return Q;
}

Box quad_box(const DQuad& Q) {
    float c[4][3];
    for (int k = 0; k < 3; k++) {
        c[0][k] = Q.q[k]; c[1][k] = Q.q[k] + Q.u[k]; c[2][k] = Q.q[k] + Q.v[k]; c[3][k] = (Q.q[k] + Q.u[k]) + Q.v[k];
    }
    Box b;
    for (int k = 0; k < 3; k++) { b.mn[k] = c[0][k]; b.mx[k] = c[0][k]; }
    for (int i = 1; i < 4; i++)
        for (int k = 0; k < 3; k++) { b.mn[k] = std::fmin(b.mn[k], c[i][k]); b.mx[k] = std::fmax(b.mx[k], c[i][k]); }
    return rtb::box_pad(b);
}

struct Flat {
    std::vector<rtb::Node> nodes;
    std::vector<DObject> objs_input;     // input order
    std::vector<Box> obj_boxes;
    std::vector<uint32_t> obj_stack;     // nested traversal extra stack
    std::vector<DObject> lights;
    std::vector<DXform> xforms;
    std::vector<float4> tris;
    std::vector<int32_t> tri_tag;
    std::vector<float> tri_n, tri_t, tri_uv;
    std::vector<DQuad> quads;
    std::vector<float4> planes;
    std::vector<DMedium> media;
    std::vector<uint32_t> media_src;      // per device medium: its rt_scene_desc.media index
    std::vector<uint8_t> media_xf;        // per device medium: 1 when its boundary carries transforms
    std::vector<float> grids;
    std::string err;
};

// spectral_to_rgb (src/spectrum.rs:104-114) on the host: the CIE table lerp,
// XYZ -> linear sRGB clamped, times the per-channel normalisation.
static void host_spectral_rgb(const float norms[3], float lambda, float out[3], int n = RT_CIE_N) {
    float t = std::fmin(std::fmax((lambda - RT_CIE_START) / RT_CIE_STEP, 0.0f), (float)(n - 1));
    int i0 = std::min((int)t, n - 2);
    float fr = t - (float)i0;
    float x = RT_CIE_WX[i0] + fr * (RT_CIE_WX[i0 + 1] - RT_CIE_WX[i0]);
    float y = RT_CIE_WY[i0] + fr * (RT_CIE_WY[i0 + 1] - RT_CIE_WY[i0]);
    float z = RT_CIE_WZ[i0] + fr * (RT_CIE_WZ[i0 + 1] - RT_CIE_WZ[i0]);
    float c[3]; xyz_srgb(x, y, z, c);
    for (int a = 0; a < 3; a++) out[a] = c[a] * norms[a];
}

// NACRE_LUT (src/material.rs:1091-1117): the thin-film interference integral
// over 65 wavelengths at 256 optical path differences in [0, 2500] nm.
static std::vector<float> nacre_table(const float norms[3]) {
    std::vector<float> lut(256 * 3);
    for (int i = 0; i < 256; i++) {
        float opd = (float)i / 255.0f * 2500.0f;
        float c[3] = {0, 0, 0};
        float lambda = 380.0f;
        while (lambda <= 700.01f) {
            float irid = 0.5f * (1.0f + std::cos(2.0f * PI * opd / lambda));
            float s[3]; host_spectral_rgb(norms, lambda, s);
            for (int a = 0; a < 3; a++) c[a] += s[a] * irid;
            lambda += 5.0f;
        }
        for (int a = 0; a < 3; a++) lut[3 * i + a] = c[a] / 65.0f;
    }
    return lut;
}

// Perlin tables (ranvec[256][3], perm_x, perm_y, perm_z as floats): the
// permutations index the tables, so they must be integers in 0..255.
static bool perlin_perms_ok(const float* t) {
    for (int k = 768; k < 1536; k++)
        if (!(t[k] >= 0.0f && t[k] <= 255.0f && t[k] == std::floor(t[k]))) return false;
    return true;
}

bool flatten_object(Flat& F, const rt_scene_desc& d, const rt_object& o, DObject& D, Box& bb, uint32_t& extra, bool is_light) {
    memset(&D, 0, sizeof D);
    D.shape = o.shape;
    D.material = o.material;
    for (int k = 0; k < 16; k++) D.p[k] = o.p[k];
    extra = 0;
    if (o.medium == 0 && o.material >= d.num_materials) { F.err = "material index out of range"; return false; }
    switch (o.shape) {
        case RT_SHAPE_SPHERE:
            bb = sphere_box(o.p);
            break;
        case RT_SHAPE_QUAD: {
            D.data_first = (uint32_t)F.quads.size();
            DQuad Q = make_quad(o.p, o.p + 3, o.p + 6);
            F.quads.push_back(Q);
            bb = quad_box(Q);
            break;
        }
        case RT_SHAPE_BOX: {
            if (is_light) { F.err = "box lights are not supported"; return false; }
            float mn[3], mx[3];
            for (int k = 0; k < 3; k++) { mn[k] = std::fmin(o.p[k], o.p[3 + k]); mx[k] = std::fmax(o.p[k], o.p[3 + k]); }
            float dx[3] = {mx[0] - mn[0], 0, 0}, dy[3] = {0, mx[1] - mn[1], 0}, dz[3] = {0, 0, mx[2] - mn[2]};
            float ndx[3] = {-dx[0], -0.0f, -0.0f}, ndz[3] = {-0.0f, -0.0f, -dz[2]};
            float c0[3] = {mn[0], mn[1], mx[2]}, c1[3] = {mx[0], mn[1], mx[2]}, c2[3] = {mx[0], mn[1], mn[2]};
            float c3[3] = {mn[0], mn[1], mn[2]}, c4[3] = {mn[0], mx[1], mx[2]}, c5[3] = {mn[0], mn[1], mn[2]};
            DQuad side[6] = {make_quad(c0, dx, dy), make_quad(c1, ndz, dy), make_quad(c2, ndx, dy),
                             make_quad(c3, dz, dy), make_quad(c4, dx, ndz), make_quad(c5, dx, dz)};
            std::vector<Box> boxes(6);
            for (int k = 0; k < 6; k++) { side[k].pad[0] = (float)k; boxes[k] = quad_box(side[k]); }
            std::vector<uint32_t> leaf;
            D.root = (uint32_t)rtb::Builder::build_tree(boxes, F.nodes, leaf, &bb);
            D.data_first = (uint32_t)F.quads.size();
            D.data_count = 6;
            for (uint32_t k : leaf) F.quads.push_back(side[k]);
            extra = std::max(1u, rtb::Builder::stack_bound(F.nodes, (int32_t)D.root));
            break;
        }
        case RT_SHAPE_MESH: {
            if (is_light) { F.err = "mesh lights are not supported"; return false; }
            if (!d.tri_v || o.data_count == 0 || (uint64_t)o.data_first + o.data_count > d.num_tris) { F.err = "mesh triangle range"; return false; }
            std::vector<Box> boxes(o.data_count);
            for (uint32_t k = 0; k < o.data_count; k++) {
                const float* v = d.tri_v + ((size_t)o.data_first + k) * 9;
                Box b;
                for (int a = 0; a < 3; a++) {
                    b.mn[a] = std::fmin(std::fmin(v[a], v[3 + a]), v[6 + a]);
                    b.mx[a] = std::fmax(std::fmax(v[a], v[3 + a]), v[6 + a]);
                }
                boxes[k] = rtb::box_pad(b);
            }
            std::vector<uint32_t> leaf;
            leaf.reserve(o.data_count);
            D.root = (uint32_t)rtb::Builder::build_tree(boxes, F.nodes, leaf, &bb);
            D.data_first = (uint32_t)F.tri_tag.size();
            D.data_count = o.data_count;
            for (uint32_t k : leaf) {
                size_t t = (size_t)o.data_first + k;
                const float* v = d.tri_v + t * 9;
                F.tris.push_back(make_float4(v[0], v[1], v[2], 0.0f));
                F.tris.push_back(make_float4(v[3] - v[0], v[4] - v[1], v[5] - v[2], 0.0f));
                F.tris.push_back(make_float4(v[6] - v[0], v[7] - v[1], v[8] - v[2], 0.0f));
                F.tri_tag.push_back((int32_t)k);
                if (d.tri_n) F.tri_n.insert(F.tri_n.end(), d.tri_n + t * 9, d.tri_n + t * 9 + 9);
                if (d.tri_t) F.tri_t.insert(F.tri_t.end(), d.tri_t + t * 9, d.tri_t + t * 9 + 9);
                if (d.tri_uv) F.tri_uv.insert(F.tri_uv.end(), d.tri_uv + t * 6, d.tri_uv + t * 6 + 6);
            }
            extra = std::max(1u, rtb::Builder::stack_bound(F.nodes, (int32_t)D.root));
            break;
        }
        case RT_SHAPE_CONVEX: {
            if ((uint64_t)(o.data_first + o.data_count) * 4 > d.pool_len || !d.pool) { F.err = "convex plane range"; return false; }
            D.data_first = (uint32_t)F.planes.size();
            D.data_count = o.data_count;
            for (uint32_t k = 0; k < o.data_count; k++) {
                const float* q = d.pool + ((size_t)o.data_first + k) * 4;
                F.planes.push_back(make_float4(q[0], q[1], q[2], q[3]));
            }
            for (int k = 0; k < 3; k++) { bb.mn[k] = o.p[7 + k]; bb.mx[k] = o.p[10 + k]; }
            break;
        }
        case RT_SHAPE_CYLINDER: case RT_SHAPE_CONE: {
            if (!(o.p[3] > 0.0f) || !(o.p[4] > 0.0f)) { F.err = "cylinder / cone radius and height must be positive"; return false; }
            if (is_light) { F.err = "cylinder / cone lights are not supported (no pdf in the reference)"; return false; }
            // bounding_box (src/cylinder.rs:217-223, src/cone.rs:124-129)
            for (int k = 0; k < 3; k++) {
                float rk = k == 1 ? 0.0f : o.p[3];
                bb.mn[k] = o.p[k] - rk;
                bb.mx[k] = (o.p[k] + rk) + (k == 1 ? o.p[4] : 0.0f);
            }
            break;
        }
        case RT_SHAPE_DISK: case RT_SHAPE_PLANE: {
            float nx = o.p[3], ny = o.p[4], nz = o.p[5];
            float l2 = nx * nx + ny * ny + nz * nz;
            if (!(l2 > 1e-12f)) { F.err = "disk / plane normal must be nonzero"; return false; }
            // Vec3::unit and Vec3::onb (src/vec3.rs:30-42)
            float l = std::sqrt(l2);
            V3 n = mk(nx / l, ny / l, nz / l);
            V3 up = std::fabs(n.y) < 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
            V3 c = cross(n, up);
            float cl = std::sqrt(len2(c));
            V3 u = mk(c.x / cl, c.y / cl, c.z / cl);
            V3 v = cross(n, u);
            D.p[3] = n.x; D.p[4] = n.y; D.p[5] = n.z;
            if (o.shape == RT_SHAPE_DISK) {
                if (!(o.p[6] > 0.0f)) { F.err = "disk radius must be positive"; return false; }
                float rad = o.p[6];
                D.p[7] = u.x; D.p[8] = u.y; D.p[9] = u.z; D.p[10] = v.x; D.p[11] = v.y; D.p[12] = v.z;
                D.p[13] = PI * rad * rad;
                // bounding_box: extent R sqrt(1 - n_i^2), padded (src/disk.rs:172-181)
                float nn[3] = {n.x, n.y, n.z};
                Box b;
                for (int k = 0; k < 3; k++) {
                    float e = rad * std::sqrt(std::fmax(1.0f - nn[k] * nn[k], 0.0f));
                    b.mn[k] = o.p[k] - e; b.mx[k] = o.p[k] + e;
                }
                bb = rtb::box_pad(b);
            } else {
                if (is_light) { F.err = "plane lights are not supported"; return false; }
                D.p[6] = o.p[6]; D.p[7] = o.p[7];
                D.p[8] = u.x; D.p[9] = u.y; D.p[10] = u.z; D.p[11] = v.x; D.p[12] = v.y; D.p[13] = v.z;
                // bounding_box (src/plane.rs:86-98): thin slab across the normal axis, 1e5 elsewhere
                float nn[3] = {n.x, n.y, n.z};
                for (int k = 0; k < 3; k++) {
                    if (std::fabs(nn[k]) > 0.999f) { bb.mn[k] = o.p[k] - 1e-3f; bb.mx[k] = o.p[k] + 1e-3f; }
                    else { bb.mn[k] = -1e5f; bb.mx[k] = 1e5f; }
                }
            }
            break;
        }
        default:
            F.err = "unsupported shape kind " + std::to_string(o.shape);
            return false;
    }
    if (o.medium) {
        if (is_light) { F.err = "media cannot be lights"; return false; }
        if (o.medium > d.num_media || !d.media) { F.err = "medium index"; return false; }
        if (o.shape != RT_SHAPE_SPHERE && o.shape != RT_SHAPE_BOX) { F.err = "medium boundary must be a sphere or a box"; return false; }
        const rt_medium& m = d.media[o.medium - 1];
        DMedium M;
        memset(&M, 0, sizeof M);
        M.kind = m.kind; M.g = m.g;
        for (int k = 0; k < 3; k++) { M.albedo[k] = m.albedo[k]; M.bmin[k] = bb.mn[k]; M.bmax[k] = bb.mx[k]; }
        if (m.kind == RT_MEDIUM_CONSTANT) M.density = m.density;
        else if (m.kind == RT_MEDIUM_GRID) {
            if (!m.grid || !m.grid_n[0] || !m.grid_n[1] || !m.grid_n[2]) { F.err = "grid medium without grid"; return false; }
            M.density = std::fmax(m.density, 1.1920929e-7f);
            M.nx = m.grid_n[0]; M.ny = m.grid_n[1]; M.nz = m.grid_n[2];
            M.grid_off = (uint32_t)F.grids.size();
            F.grids.insert(F.grids.end(), m.grid, m.grid + (size_t)M.nx * M.ny * M.nz);
        } else if (m.kind == RT_MEDIUM_NOISE) {
            if (!m.perlin) { F.err = "noise medium without Perlin tables"; return false; }
            if (!perlin_perms_ok(m.perlin)) { F.err = "noise medium permutation entries must be integers in 0..255"; return false; }
            // NoiseMedium::new (src/volume.rs:170-188)
            M.density = std::fmax(m.density, 1.1920929e-7f);
            M.threshold = std::fmin(std::fmax(m.threshold, 0.0f), 0.99f);
            M.noise_scale = m.noise_scale;
            M.grid_off = (uint32_t)F.grids.size();
            F.grids.insert(F.grids.end(), m.perlin, m.perlin + 1536);
        } else { F.err = "medium kind"; return false; }
        F.media.push_back(M);
        F.media_src.push_back(o.medium - 1);
        F.media_xf.push_back(o.xform_count ? 1 : 0);
        D.medium = (uint32_t)F.media.size();
    }
    D.xf_first = (uint32_t)F.xforms.size();
    D.xf_count = o.xform_count;
    for (uint32_t k = 0; k < o.xform_count; k++) {
        if (!d.xforms || o.xform_first + k >= d.num_xforms) { F.err = "xform range"; return false; }
        const rt_xform& x = d.xforms[o.xform_first + k];
        DXform X;
        X.kind = x.kind;
        for (int j = 0; j < 9; j++) X.v[j] = x.v[j];
        if (x.kind == RT_XFORM_TRANSLATE) {
            for (int a = 0; a < 3; a++) { bb.mn[a] = bb.mn[a] + x.v[a]; bb.mx[a] = bb.mx[a] + x.v[a]; }
        } else if (x.kind == RT_XFORM_ROTATE) {
            Box r;
            for (int a = 0; a < 3; a++) { r.mn[a] = FINF; r.mx[a] = -FINF; }
            for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) for (int k = 0; k < 2; k++) {
                V3 c = mk(i == 0 ? bb.mn[0] : bb.mx[0], j == 0 ? bb.mn[1] : bb.mx[1], k == 0 ? bb.mn[2] : bb.mx[2]);
                V3 q = rot_apply(x.v, c);
                r.mn[0] = std::fmin(r.mn[0], q.x); r.mx[0] = std::fmax(r.mx[0], q.x);
                r.mn[1] = std::fmin(r.mn[1], q.y); r.mx[1] = std::fmax(r.mx[1], q.y);
                r.mn[2] = std::fmin(r.mn[2], q.z); r.mx[2] = std::fmax(r.mx[2], q.z);
            }
            bb = r;
        } else if (x.kind == RT_XFORM_SCALE) {
            if (!(x.v[0] > 0.0f)) { F.err = "UniformScale factor must be positive"; return false; }
            for (int a = 0; a < 3; a++) { bb.mn[a] = x.v[0] * bb.mn[a]; bb.mx[a] = x.v[0] * bb.mx[a]; }
        } else { F.err = "xform kind"; return false; }
        F.xforms.push_back(X);
    }
    return true;
}

}  // namespace

template <class T>
static int upload_vec(rt_scene* s, const std::vector<T>& v, const T** out) { return upload(s, v.data(), v.size(), out); }

extern "C" int rt_scene_upload(rt_ctx* ctx, const rt_scene_desc* d, rt_scene** out) {
    if (!ctx || !d || !out) return RT_ERR_INVALID;
    *out = nullptr;
    if (d->num_objects == 0 || !d->objects) return fail(ctx, RT_ERR_INVALID, "empty world");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    Flat F;
    for (uint32_t i = 0; i < d->num_materials; i++) {
        uint32_t k = d->materials[i].kind;
        if (k < RT_MAT_LAMBERTIAN || k > RT_MAT_PEARL) return fail(ctx, RT_ERR_UNSUPPORTED, "unsupported material kind " + std::to_string(k));
    }
    F.objs_input.resize(d->num_objects);
    F.obj_boxes.resize(d->num_objects);
    F.obj_stack.resize(d->num_objects);
    for (uint32_t i = 0; i < d->num_objects; i++) {
        if (!flatten_object(F, *d, d->objects[i], F.objs_input[i], F.obj_boxes[i], F.obj_stack[i], false))
            return fail(ctx, F.err.find("unsupported") != std::string::npos || F.err.find("not supported") != std::string::npos ? RT_ERR_UNSUPPORTED : RT_ERR_INVALID, F.err);
        F.objs_input[i].tag = i;
    }
    std::vector<uint32_t> leaf;
    rtb::Box wb;
    size_t world_first = F.nodes.size();
    int32_t root = rtb::Builder::build_tree(F.obj_boxes, F.nodes, leaf, &wb);
    // the kernels start the world traversal at node 0: move the world tree there
    if (world_first != 0 || root != 0) {
        // renumber: world nodes are appended after nested trees; swap blocks so
        // the world tree starts at index 0 and fix child indices.
        size_t nested = world_first;
        size_t world_n = F.nodes.size() - world_first;
        std::vector<rtb::Node> re;
        re.reserve(F.nodes.size());
        auto remap = [&](int32_t c, bool in_world) -> int32_t {
            if (c < 0) return c;
            if (in_world) return c - (int32_t)nested;
            return c + (int32_t)world_n;
        };
        for (size_t k = world_first; k < F.nodes.size(); k++) {
            rtb::Node n = F.nodes[k];
            for (int c = 0; c < 8; c++) n.child[c] = remap(n.child[c], true);
            re.push_back(n);
        }
        for (size_t k = 0; k < nested; k++) {
            rtb::Node n = F.nodes[k];
            for (int c = 0; c < 8; c++) n.child[c] = remap(n.child[c], false);
            re.push_back(n);
        }
        F.nodes.swap(re);
        for (auto& o : F.objs_input) if (o.shape == RT_SHAPE_MESH || o.shape == RT_SHAPE_BOX) o.root += (uint32_t)world_n;
        root = remap(root, true);
        if (root != 0) return fail(ctx, RT_ERR_INVALID, "world root is not the first node");
    }
    std::vector<DObject> objs_leaf;
    std::vector<uint32_t> leaf_extra;
    for (uint32_t k : leaf) { objs_leaf.push_back(F.objs_input[k]); leaf_extra.push_back(F.obj_stack[k]); }
    uint32_t bound = rtb::Builder::stack_bound(F.nodes, 0, &leaf_extra);
    bound = std::max(bound, 1u);
    if (bound > RT_LDS_STACK + RT_PRIV_STACK)
        return fail(ctx, RT_ERR_UNSUPPORTED, "BVH needs a traversal stack of " + std::to_string(bound) + " entries");
    for (uint32_t i = 0; i < d->num_lights; i++) {
        DObject L; rtb::Box lb; uint32_t ex;
        if (!flatten_object(F, *d, d->lights[i], L, lb, ex, true)) return fail(ctx, RT_ERR_UNSUPPORTED, F.err);
        L.tag = i;
        F.lights.push_back(L);
    }
    std::unique_ptr<rt_scene> s(new rt_scene());
    s->ctx = ctx;
    DScene& S = s->S;
    memset(&S, 0, sizeof S);
    int rc = 0;
    const float4* nodes = nullptr;
    static_assert(sizeof(rtb::Node) == 16 * sizeof(float4), "node layout");
    if ((rc = upload(s.get(), reinterpret_cast<const float4*>(F.nodes.data()), F.nodes.size() * 16, &nodes))) return rc;
    S.nodes = nodes;
    if ((rc = upload_vec(s.get(), objs_leaf, &S.objs))) return rc;
    S.num_objs = (uint32_t)objs_leaf.size();
    if ((rc = upload_vec(s.get(), F.lights, &S.lights))) return rc;
    S.num_lights = (uint32_t)F.lights.size();
    if ((rc = upload_vec(s.get(), F.xforms, &S.xforms))) return rc;
    if ((rc = upload_vec(s.get(), F.tris, &S.tris))) return rc;
    if ((rc = upload_vec(s.get(), F.tri_tag, &S.tri_tag))) return rc;
    if ((rc = upload_vec(s.get(), F.tri_n, &S.tri_n))) return rc;
    if ((rc = upload_vec(s.get(), F.tri_t, &S.tri_t))) return rc;
    if ((rc = upload_vec(s.get(), F.tri_uv, &S.tri_uv))) return rc;
    if ((rc = upload_vec(s.get(), F.quads, &S.quads))) return rc;
    if ((rc = upload_vec(s.get(), F.planes, &S.planes))) return rc;
    if ((rc = upload_vec(s.get(), F.media, &S.media))) return rc;
    s->media_host = F.media; s->media_src = F.media_src; s->media_xf = F.media_xf; s->desc_media = d->num_media;
    s->media_slot.assign(F.media.size(), -1);
    for (size_t k = 0; k < objs_leaf.size(); k++)
        if (objs_leaf[k].medium) s->media_slot[objs_leaf[k].medium - 1] = (int32_t)k;
    std::vector<DMaterial> mats(d->num_materials);
    std::vector<uint8_t> texels;
    std::vector<int64_t> image_off(d->num_images, -1), perlin_off(d->num_perlin, -1);
    for (uint32_t i = 0; i < d->num_materials; i++) {
        const rt_material& dm = d->materials[i];
        memset(&mats[i], 0, sizeof(DMaterial));
        mats[i].kind = dm.kind; mats[i].tex = dm.tex;
        for (int k = 0; k < 24; k++) mats[i].p[k] = dm.p[k];
        // an image of the descriptor in the texel pool (each image once)
        auto place_image = [&](uint32_t k, uint32_t& off, uint32_t& w, uint32_t& h) -> int {
            const rt_image& im = d->images[k];
            if (!im.rgb || !im.width || !im.height) return fail(ctx, RT_ERR_INVALID, "material " + std::to_string(i) + ": empty image");
            size_t bytes = (size_t)im.width * im.height * 3;
            if (image_off[k] < 0) {
                if (texels.size() + bytes > 0xffffffffull) return fail(ctx, RT_ERR_UNSUPPORTED, "texture images exceed 4 GiB");
                image_off[k] = (int64_t)texels.size();
                texels.insert(texels.end(), im.rgb, im.rgb + bytes);
            }
            off = (uint32_t)image_off[k]; w = im.width; h = im.height;
            return RT_OK;
        };
        if (dm.kind == RT_MAT_PBR) {
            for (int k = 0; k < 5; k++) {
                float f = dm.p[17 + k];
                if (f == 0.0f) continue;
                if (!d->images || !(f >= 1.0f && f <= (float)d->num_images && f == std::floor(f)))
                    return fail(ctx, RT_ERR_INVALID, "material " + std::to_string(i) + ": PBR map image index");
                int rc2 = place_image((uint32_t)f - 1, mats[i].map_off[k], mats[i].map_w[k], mats[i].map_h[k]);
                if (rc2) return rc2;
            }
            if (dm.p[15] > 0.0f && !(dm.p[16] > 0.0f)) return fail(ctx, RT_ERR_INVALID, "material " + std::to_string(i) + ": film_ior must be positive");
        }
        if (dm.tex != RT_TEX_SOLID) {
            std::string tag = "material " + std::to_string(i) + ": ";
            if (dm.kind != RT_MAT_LAMBERTIAN && dm.kind != RT_MAT_DIFFUSE_LIGHT)
                return fail(ctx, RT_ERR_INVALID, tag + "only LAMBERTIAN and DIFFUSE_LIGHT take a texture");
            if (dm.tex > RT_TEX_MOTTLE) return fail(ctx, RT_ERR_UNSUPPORTED, tag + "unsupported texture kind " + std::to_string(dm.tex));
            if (dm.tex == RT_TEX_IMAGE) {
                float f = dm.p[7];
                if (!d->images || !(f >= 0.0f && f < (float)d->num_images && f == std::floor(f))) return fail(ctx, RT_ERR_INVALID, tag + "image index");
                int rc2 = place_image((uint32_t)f, mats[i].toff, mats[i].tw, mats[i].th);
                if (rc2) return rc2;
            } else if (dm.tex == RT_TEX_NOISE || dm.tex == RT_TEX_MOTTLE) {
                float f = dm.p[8];
                if (!d->perlin || !(f >= 0.0f && f < (float)d->num_perlin && f == std::floor(f))) return fail(ctx, RT_ERR_INVALID, tag + "Perlin table index");
                uint32_t k = (uint32_t)f;
                const float* t = d->perlin + (size_t)1536 * k;
                if (!perlin_perms_ok(t)) return fail(ctx, RT_ERR_INVALID, tag + "Perlin permutation entries must be integers in 0..255");
                if (perlin_off[k] < 0) {
                    perlin_off[k] = (int64_t)F.grids.size();
                    F.grids.insert(F.grids.end(), t, t + 1536);
                }
                mats[i].toff = (uint32_t)perlin_off[k];
            }
        }
        if (mats[i].kind == RT_MAT_BLACKBODY) {
            // BlackbodyLight::new (src/material.rs:49-61): normalising mean of
            // Planck over the 65 CIE samples and the white-point average colour
            float T = mats[i].p[0], intensity = mats[i].p[1];
            float raw[RT_CIE_N], sum = 0.0f;
            for (int j = 0; j < RT_CIE_N; j++) { raw[j] = planck_raw(380.0f + (float)j * 5.0f, T); sum += raw[j]; }
            float mean = sum / (float)RT_CIE_N;
            float norm = 1.0f / std::fmax(mean, 1e-30f);
            float acc[3] = {0, 0, 0};
            for (int j = 0; j < RT_CIE_N; j++) {
                float c[3]; host_spectral_rgb(ctx->norms, 380.0f + (float)j * 5.0f, c);
                float w = raw[j] * norm;
                for (int a = 0; a < 3; a++) acc[a] += c[a] * w;
            }
            mats[i].p[2] = norm;
            for (int a = 0; a < 3; a++) mats[i].p[3 + a] = acc[a] / (float)RT_CIE_N * intensity;
            // the same over the table extended to 780 nm (the wide wavelength
            // modes): p6 normalising mean, p7-9 average colour
            float raw_w[RT_CIE_WIDE_N], sum_w = 0.0f;
            for (int j = 0; j < RT_CIE_WIDE_N; j++) { raw_w[j] = planck_raw(380.0f + (float)j * 5.0f, T); sum_w += raw_w[j]; }
            float norm_w = 1.0f / std::fmax(sum_w / (float)RT_CIE_WIDE_N, 1e-30f);
            float acc_w[3] = {0, 0, 0};
            for (int j = 0; j < RT_CIE_WIDE_N; j++) {
                float c[3]; host_spectral_rgb(ctx->norms_w, 380.0f + (float)j * 5.0f, c, RT_CIE_WIDE_N);
                for (int a = 0; a < 3; a++) acc_w[a] += c[a] * raw_w[j] * norm_w;
            }
            mats[i].p[6] = norm_w;
            for (int a = 0; a < 3; a++) mats[i].p[7 + a] = acc_w[a] / (float)RT_CIE_WIDE_N * intensity;
        }
    }
    if ((rc = upload_vec(s.get(), F.grids, &S.grids))) return rc;
    if ((rc = upload_vec(s.get(), texels, &S.tex_pix))) return rc;
    {
        // SRGB_DECODE_LUT (src/output.rs:88-95), built with the host's powf
        std::vector<float> lut(256);
        for (int c = 0; c < 256; c++) {
            float v = (float)c / 255.0f;
            lut[c] = v <= 0.04045f ? v / 12.92f : std::pow((v + 0.055f) / 1.055f, 2.4f);
        }
        if ((rc = upload_vec(s.get(), lut, &S.srgb_lut))) return rc;
    }
    if ((rc = upload_vec(s.get(), nacre_table(ctx->norms), &S.nacre))) return rc;
    if ((rc = upload_vec(s.get(), mats, &S.mats))) return rc;
    s->mats_host = mats;
    // background
    const rt_background& b = d->background;
    S.bg_kind = b.kind; S.bg_visible = b.visible_kind;
    for (int k = 0; k < 3; k++) {
        S.bg_color[k] = b.color[k]; S.bg_top[k] = b.top[k]; S.bg_bottom[k] = b.bottom[k]; S.sun_dir[k] = b.sun_dir[k];
    }
    S.turbidity = b.turbidity;
    // render_tiles sets the pearl sun from a Physical background (src/renderer.rs:676-679)
    S.pearl_sun_on = b.kind == RT_BG_PHYSICAL ? 1u : 0u;
    for (int k = 0; k < 3; k++) S.pearl_sun[k] = b.sun_dir[k];
    if (b.kind == RT_BG_ENVMAP || b.kind == RT_BG_ENVMAP_SPLIT) {
        if (!b.env_rgb || !b.env_w || !b.env_h) return fail(ctx, RT_ERR_INVALID, "env map missing");
        uint32_t w = b.env_w, h = b.env_h;
        // EnvMapData::new (src/renderer.rs:32-74) in f32
        std::vector<float> cond((size_t)w * h), row(h), marg(h);
        for (uint32_t j = 0; j < h; j++) {
            float sj = std::sin(PI * ((float)j + 0.5f) / (float)h);
            float sum = 0.0f;
            for (uint32_t i = 0; i < w; i++) {
                const float* p = b.env_rgb + ((size_t)j * w + i) * 3;
                sum += lum(mk(p[0], p[1], p[2])) * sj;
                cond[(size_t)j * w + i] = sum;
            }
            if (sum > 0.0f) for (uint32_t i = 0; i < w; i++) cond[(size_t)j * w + i] /= sum;
            else for (uint32_t i = 0; i < w; i++) cond[(size_t)j * w + i] = (float)(i + 1) / (float)w;
            row[j] = sum;
        }
        float total = 0.0f;
        for (uint32_t j = 0; j < h; j++) total += row[j];
        float cum = 0.0f;
        for (uint32_t j = 0; j < h; j++) { cum += row[j]; marg[j] = total > 0.0f ? cum / total : (float)(j + 1) / (float)h; }
        S.env_w = w; S.env_h = h; S.env_rot = b.env_rot_y; S.env_inv_total = total > 0.0f ? 1.0f / total : 0.0f;
        if ((rc = upload(s.get(), b.env_rgb, (size_t)w * h * 3, &S.env_rgb))) return rc;
        if ((rc = upload_vec(s.get(), cond, &S.env_cond))) return rc;
        if ((rc = upload_vec(s.get(), marg, &S.env_marg))) return rc;
        S.env_sampling = b.env_sampling == RT_ENV_ALIAS ? 1u : 0u;
        if (b.env_sampling == RT_ENV_ALIAS) {
            std::vector<float> aprob;
            std::vector<uint32_t> alias;
            if (!env_alias_table(b.env_rgb, w, h, aprob, alias)) return fail(ctx, RT_ERR_INVALID, "env map has no energy");
            if ((rc = upload_vec(s.get(), aprob, &S.env_aprob))) return rc;
            if ((rc = upload_vec(s.get(), alias, &S.env_alias))) return rc;
        } else if (b.env_sampling != RT_ENV_CDF) return fail(ctx, RT_ERR_INVALID, "env_sampling");
    }
    S.elut = ctx->elut;
    for (int k = 0; k < 3; k++) {
        S.norms[k] = ctx->norms[k]; S.fire_norms[k] = ctx->fire_norms[k];
        S.norms_w[k] = ctx->norms_w[k]; S.fire_norms_w[k] = ctx->fire_norms_w[k];
    }
    S.stack_bound = bound;
    for (int k = 0; k < 3; k++) { s->world_min[k] = wb.mn[k]; s->world_max[k] = wb.mx[k]; }
    for (uint32_t i = 0; i < d->num_objects; i++) {
        if (d->objects[i].shape == RT_SHAPE_PLANE) continue;        // InfinitePlane::photon_bounds = None (src/plane.rs:102-106)
        const rtb::Box& b = F.obj_boxes[i];
        for (int k = 0; k < 3; k++) {
            s->pb_min[k] = s->have_pb ? std::fmin(s->pb_min[k], b.mn[k]) : b.mn[k];
            s->pb_max[k] = s->have_pb ? std::fmax(s->pb_max[k], b.mx[k]) : b.mx[k];
        }
        s->have_pb = true;
    }
    size_t node_count = F.nodes.size();
    s->stats[0] = node_count;                 // nodes
    s->stats[1] = F.tri_tag.size();           // triangles
    s->stats[2] = node_count * 256;           // node bytes
    s->stats[3] = bound;                      // stack bound
    s->stats[4] = d->num_objects;
    s->stats[5] = F.tris.size() * sizeof(float4);
    *out = s.release();
    return RT_OK;
}

extern "C" {

void rt_scene_destroy(rt_scene* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    delete s;
}

int rt_scene_stats(rt_scene* s, uint64_t stats[8]) {
    if (!s || !stats) return RT_ERR_INVALID;
    memcpy(stats, s->stats, sizeof s->stats);
    return RT_OK;
}

static int trace_common(rt_ctx* ctx, rt_scene* s, const rt_ray* rays, uint32_t n, void* out, int mode) {
    if (!ctx || !s || (!rays && n) || (!out && n)) return RT_ERR_INVALID;
    if (n == 0) return RT_OK;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    Staged rin, rout;
    int rc;
    if ((rc = stage_in(ctx, rays, (size_t)n * sizeof(rt_ray), rin))) return rc;
    size_t ob = mode == 0 ? sizeof(rt_hit) : (mode == 1 ? 1 : sizeof(float));
    if ((rc = stage_out(ctx, out, (size_t)n * ob, rout, false))) return rc;
    dim3 grid((n + BLOCK - 1) / BLOCK);
    if (mode == 0) hipLaunchKernelGGL(k_trace_closest, grid, dim3(BLOCK), 0, ctx->stream, s->S, (const rt_ray*)rin.dev, n, (rt_hit*)rout.dev);
    else if (mode == 1) hipLaunchKernelGGL(k_trace_tr, grid, dim3(BLOCK), 0, ctx->stream, s->S, (const rt_ray*)rin.dev, n, (float*)nullptr, (uint8_t*)rout.dev);
    else hipLaunchKernelGGL(k_trace_tr, grid, dim3(BLOCK), 0, ctx->stream, s->S, (const rt_ray*)rin.dev, n, (float*)rout.dev, (uint8_t*)nullptr);
    HIPCHECK(ctx, hipGetLastError());
    return finish_out(ctx, out, (size_t)n * ob, rout);
}

int rt_trace_closest(rt_ctx* ctx, rt_scene* s, const rt_ray* rays, uint32_t n, rt_hit* hits) { return trace_common(ctx, s, rays, n, hits, 0); }
int rt_trace_any(rt_ctx* ctx, rt_scene* s, const rt_ray* rays, uint32_t n, uint8_t* occ) { return trace_common(ctx, s, rays, n, occ, 1); }
int rt_transmittance(rt_ctx* ctx, rt_scene* s, const rt_ray* rays, uint32_t n, float* tr) { return trace_common(ctx, s, rays, n, tr, 2); }

// rt_sample_params.wavelengths -> Ray::wl: 0 the reference's one wavelength
// over [380, 700), 1 one over [380, 780), 4 four hero wavelengths over
// [380, 780) (wl 2); -1 for any other value.
static int wl_mode(const rt_sample_params* p) {
    return p->wavelengths == 0 ? 0 : p->wavelengths == 1 ? 1 : p->wavelengths == 4 ? 2 : -1;
}
static void to_dev(const rt_camera* c, const rt_sample_params* p, DCamera& C, DParams& P) {
    for (int k = 0; k < 3; k++) {
        C.origin[k] = c->origin[k]; C.lower_left[k] = c->lower_left[k]; C.horizontal[k] = c->horizontal[k];
        C.vertical[k] = c->vertical[k]; C.u[k] = c->u[k]; C.v[k] = c->v[k];
    }
    C.lens_radius = c->lens_radius; C.blades = c->aperture_blades;
    P.sample_idx = p->sample_idx; P.strata = p->strata; P.width = p->width; P.height = p->height;
    P.max_depth = p->max_depth; P.use_frame_time = p->use_frame_time; P.use_photons = p->use_photons;
    P.bg_scale = p->bg_scale; P.t0 = p->frame_t0; P.t1 = p->frame_t1; P.seed = p->seed;
    P.wl = (uint32_t)std::max(wl_mode(p), 0);
}
extern "C++" {
template <int OCC>
static void launch_render_occ(bool knn, bool h4, dim3 grid, hipStream_t st, const DScene& S, const DCamera& C,
                              const DParams& Q, const uint8_t* conv, float* out, uint32_t m, int acc) {
    if (knn && h4) hipLaunchKernelGGL((k_render<OCC, true, true>), grid, dim3(BLOCK), 0, st, S, C, Q, conv, out, m, acc);
    else if (knn) hipLaunchKernelGGL((k_render<OCC, true, false>), grid, dim3(BLOCK), 0, st, S, C, Q, conv, out, m, acc);
    else if (h4) hipLaunchKernelGGL((k_render<OCC, false, true>), grid, dim3(BLOCK), 0, st, S, C, Q, conv, out, m, acc);
    else hipLaunchKernelGGL((k_render<OCC, false, false>), grid, dim3(BLOCK), 0, st, S, C, Q, conv, out, m, acc);
}
}
static void launch_render(int occ, bool knn, bool h4, dim3 grid, hipStream_t st, const DScene& S, const DCamera& C,
                          const DParams& Q, const uint8_t* conv, float* out, uint32_t m, int acc) {
    if (occ == 2) launch_render_occ<2>(knn, h4, grid, st, S, C, Q, conv, out, m, acc);
    else launch_render_occ<0>(knn, h4, grid, st, S, C, Q, conv, out, m, acc);
}

static int render_common(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, const rt_sample_params* p, const uint8_t* conv,
                         float* out, uint32_t spp, bool accumulate) {
    if (!ctx || !s || !cam || !p || !out) return RT_ERR_INVALID;
    if (wl_mode(p) < 0) return fail(ctx, RT_ERR_INVALID, "wavelengths must be 0, 1 or 4");
    if (p->width == 0 || p->height == 0) return fail(ctx, RT_ERR_INVALID, "empty image");
    if ((uint64_t)p->width * p->height > 0xffffffffull / 4) return fail(ctx, RT_ERR_INVALID, "image too large");
    if (p->use_photons && s->S.num_photons == 0 && s->photons_host.empty()) return fail(ctx, RT_ERR_INVALID, "use_photons without a photon map");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    uint32_t npix = p->width * p->height;
    Staged cin, rout;
    int rc;
    if ((rc = stage_in(ctx, conv, npix, cin))) return rc;
    if ((rc = stage_out(ctx, out, (size_t)npix * 12, rout, false))) return rc;
    DCamera C; DParams P;
    to_dev(cam, p, C, P);
    P.counters = ctx->counters;
    const bool knn = p->use_photons && s->S.ph_knn;
    dim3 grid((npix + BLOCK - 1) / BLOCK);
    if (ctx->timing) HIPCHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    uint32_t nlaunch = 1;
    if (accumulate) {
        HIPCHECK(ctx, hipMemsetAsync(rout.dev, 0, (size_t)npix * 12, ctx->stream));
        // chunks of passes keep single launches short
        const uint32_t chunk = 4;
        for (uint32_t k = 0; k < spp; k += chunk) {
            DParams Q = P;
            Q.sample_idx = P.sample_idx + k;
            uint32_t m = std::min(chunk, spp - k);
            launch_render(ctx->mk_occ, knn, P.wl == 2, grid, ctx->stream, s->S, C, Q, (const uint8_t*)cin.dev, (float*)rout.dev, m, 1);
            HIPCHECK(ctx, hipGetLastError());
        }
        nlaunch = (spp + chunk - 1) / chunk;
    } else {
        launch_render(ctx->mk_occ, knn, P.wl == 2, grid, ctx->stream, s->S, C, P, (const uint8_t*)cin.dev, (float*)rout.dev, 1u, 0);
        HIPCHECK(ctx, hipGetLastError());
    }
    if (ctx->timing) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHECK(ctx, hipEventSynchronize(ctx->ev1));
        float ms = 0.0f;
        HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->kernel_ms += ms;
        ctx->launches += nlaunch;
    }
    return finish_out(ctx, out, (size_t)npix * 12, rout);
}

int rt_render_sample(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, const rt_sample_params* p, const uint8_t* conv, float* out) {
    return render_common(ctx, s, cam, p, conv, out, 1, false);
}

int rt_render_accumulate(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, const rt_sample_params* p, uint32_t spp, float* out) {
    return render_common(ctx, s, cam, p, nullptr, out, spp, true);
}

// Timing of one kernel launch on the context stream (rt_ctx_set_timing).
static int time_begin(rt_ctx* ctx) {
    if (ctx->timing) HIPCHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    return RT_OK;
}
static int time_end(rt_ctx* ctx) {
    if (!ctx->timing) return RT_OK;
    HIPCHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHECK(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->kernel_ms += ms;
    ctx->launches += 1;
    return RT_OK;
}

// The four arrays of a pixel state on the device (staged when the caller
// passed host memory, copied back by state_finish).
struct StateDev {
    Staged acc, ns, m2, conv;
    size_t npix = 0;
};
static int state_stage(rt_ctx* ctx, const rt_pixel_state* st, uint32_t npix, StateDev& d) {
    if (!st || !st->accum || !st->samples || !st->var_m2 || !st->converged) return fail(ctx, RT_ERR_INVALID, "pixel state with a null array");
    bool dev = is_device_ptr(st->accum);
    if (is_device_ptr(st->samples) != dev || is_device_ptr(st->var_m2) != dev || is_device_ptr(st->converged) != dev)
        return fail(ctx, RT_ERR_INVALID, "pixel state mixes host and device arrays");
    d.npix = npix;
    int rc;
    if ((rc = stage_out(ctx, st->accum, (size_t)npix * 12, d.acc, true))) return rc;
    if ((rc = stage_out(ctx, st->samples, (size_t)npix * 4, d.ns, true))) return rc;
    if ((rc = stage_out(ctx, st->var_m2, (size_t)npix * 12, d.m2, true))) return rc;
    if ((rc = stage_out(ctx, st->converged, (size_t)npix, d.conv, true))) return rc;
    return RT_OK;
}
static int state_finish(rt_ctx* ctx, const rt_pixel_state* st, StateDev& d) {
    int rc;
    if ((rc = finish_out(ctx, st->accum, d.npix * 12, d.acc))) return rc;
    if ((rc = finish_out(ctx, st->samples, d.npix * 4, d.ns))) return rc;
    if ((rc = finish_out(ctx, st->var_m2, d.npix * 12, d.m2))) return rc;
    return finish_out(ctx, st->converged, d.npix, d.conv);
}
static int check_accum_opts(rt_ctx* ctx, const rt_accum_opts* o) {
    if (!o) return fail(ctx, RT_ERR_INVALID, "null accumulation options");
    if (!std::isfinite(o->firefly_clamp) || !std::isfinite(o->firefly_abs_floor) || !std::isfinite(o->threshold) ||
        !std::isfinite(o->convergence_pct))
        return fail(ctx, RT_ERR_INVALID, "accumulation options must be finite");
    return RT_OK;
}
static int ctl_buffer(rt_ctx* ctx) {
    if (!ctx->prog_ctl) HIPCHECK(ctx, hipMalloc((void**)&ctx->prog_ctl, 4 * sizeof(uint32_t)));
    return RT_OK;
}

int rt_accumulate_sample(rt_ctx* ctx, const float* sc, uint32_t npix, const rt_accum_opts* o, const rt_pixel_state* st) {
    if (!ctx || (!sc && npix)) return RT_ERR_INVALID;
    int rc;
    if ((rc = check_accum_opts(ctx, o))) return rc;
    if (npix == 0) return RT_OK;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    Staged in;
    StateDev d;
    if ((rc = stage_in(ctx, sc, (size_t)npix * 12, in))) return rc;
    if ((rc = state_stage(ctx, st, npix, d))) return rc;
    if ((rc = time_begin(ctx))) return rc;
    hipLaunchKernelGGL(k_accumulate, dim3((npix + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, (const float*)in.dev, npix, *o,
                       (float*)d.acc.dev, (uint32_t*)d.ns.dev, (float*)d.m2.dev, (const uint8_t*)d.conv.dev);
    HIPCHECK(ctx, hipGetLastError());
    if ((rc = time_end(ctx))) return rc;
    return state_finish(ctx, st, d);
}

int rt_mark_converged(rt_ctx* ctx, uint32_t npix, const rt_accum_opts* o, const rt_pixel_state* st, uint32_t* n_converged) {
    if (!ctx) return RT_ERR_INVALID;
    int rc;
    if ((rc = check_accum_opts(ctx, o))) return rc;
    if (npix == 0) { if (n_converged) *n_converged = 0; return RT_OK; }
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    StateDev d;
    if ((rc = state_stage(ctx, st, npix, d))) return rc;
    if ((rc = ctl_buffer(ctx))) return rc;
    HIPCHECK(ctx, hipMemsetAsync(ctx->prog_ctl, 0, 4 * sizeof(uint32_t), ctx->stream));
    if ((rc = time_begin(ctx))) return rc;
    hipLaunchKernelGGL(k_mark_converged, dim3(std::min((npix + BLOCK - 1) / BLOCK, RT_COUNT_GRID)), dim3(BLOCK), 0, ctx->stream, npix, o->threshold,
                       (const float*)d.acc.dev, (const uint32_t*)d.ns.dev, (const float*)d.m2.dev, (uint8_t*)d.conv.dev, ctx->prog_ctl);
    HIPCHECK(ctx, hipGetLastError());
    if ((rc = time_end(ctx))) return rc;
    uint32_t cnt = 0;
    HIPCHECK(ctx, hipMemcpyAsync(&cnt, ctx->prog_ctl, 4, hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = state_finish(ctx, st, d))) return rc;
    if (n_converged) *n_converged = cnt;
    return RT_OK;
}

int rt_render_aux(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, uint32_t w, uint32_t h, float* albedo, float* normal) {
    if (!ctx || !s || !cam || !albedo || !normal) return RT_ERR_INVALID;
    if (w == 0 || h == 0) return fail(ctx, RT_ERR_INVALID, "empty image");
    if ((uint64_t)w * h > 0xffffffffull / 12) return fail(ctx, RT_ERR_INVALID, "image too large");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    uint32_t npix = w * h;
    Staged a, n;
    int rc;
    if ((rc = stage_out(ctx, albedo, (size_t)npix * 12, a, false))) return rc;
    if ((rc = stage_out(ctx, normal, (size_t)npix * 12, n, false))) return rc;
    rt_sample_params p;
    memset(&p, 0, sizeof p);
    DCamera C; DParams P;
    to_dev(cam, &p, C, P);
    if ((rc = time_begin(ctx))) return rc;
    hipLaunchKernelGGL(k_aux, dim3((npix + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, s->S, C, w, h,
                       (float*)a.dev, (float*)n.dev);
    HIPCHECK(ctx, hipGetLastError());
    if ((rc = time_end(ctx))) return rc;
    if ((rc = finish_out(ctx, albedo, (size_t)npix * 12, a))) return rc;
    return finish_out(ctx, normal, (size_t)npix * 12, n);
}

int rt_resolve_rgb8(rt_ctx* ctx, const float* acc, const uint32_t* ns, uint32_t spp, uint32_t w, uint32_t h,
                    const rt_resolve_opts* o, uint8_t* out) {
    if (!ctx || !o) return RT_ERR_INVALID;
    if ((uint64_t)w * h > 0xffffffffull / 12) return fail(ctx, RT_ERR_INVALID, "image too large");
    if (o->tonemapper != RT_TONE_ACES && o->tonemapper != RT_TONE_AGX) return fail(ctx, RT_ERR_INVALID, "unknown tone mapper");
    if (!std::isfinite(o->exposure) || !std::isfinite(o->vignette) || !std::isfinite(o->distortion))
        return fail(ctx, RT_ERR_INVALID, "resolve options must be finite");
    uint32_t npix = w * h;
    if (npix == 0) return RT_OK;
    if (!acc || !out) return RT_ERR_INVALID;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    Staged a, n, r;
    int rc;
    if ((rc = stage_in(ctx, acc, (size_t)npix * 12, a))) return rc;
    if ((rc = stage_in(ctx, ns, ns ? (size_t)npix * 4 : 0, n))) return rc;
    if ((rc = stage_out(ctx, out, (size_t)npix * 3, r, false))) return rc;
    if ((rc = time_begin(ctx))) return rc;
    hipLaunchKernelGGL(k_resolve, dim3((npix + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, (const float*)a.dev,
                       (const uint32_t*)n.dev, spp, w, h, *o, (uint8_t*)r.dev);
    HIPCHECK(ctx, hipGetLastError());
    if ((rc = time_end(ctx))) return rc;
    return finish_out(ctx, out, (size_t)npix * 3, r);
}

int rt_render_progressive(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, const rt_sample_params* p, uint32_t samples_max,
                          const rt_accum_opts* o, const rt_pixel_state* st, uint32_t* samples_done, uint32_t* n_converged) {
    if (!ctx || !s || !cam || !p) return RT_ERR_INVALID;
    if (wl_mode(p) < 0) return fail(ctx, RT_ERR_INVALID, "wavelengths must be 0, 1 or 4");
    int rc;
    if ((rc = check_accum_opts(ctx, o))) return rc;
    if (p->width == 0 || p->height == 0) return fail(ctx, RT_ERR_INVALID, "empty image");
    if ((uint64_t)p->width * p->height > 0xffffffffull / 12) return fail(ctx, RT_ERR_INVALID, "image too large");
    if (p->use_photons && s->S.num_photons == 0 && s->photons_host.empty()) return fail(ctx, RT_ERR_INVALID, "use_photons without a photon map");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    uint32_t npix = p->width * p->height;
    StateDev d;
    if ((rc = state_stage(ctx, st, npix, d))) return rc;
    if ((rc = ctl_buffer(ctx))) return rc;
    DCamera C; DParams P;
    to_dev(cam, p, C, P);
    P.counters = nullptr;
    const bool knn = p->use_photons && s->S.ph_knn;
    dim3 grid((npix + BLOCK - 1) / BLOCK);
    HIPCHECK(ctx, hipMemsetAsync(ctx->prog_ctl, 0, 4 * sizeof(uint32_t), ctx->stream));
    hipLaunchKernelGGL(k_count_flags, dim3(std::min((npix + BLOCK - 1) / BLOCK, RT_COUNT_GRID)), dim3(BLOCK), 0, ctx->stream, npix,
                       (const uint8_t*)d.conv.dev, ctx->prog_ctl + 1);
    if (ctx->timing) HIPCHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    float* acc = (float*)d.acc.dev; uint32_t* ns = (uint32_t*)d.ns.dev; float* m2 = (float*)d.m2.dev; uint8_t* conv = (uint8_t*)d.conv.dev;
    for (uint32_t k = 0; k < samples_max; k++) {
        DParams Q = P;
        Q.sample_idx = P.sample_idx + k;
        if (knn && Q.wl == 2) hipLaunchKernelGGL((k_progressive<true, true>), grid, dim3(BLOCK), 0, ctx->stream, s->S, C, Q, *o, k, acc, ns, m2, conv, ctx->prog_ctl);
        else if (knn) hipLaunchKernelGGL((k_progressive<true, false>), grid, dim3(BLOCK), 0, ctx->stream, s->S, C, Q, *o, k, acc, ns, m2, conv, ctx->prog_ctl);
        else if (Q.wl == 2) hipLaunchKernelGGL((k_progressive<false, true>), grid, dim3(BLOCK), 0, ctx->stream, s->S, C, Q, *o, k, acc, ns, m2, conv, ctx->prog_ctl);
        else hipLaunchKernelGGL((k_progressive<false, false>), grid, dim3(BLOCK), 0, ctx->stream, s->S, C, Q, *o, k, acc, ns, m2, conv, ctx->prog_ctl);
        hipLaunchKernelGGL(k_progressive_step, dim3(1), dim3(64), 0, ctx->stream, ctx->prog_ctl, k, npix, *o);
        HIPCHECK(ctx, hipGetLastError());
    }
    if (ctx->timing) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHECK(ctx, hipEventSynchronize(ctx->ev1));
        float ms = 0.0f;
        HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->kernel_ms += ms;
        ctx->launches += 2ull * samples_max;
    }
    uint32_t ctl[4] = {0, 0, 0, 0};
    HIPCHECK(ctx, hipMemcpyAsync(ctl, ctx->prog_ctl, sizeof ctl, hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = state_finish(ctx, st, d))) return rc;
    if (samples_done) *samples_done = ctl[2];
    if (n_converged) *n_converged = ctl[1];
    return RT_OK;
}


}  // extern "C"

// ---- wavefront renderer --------------------------------------------------------------------------
static int wf_buffers(rt_ctx* ctx, uint32_t cap, WfQueues& Q, int slot = 0) {
    size_t total = wf_layout_bytes(cap);
    void*& mem = slot ? ctx->xwf_mem[slot - 1] : ctx->wf_mem;
    size_t& bytes = slot ? ctx->xwf_bytes[slot - 1] : ctx->wf_bytes;
    if (bytes < total) {
        if (mem) HIPCHECK(ctx, hipFree(mem));
        mem = nullptr; bytes = 0;
        hipError_t e = hipMalloc(&mem, total);
        if (e != hipSuccess) return fail(ctx, RT_ERR_NOMEM, std::string("wavefront buffers: ") + hipGetErrorString(e));
        bytes = total;
    }
    wf_layout((char*)mem, cap, Q);
    ctx->wf_cap = cap;
    return RT_OK;
}

// The traversal stages at the chosen register budget and block size; grid is
// in blocks of 128 lanes, so the launch keeps the same lanes in flight.
template <bool BT, int BS>
static void launch_trace_bs(bool shadow, int occ, uint32_t grid, hipStream_t st, const DScene& S, const WfQueues& Q,
                            const WfBatch& B) {
    dim3 g(std::max(1u, (uint32_t)((uint64_t)grid * BLOCK / BS))), b(BS);
    if (!shadow) {
        if (occ >= 5) hipLaunchKernelGGL((k_wf_extend<5, BT, BS>), g, b, 0, st, S, Q, B);
        else if (occ == 4) hipLaunchKernelGGL((k_wf_extend<4, BT, BS>), g, b, 0, st, S, Q, B);
        else hipLaunchKernelGGL((k_wf_extend<0, BT, BS>), g, b, 0, st, S, Q, B);
    } else {
        if (occ >= 5) hipLaunchKernelGGL((k_wf_shadow<6, BT, BS>), g, b, 0, st, S, Q, B);
        else if (occ == 4) hipLaunchKernelGGL((k_wf_shadow<5, BT, BS>), g, b, 0, st, S, Q, B);
        else hipLaunchKernelGGL((k_wf_shadow<0, BT, BS>), g, b, 0, st, S, Q, B);
    }
}
template <bool BT>
static void launch_trace(bool shadow, int occ, uint32_t block, uint32_t grid, hipStream_t st, const DScene& S,
                         const WfQueues& Q, const WfBatch& B) {
    if (block == 64) launch_trace_bs<BT, 64>(shadow, occ, grid, st, S, Q, B);
    else if (block == 256) launch_trace_bs<BT, 256>(shadow, occ, grid, st, S, Q, B);
    else launch_trace_bs<BT, 128>(shadow, occ, grid, st, S, Q, B);
}

// Launch one chunk: generate `n` paths at global position `first` (~0: read
// from the counters, for graph replays) and run `depth` bounces.
template <bool BT>
static void wf_chunk(rt_ctx* ctx, hipStream_t st, const DScene& S, const DCamera& C, const DParams& P, const WfQueues& Q0,
                     uint64_t first, uint32_t n, uint32_t grid, int occ, uint32_t block, uint32_t depth, const WfBatch& Bt,
                     bool knn) {
    if (Q0.conv) hipLaunchKernelGGL(k_wf_gen_zero, dim3(1), dim3(64), 0, st, Q0.counters);
    hipLaunchKernelGGL((k_wf_generate<BT>), dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, C, P, Q0, first, n, Bt);
    hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(32), 0, st, Q0.counters, n, Q0.stop, Q0.conv ? 1 : 0);
    WfQueues A = Q0, B = Q0;
    B.active = Q0.next; B.next = Q0.active;
    for (uint32_t d = 0; d < depth; d++) {
        const WfQueues& Q = (d & 1) ? B : A;
        // per-stage events (timing on, no graph): before extend, after extend,
        // after the k-NN stage, after shade, after shadow
        bool ev = ctx->timing && !ctx->no_alloc && first != ~0ull;
        if (ev) {
            while (ctx->ext_ev.size() < ctx->ext_used + 5) { hipEvent_t e; if (hipEventCreate(&e) != hipSuccess) { ev = false; break; } ctx->ext_ev.push_back(e); }
        }
        hipEvent_t* E = ev ? &ctx->ext_ev[ctx->ext_used] : nullptr;
        if (ev) { ctx->ext_used += 5; ctx->ext_knn.push_back(knn); }
        if (ev) (void)hipEventRecord(E[0], st);
        launch_trace<BT>(false, occ, block, grid, st, S, Q, Bt);
        if (ev) (void)hipEventRecord(E[1], st);
        if (knn) {
            hipLaunchKernelGGL((k_wf_knn<BT>), dim3(grid), dim3(BLOCK), 0, st, S, P, Q, Bt);
            hipLaunchKernelGGL((k_wf_knn_dense<BT>), dim3(grid), dim3(BLOCK), 0, st, S, Q, Bt);
            hipLaunchKernelGGL((k_wf_knn_gather<BT>), dim3(grid), dim3(BLOCK), 0, st, S, Q, Bt);
            hipLaunchKernelGGL((k_wf_knn_wave<BT>), dim3(grid), dim3(BLOCK), 0, st, S, Q, Bt);
        }
        if (ev) (void)hipEventRecord(E[2], st);
        if (Q.sort == 2 && Q.h4) {
            hipLaunchKernelGGL((k_wf_shade<BT, 1, true>), dim3(grid), dim3(BLOCK), 0, st, S, P, Q, (int)d, Bt);
            hipLaunchKernelGGL((k_wf_shade<BT, 2, true>), dim3(grid), dim3(BLOCK), 0, st, S, P, Q, (int)d, Bt);
        } else if (Q.sort == 2) {
            hipLaunchKernelGGL((k_wf_shade<BT, 1, false>), dim3(grid), dim3(BLOCK), 0, st, S, P, Q, (int)d, Bt);
            hipLaunchKernelGGL((k_wf_shade<BT, 2, false>), dim3(grid), dim3(BLOCK), 0, st, S, P, Q, (int)d, Bt);
        } else if (Q.h4) {
            hipLaunchKernelGGL((k_wf_shade<BT, 0, true>), dim3(grid), dim3(BLOCK), 0, st, S, P, Q, (int)d, Bt);
        } else {
            hipLaunchKernelGGL((k_wf_shade<BT, 0, false>), dim3(grid), dim3(BLOCK), 0, st, S, P, Q, (int)d, Bt);
        }
        if (ev) (void)hipEventRecord(E[3], st);
        launch_trace<BT>(true, occ, block, grid, st, S, Q, Bt);
        if (ev) (void)hipEventRecord(E[4], st);
        hipLaunchKernelGGL(k_wf_advance, dim3(1), dim3(32), 0, st, Q.counters);
    }
}

struct WfOpts { uint32_t cap, grid, sort, use_graph, streams, block, knn_wave_max; int occ; };

static int wf_opts(rt_ctx* ctx, const rt_wavefront_opts* opts, uint64_t total, WfOpts& o) {
    o.cap = opts && opts->capacity ? opts->capacity : (1u << 22);
    o.grid = opts && opts->grid ? opts->grid : 2048u;
    o.sort = opts ? opts->sort : 1u;
    o.use_graph = opts ? opts->use_graph : 0u;
    o.occ = opts ? (int)opts->occupancy : 0;
    o.streams = opts && opts->streams > 1 ? opts->streams : 1u;
    o.block = opts && opts->block ? opts->block : (uint32_t)BLOCK;
    o.knn_wave_max = opts && opts->knn_wave_max ? opts->knn_wave_max : 65536u;
    if (o.block != 64 && o.block != 128 && o.block != 256) return fail(ctx, RT_ERR_INVALID, "wavefront block must be 64, 128 or 256");
    if (o.streams > RT_WF_MAX_STREAMS) return fail(ctx, RT_ERR_INVALID, "wavefront streams above RT_WF_MAX_STREAMS");
    if (o.cap < BLOCK || o.cap > (1u << 26)) return fail(ctx, RT_ERR_INVALID, "wavefront capacity out of range");
    if (o.sort > 2) return fail(ctx, RT_ERR_INVALID, "wavefront sort must be 0, 1 or 2");
    if (o.grid == 0 || o.grid > (1u << 20)) return fail(ctx, RT_ERR_INVALID, "wavefront grid out of range");
    if (o.cap > total) o.cap = (uint32_t)((total + BLOCK - 1) / BLOCK * BLOCK);
    return RT_OK;
}

// The extra streams and their events, created on first use.
static int wf_streams(rt_ctx* ctx, uint32_t n) {
    if (!ctx->ev_fork) HIPCHECK(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    for (uint32_t k = 0; k + 1 < n; k++) {
        if (!ctx->xstream[k]) HIPCHECK(ctx, hipStreamCreate(&ctx->xstream[k]));
        if (!ctx->ev_join[k]) HIPCHECK(ctx, hipEventCreateWithFlags(&ctx->ev_join[k], hipEventDisableTiming));
    }
    return RT_OK;
}

// Every chunk of the global path range through every bounce.  fb must be
// zeroed.  With L lanes (Qs[0..L), L > 1: more streams), chunk c runs on lane
// c % L: that lane's stream and queues; all add into the same accumulator.
// The extra streams start after the context stream's earlier work, and the
// context stream waits for all of them at the end.
template <bool BT>
static int wf_run(rt_ctx* ctx, const DScene& S, const DCamera& C, const DParams& P, WfQueues* Qs, uint32_t lanes,
                  const WfOpts& o, uint64_t total, uint32_t depth, const WfBatch& Bt, bool knn) {
    uint32_t cap = o.cap;
    uint64_t full = total / cap, rem = total % cap;
    uint64_t nchunks = full + (rem ? 1 : 0);
    const uint32_t L = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(lanes, nchunks));
    auto stream_of = [&](uint32_t k) { return k ? ctx->xstream[k - 1] : ctx->stream; };
    hipStream_t s0 = ctx->stream;
    if (ctx->timing) HIPCHECK(ctx, hipEventRecord(ctx->ev0, s0));
    if (L > 1) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev_fork, s0));
        for (uint32_t k = 1; k < L; k++) HIPCHECK(ctx, hipStreamWaitEvent(stream_of(k), ctx->ev_fork, 0));
    }
    if (o.use_graph && full > 0) {
        const uint32_t G = (uint32_t)std::min<uint64_t>(L, full);     // lanes that replay a captured chunk
        hipGraph_t graph[RT_WF_MAX_STREAMS] = {};
        hipGraphExec_t exec[RT_WF_MAX_STREAMS] = {};
        for (uint32_t k = 0; k < G; k++) {
            hipStream_t sk = stream_of(k);
            hipLaunchKernelGGL(k_wf_set_first, dim3(1), dim3(32), 0, sk, Qs[k].counters, (uint64_t)k * cap);
            HIPCHECK(ctx, hipStreamBeginCapture(sk, hipStreamCaptureModeThreadLocal));
            wf_chunk<BT>(ctx, sk, S, C, P, Qs[k], ~0ull, cap, o.grid, o.occ, o.block, depth, Bt, knn);
            hipLaunchKernelGGL(k_wf_bump_first, dim3(1), dim3(32), 0, sk, Qs[k].counters, G * cap);
            HIPCHECK(ctx, hipStreamEndCapture(sk, &graph[k]));
            HIPCHECK(ctx, hipGraphInstantiate(&exec[k], graph[k], nullptr, nullptr, 0));
        }
        for (uint64_t c = 0; c < full; c++) HIPCHECK(ctx, hipGraphLaunch(exec[c % G], stream_of((uint32_t)(c % G))));
        for (uint32_t k = 0; k < G; k++) {
            HIPCHECK(ctx, hipGraphExecDestroy(exec[k]));
            HIPCHECK(ctx, hipGraphDestroy(graph[k]));
        }
    } else {
        for (uint64_t c = 0; c < full; c++) {
            uint32_t k = (uint32_t)(c % L);
            wf_chunk<BT>(ctx, stream_of(k), S, C, P, Qs[k], c * (uint64_t)cap, cap, o.grid, o.occ, o.block, depth, Bt, knn);
        }
    }
    if (rem) {
        uint32_t k = (uint32_t)(full % L);
        wf_chunk<BT>(ctx, stream_of(k), S, C, P, Qs[k], full * (uint64_t)cap, (uint32_t)rem, o.grid, o.occ, o.block, depth, Bt, knn);
    }
    HIPCHECK(ctx, hipGetLastError());
    for (uint32_t k = 1; k < L; k++) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev_join[k - 1], stream_of(k)));
        HIPCHECK(ctx, hipStreamWaitEvent(s0, ctx->ev_join[k - 1], 0));
    }
    if (ctx->timing) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev1, s0));
        HIPCHECK(ctx, hipEventSynchronize(ctx->ev1));
        float ms = 0.0f;
        HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->kernel_ms += ms;
        ctx->launches += nchunks;   // one "launch" = one chunk of paths through every bounce
        for (size_t g = 0; 5 * g + 4 < ctx->ext_used; g++) {
            const hipEvent_t* E = &ctx->ext_ev[5 * g];
            for (int k = 0; k < 4; k++) {
                if (k == 1 && !ctx->ext_knn[g]) continue;
                float e = 0.0f;
                if (hipEventElapsedTime(&e, E[k], E[k + 1]) == hipSuccess) { ctx->stage_ms[k] += e; ctx->stage_n[k]++; }
            }
        }
    }
    ctx->ext_used = 0;
    ctx->ext_knn.clear();
    return RT_OK;
}

// Queues of every lane the options ask for (Qs[0] already set up).
static int wf_lanes(rt_ctx* ctx, const WfOpts& o, WfQueues* Qs) {
    if (o.streams <= 1) return RT_OK;
    int rc;
    if ((rc = wf_streams(ctx, o.streams))) return rc;
    for (uint32_t k = 1; k < o.streams; k++) {
        if ((rc = wf_buffers(ctx, o.cap, Qs[k], (int)k))) return rc;
        Qs[k].sort = Qs[0].sort;
        Qs[k].fb = Qs[0].fb;
        Qs[k].knn_wave_max = Qs[0].knn_wave_max;
        Qs[k].h4 = Qs[0].h4;
    }
    return RT_OK;
}

extern "C" int rt_render_wavefront(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, const rt_sample_params* p, uint32_t spp,
                                   float* out, const rt_wavefront_opts* opts) {
    if (!ctx || !s || !cam || !p || !out) return RT_ERR_INVALID;
    if (wl_mode(p) < 0) return fail(ctx, RT_ERR_INVALID, "wavelengths must be 0, 1 or 4");
    if (p->width == 0 || p->height == 0 || spp == 0) return fail(ctx, RT_ERR_INVALID, "empty frame");
    if (p->max_depth == 0) return fail(ctx, RT_ERR_INVALID, "max_depth = 0");
    if (p->use_photons && s->S.num_photons == 0) return fail(ctx, RT_ERR_INVALID, "use_photons without a photon map");
    if ((uint64_t)p->width * p->height > 0xffffffffull / 4) return fail(ctx, RT_ERR_INVALID, "image too large");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    uint32_t npix = p->width * p->height;
    uint64_t total = (uint64_t)npix * spp;
    WfOpts o;
    int rc;
    if ((rc = wf_opts(ctx, opts, total, o))) return rc;
    WfQueues Q;
    if ((rc = wf_buffers(ctx, o.cap, Q))) return rc;
    Q.sort = o.sort;
    Q.knn_wave_max = o.knn_wave_max;
    Staged rout;
    if ((rc = stage_out(ctx, out, (size_t)npix * 12, rout, false))) return rc;
    Q.fb = (float*)rout.dev;
    DCamera C; DParams P;
    to_dev(cam, p, C, P);
    P.counters = nullptr;
    Q.h4 = P.wl == 2;
    HIPCHECK(ctx, hipMemsetAsync(rout.dev, 0, (size_t)npix * 12, ctx->stream));
    WfBatch none = {nullptr, nullptr, 0};
    WfQueues Qs[RT_WF_MAX_STREAMS];
    Qs[0] = Q;
    if ((rc = wf_lanes(ctx, o, Qs))) return rc;
    if ((rc = wf_run<false>(ctx, s->S, C, P, Qs, o.streams, o, total, P.max_depth, none, P.use_photons && s->S.ph_knn))) return rc;
    return finish_out(ctx, out, (size_t)npix * 12, rout);
}

// rt_render_wavefront without any allocation: one lane of queues in this
// device's `scratch` array, the frame zeroed by a fill kernel, no streams,
// graphs or per-launch events created.  Calls overlapping on one device share
// the array and race.
extern "C" int rt_render_wavefront_noalloc(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, const rt_sample_params* p,
                                           uint32_t spp, float* out, const rt_wavefront_opts* opts) {
    if (!ctx || !s || !cam || !p || !out) return RT_ERR_INVALID;
    if (wl_mode(p) < 0) return fail(ctx, RT_ERR_INVALID, "wavelengths must be 0, 1 or 4");
    if (p->width == 0 || p->height == 0 || spp == 0) return fail(ctx, RT_ERR_INVALID, "empty frame");
    if (p->max_depth == 0) return fail(ctx, RT_ERR_INVALID, "max_depth = 0");
    if (p->use_photons && s->S.num_photons == 0) return fail(ctx, RT_ERR_INVALID, "use_photons without a photon map");
    if ((uint64_t)p->width * p->height > 0xffffffffull / 4) return fail(ctx, RT_ERR_INVALID, "image too large");
    if (!ctx->scratch) return fail(ctx, RT_ERR_UNSUPPORTED, "no scratch array on this device");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    if (!is_device_ptr(out)) return fail(ctx, RT_ERR_INVALID, "rt_render_wavefront_noalloc: out must be device memory");
    uint32_t npix = p->width * p->height;
    uint64_t total = (uint64_t)npix * spp;
    rt_wavefront_opts oo = opts ? *opts : rt_wavefront_opts{0, 0, 1, 0, 0, 1, 0, 0};
    if (oo.capacity == 0) oo.capacity = RT_NOALLOC_CAPACITY;
    if (oo.capacity > RT_NOALLOC_CAPACITY)
        return fail(ctx, RT_ERR_INVALID, "capacity " + std::to_string(oo.capacity) + " exceeds the no-allocation scratch (" +
                                             std::to_string(RT_NOALLOC_CAPACITY) + " paths)");
    if (oo.streams > 1 || oo.use_graph) return fail(ctx, RT_ERR_INVALID, "rt_render_wavefront_noalloc runs one stream without graphs");
    WfOpts o;
    int rc;
    if ((rc = wf_opts(ctx, &oo, total, o))) return rc;
    WfQueues Q;
    wf_layout((char*)ctx->scratch, o.cap, Q);
    Q.sort = o.sort;
    Q.knn_wave_max = o.knn_wave_max;
    Q.fb = out;
    DCamera C; DParams P;
    to_dev(cam, p, C, P);
    P.counters = nullptr;
    Q.h4 = P.wl == 2;
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min((npix * 3 + BLOCK - 1) / BLOCK, 4096u)), dim3(BLOCK), 0, ctx->stream,
                       (uint32_t*)out, (uint64_t)npix * 3, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(BLOCK), 0, ctx->stream, Q.counters, (uint64_t)WF_CNT_WORDS, 0u);
    HIPCHECK(ctx, hipGetLastError());
    WfBatch none = {nullptr, nullptr, 0};
    ctx->no_alloc = true;
    rc = wf_run<false>(ctx, s->S, C, P, &Q, 1, o, total, P.max_depth, none, P.use_photons && s->S.ph_knn);
    ctx->no_alloc = false;
    if (rc) return rc;
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    return RT_OK;
}

// rt_render_progressive through the wavefront: chunks of K whole passes
// (K = capacity / pixels, 1 when adaptive) render their per-sample radiance
// into a buffer of the chunk's lane; k_accum_passes then folds the passes
// into the pixel state in order.  Chunks go round-robin over the streams;
// each chunk's bookkeeping waits for the previous chunk's (an event per
// lane), its rendering does not.
extern "C" int rt_render_progressive_wavefront(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, const rt_sample_params* p,
                                               uint32_t samples_max, const rt_accum_opts* ao, const rt_pixel_state* state,
                                               const rt_wavefront_opts* opts, uint32_t* samples_done, uint32_t* n_converged) {
    if (!ctx || !s || !cam || !p) return RT_ERR_INVALID;
    if (wl_mode(p) < 0) return fail(ctx, RT_ERR_INVALID, "wavelengths must be 0, 1 or 4");
    int rc;
    if ((rc = check_accum_opts(ctx, ao))) return rc;
    if (p->width == 0 || p->height == 0) return fail(ctx, RT_ERR_INVALID, "empty image");
    if (p->max_depth == 0) return fail(ctx, RT_ERR_INVALID, "max_depth = 0");
    if ((uint64_t)p->width * p->height > 0xffffffffull / 12) return fail(ctx, RT_ERR_INVALID, "image too large");
    if (p->use_photons && s->S.num_photons == 0) return fail(ctx, RT_ERR_INVALID, "use_photons without a photon map");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    uint32_t npix = p->width * p->height;
    StateDev d;
    if ((rc = state_stage(ctx, state, npix, d))) return rc;
    if ((rc = ctl_buffer(ctx))) return rc;
    WfOpts o;
    if ((rc = wf_opts(ctx, opts, (uint64_t)npix * std::max(samples_max, 1u), o))) return rc;
    uint32_t K = ao->adaptive ? 1u : std::max(1u, o.cap / npix);
    if (o.cap < npix) o.cap = (npix + BLOCK - 1) / BLOCK * BLOCK;   // a chunk holds at least one whole pass
    o.cap = std::max(o.cap, K * npix);
    WfQueues Qs[RT_WF_MAX_STREAMS];
    if ((rc = wf_buffers(ctx, o.cap, Qs[0]))) return rc;
    Qs[0].sort = o.sort;
    Qs[0].knn_wave_max = o.knn_wave_max;
    Qs[0].h4 = wl_mode(p) == 2;
    if ((rc = wf_lanes(ctx, o, Qs))) return rc;
    const uint32_t nchunks = (samples_max + K - 1) / K;
    const uint32_t L = std::max(1u, std::min(o.streams, nchunks));
    size_t b_l = ((size_t)K * npix * 12 + 255) & ~(size_t)255;
    char* side = nullptr;
    HIPCHECK(ctx, hipMalloc(&side, std::max<size_t>(L * b_l, 256)));
    std::unique_ptr<void, void (*)(void*)> guard(side, [](void* q) { (void)hipFree(q); });
    std::vector<hipEvent_t> done_ev(L, nullptr);
    for (uint32_t k = 0; k < L; k++) HIPCHECK(ctx, hipEventCreateWithFlags(&done_ev[k], hipEventDisableTiming));
    struct EvGuard { std::vector<hipEvent_t>& v; ~EvGuard() { for (hipEvent_t e : v) if (e) (void)hipEventDestroy(e); } } eg{done_ev};
    DCamera C; DParams P;
    to_dev(cam, p, C, P);
    P.counters = nullptr;
    const bool knn = p->use_photons && s->S.ph_knn;
    auto stream_of = [&](uint32_t k) { return k ? ctx->xstream[k - 1] : ctx->stream; };
    float* acc = (float*)d.acc.dev; uint32_t* ns = (uint32_t*)d.ns.dev; float* m2 = (float*)d.m2.dev; uint8_t* conv = (uint8_t*)d.conv.dev;
    HIPCHECK(ctx, hipMemsetAsync(ctx->prog_ctl, 0, 4 * sizeof(uint32_t), ctx->stream));
    hipLaunchKernelGGL(k_count_flags, dim3(std::min((npix + BLOCK - 1) / BLOCK, RT_COUNT_GRID)), dim3(BLOCK), 0, ctx->stream, npix,
                       (const uint8_t*)conv, ctx->prog_ctl + 1);
    if (ctx->timing) HIPCHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (L > 1) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
        for (uint32_t k = 1; k < L; k++) HIPCHECK(ctx, hipStreamWaitEvent(stream_of(k), ctx->ev_fork, 0));
    }
    WfBatch none = {nullptr, nullptr, 0};
    uint32_t grid_acc = std::min((npix + BLOCK - 1) / BLOCK, 4096u);
    for (uint32_t c = 0; c < nchunks; c++) {
        const uint32_t k = c % L;
        hipStream_t st = stream_of(k);
        uint32_t pass0 = c * K, kk = std::min(K, samples_max - pass0);
        float* lsamp = (float*)(side + k * b_l);
        HIPCHECK(ctx, hipMemsetAsync(lsamp, 0, (size_t)kk * npix * 12, st));
        WfQueues Qc = Qs[k];
        Qc.fb = nullptr;
        Qc.lsamp = lsamp;
        Qc.stop = ctx->prog_ctl;
        Qc.conv = conv;                   // converged pixels: no path (their samples are skipped anyway)
        DParams Pc = P;
        Pc.sample_idx = P.sample_idx + pass0;
        wf_chunk<false>(ctx, st, s->S, C, Pc, Qc, 0, kk * npix, o.grid, o.occ, o.block, Pc.max_depth, none, knn);
        if (c > 0 && L > 1) HIPCHECK(ctx, hipStreamWaitEvent(st, done_ev[(c - 1) % L], 0));   // passes in order
        hipLaunchKernelGGL(k_accum_passes, dim3(grid_acc), dim3(BLOCK), 0, st, (const float*)lsamp, kk, p->width, p->height,
                           P.sample_idx + pass0, *ao, acc, ns, m2, conv, ctx->prog_ctl);
        for (uint32_t q = 0; q < kk; q++)
            hipLaunchKernelGGL(k_progressive_step, dim3(1), dim3(64), 0, st, ctx->prog_ctl, pass0 + q, npix, *ao);
        HIPCHECK(ctx, hipEventRecord(done_ev[k], st));
        HIPCHECK(ctx, hipGetLastError());
    }
    for (uint32_t k = 1; k < L; k++) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev_join[k - 1], stream_of(k)));
        HIPCHECK(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_join[k - 1], 0));
    }
    ctx->ext_used = 0;
    ctx->ext_knn.clear();
    if (ctx->timing) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHECK(ctx, hipEventSynchronize(ctx->ev1));
        float ms = 0.0f;
        HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->kernel_ms += ms;
        ctx->launches += nchunks;
    }
    uint32_t ctl[4] = {0, 0, 0, 0};
    HIPCHECK(ctx, hipMemcpyAsync(ctl, ctx->prog_ctl, sizeof ctl, hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = state_finish(ctx, state, d))) return rc;
    if (samples_done) *samples_done = ctl[2];
    if (n_converged) *n_converged = ctl[1];
    return RT_OK;
}

extern "C" int rt_render_batch(rt_ctx* ctx, const rt_batch_view* views, uint32_t nviews, const rt_wavefront_opts* opts) {
    if (!ctx || !views || nviews == 0) return RT_ERR_INVALID;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    std::vector<DView> dv(nviews);
    std::vector<DScene> ds(nviews);
    uint64_t total = 0, fb_pix = 0;
    uint32_t depth = 0;
    bool knn = false, h4 = false;
    for (uint32_t v = 0; v < nviews; v++) {
        const rt_batch_view& b = views[v];
        if (!b.scene || !b.accum_rgb || b.spp == 0) return fail(ctx, RT_ERR_INVALID, "batch view " + std::to_string(v) + ": missing scene/output/spp");
        if (b.scene->ctx != ctx) return fail(ctx, RT_ERR_INVALID, "batch view " + std::to_string(v) + ": scene of another context");
        const rt_sample_params& p = b.params;
        if (p.width == 0 || p.height == 0 || p.max_depth == 0) return fail(ctx, RT_ERR_INVALID, "batch view " + std::to_string(v) + ": empty frame");
        if (wl_mode(&p) < 0) return fail(ctx, RT_ERR_INVALID, "batch view " + std::to_string(v) + ": wavelengths must be 0, 1 or 4");
        h4 = h4 || wl_mode(&p) == 2;
        if (p.use_photons && b.scene->S.num_photons == 0) return fail(ctx, RT_ERR_INVALID, "batch view " + std::to_string(v) + ": no photon map");
        to_dev(&b.camera, &p, dv[v].cam, dv[v].P);
        dv[v].P.counters = nullptr;
        dv[v].npix = p.width * p.height;
        dv[v].first = total;
        dv[v].count = (uint64_t)dv[v].npix * b.spp;
        dv[v].fb_off = (uint32_t)fb_pix;
        total += dv[v].count;
        fb_pix += dv[v].npix;
        if (fb_pix > 0xffffffffull / 4) return fail(ctx, RT_ERR_INVALID, "batch frames too large");
        depth = std::max(depth, p.max_depth);
        ds[v] = b.scene->S;
        knn = knn || (p.use_photons && ds[v].ph_knn);
    }
    WfOpts o;
    int rc;
    if ((rc = wf_opts(ctx, opts, total, o))) return rc;
    WfQueues Q;
    if ((rc = wf_buffers(ctx, o.cap, Q))) return rc;
    Q.sort = o.sort;
    Q.knn_wave_max = o.knn_wave_max;
    Q.h4 = h4;
    // batch side data: views, scenes, concatenated accumulator
    size_t bv = (sizeof(DView) * nviews + 255) & ~(size_t)255, bs = (sizeof(DScene) * nviews + 255) & ~(size_t)255;
    size_t bf = (size_t)fb_pix * 12;
    char* side = nullptr;
    HIPCHECK(ctx, hipMalloc(&side, bv + bs + bf));
    std::unique_ptr<void, void (*)(void*)> guard(side, [](void* p) { (void)hipFree(p); });
    HIPCHECK(ctx, hipMemcpyAsync(side, dv.data(), sizeof(DView) * nviews, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemcpyAsync(side + bv, ds.data(), sizeof(DScene) * nviews, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemsetAsync(side + bv + bs, 0, bf, ctx->stream));
    Q.fb = (float*)(side + bv + bs);
    WfBatch Bt = {(const DScene*)(side + bv), (const DView*)side, nviews};
    DCamera C = dv[0].cam;
    DParams P = dv[0].P;
    WfQueues Qs[RT_WF_MAX_STREAMS];
    Qs[0] = Q;
    if ((rc = wf_lanes(ctx, o, Qs))) return rc;
    if ((rc = wf_run<true>(ctx, ds[0], C, P, Qs, o.streams, o, total, depth, Bt, knn))) return rc;
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (uint32_t v = 0; v < nviews; v++) {
        const float* src = Q.fb + 3 * (size_t)dv[v].fb_off;
        size_t bytes = (size_t)dv[v].npix * 12;
        hipMemcpyKind kind = is_device_ptr(views[v].accum_rgb) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        HIPCHECK(ctx, hipMemcpy(views[v].accum_rgb, src, bytes, kind));
    }
    return RT_OK;
}

// ---- differentiable material parameters ----------------------------------------------------
// Field -> parameter table (24 int8 per material, -1 = constant) after
// checking that every mapped field is differentiable.
// fmap[24 * material + field] = parameter or -1; mpar: one entry per medium
// object of each density parameter (de: the evaluation values, or the scene's).
static int param_map(rt_ctx* ctx, const rt_scene* s, const rt_param* prm, uint32_t n, std::vector<int8_t>& fmap,
                     std::vector<MedPar>* mpar = nullptr, const float* de = nullptr) {
    if (!prm || n == 0) return fail(ctx, RT_ERR_INVALID, "no parameters");
    if (n > RT_MAX_PARAMS) return fail(ctx, RT_ERR_INVALID, "at most " + std::to_string(RT_MAX_PARAMS) + " parameters");
    fmap.assign(24 * s->mats_host.size(), (int8_t)-1);
    if (mpar) mpar->clear();
    std::vector<int> seen(s->desc_media, -1);
    for (uint32_t i = 0; i < n; i++) {
        const rt_param& q = prm[i];
        std::string tag = "parameter " + std::to_string(i) + ": ";
        if (q.material & RT_PARAM_MEDIUM) {
            uint32_t k = q.material & ~RT_PARAM_MEDIUM;
            if (k >= s->desc_media) return fail(ctx, RT_ERR_INVALID, tag + "medium out of range");
            if (q.field != 0 || q.count != 1) return fail(ctx, RT_ERR_UNSUPPORTED, tag + "a medium has one parameter: field 0, count 1");
            if (seen[k] >= 0) return fail(ctx, RT_ERR_INVALID, tag + "medium already mapped");
            seen[k] = (int)i;
            int found = 0;
            for (size_t j = 0; j < s->media_host.size(); j++) {
                if (s->media_src[j] != k) continue;
                if (s->media_host[j].kind != RT_MEDIUM_CONSTANT) return fail(ctx, RT_ERR_UNSUPPORTED, tag + "only CONSTANT media are differentiable");
                if (s->media_xf[j]) return fail(ctx, RT_ERR_UNSUPPORTED, tag + "the medium's boundary carries a transform");
                found++;
                if (mpar) {
                    float d0 = s->media_host[j].density;
                    mpar->push_back(MedPar{s->media_slot[j], (int32_t)i, d0, de ? de[i] : d0});
                }
            }
            if (!found) return fail(ctx, RT_ERR_INVALID, tag + "no object uses that medium");
            continue;
        }
        if (q.material >= s->mats_host.size()) return fail(ctx, RT_ERR_INVALID, tag + "material out of range");
        if (q.count == 0 || q.field >= 24 || q.count > 24 - q.field) return fail(ctx, RT_ERR_INVALID, tag + "field range outside p[0..24)");
        const DMaterial& m = s->mats_host[q.material];
        for (uint32_t f = q.field; f < q.field + q.count; f++) {
            bool ok = (m.kind == RT_MAT_LAMBERTIAN && m.tex == 0 && f < 3) ||
                      (m.kind == RT_MAT_ROUGH_SPECTRAL && (f <= 3 || f >= 7));
            if (!ok) return fail(ctx, RT_ERR_UNSUPPORTED, tag + "field " + std::to_string(f) + " of material kind " +
                                                         std::to_string(m.kind) + " is not differentiable");
            int8_t& slot = fmap[24 * (size_t)q.material + f];
            if (slot >= 0) return fail(ctx, RT_ERR_INVALID, tag + "field " + std::to_string(f) + " already mapped");
            slot = (int8_t)i;
        }
    }
    return RT_OK;
}

static int render_args(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, const rt_sample_params* p, uint32_t spp) {
    if (!cam || !p || spp == 0) return fail(ctx, RT_ERR_INVALID, "missing camera/params or spp = 0");
    if (p->width == 0 || p->height == 0 || p->max_depth == 0) return fail(ctx, RT_ERR_INVALID, "empty frame");
    if ((uint64_t)p->width * p->height > 0xffffffffull / 4) return fail(ctx, RT_ERR_INVALID, "image too large");
    if (p->use_photons && s->S.num_photons == 0) return fail(ctx, RT_ERR_INVALID, "use_photons without a photon map");
    if (p->use_photons && s->S.ph_knn) return fail(ctx, RT_ERR_UNSUPPORTED, "the adjoint gathers photons in a fixed radius only");
    if (wl_mode(p) < 0) return fail(ctx, RT_ERR_INVALID, "wavelengths must be 0, 1 or 4");
    if (wl_mode(p) == 2) return fail(ctx, RT_ERR_UNSUPPORTED, "the adjoint differentiates one wavelength per path (wavelengths 0 or 1)");
    return RT_OK;
}

extern "C" int rt_scene_set_params(rt_ctx* ctx, rt_scene* s, const rt_param* prm, uint32_t n, const float* values) {
    if (!ctx || !s || !values) return RT_ERR_INVALID;
    std::vector<int8_t> fmap;
    int rc;
    if ((rc = param_map(ctx, s, prm, n, fmap))) return rc;
    for (uint32_t i = 0; i < n; i++)
        if (!std::isfinite(values[i])) return fail(ctx, RT_ERR_INVALID, "parameter " + std::to_string(i) + " is not finite");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    bool media = false;
    for (uint32_t i = 0; i < n; i++) {
        if (prm[i].material & RT_PARAM_MEDIUM) {
            if (!(values[i] > 0.0f)) return fail(ctx, RT_ERR_INVALID, "parameter " + std::to_string(i) + ": a density must be positive");
            for (size_t j = 0; j < s->media_host.size(); j++)
                if (s->media_src[j] == (prm[i].material & ~RT_PARAM_MEDIUM)) s->media_host[j].density = values[i];
            media = true;
            continue;
        }
        for (uint32_t f = prm[i].field; f < prm[i].field + prm[i].count; f++) s->mats_host[prm[i].material].p[f] = values[i];
    }
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHECK(ctx, hipMemcpy((void*)s->S.mats, s->mats_host.data(), sizeof(DMaterial) * s->mats_host.size(), hipMemcpyHostToDevice));
    if (media) HIPCHECK(ctx, hipMemcpy((void*)s->S.media, s->media_host.data(), sizeof(DMedium) * s->media_host.size(), hipMemcpyHostToDevice));
    return RT_OK;
}

extern "C" int rt_scene_get_params(rt_ctx* ctx, rt_scene* s, const rt_param* prm, uint32_t n, float* values) {
    if (!ctx || !s || !values) return RT_ERR_INVALID;
    std::vector<int8_t> fmap;
    int rc;
    if ((rc = param_map(ctx, s, prm, n, fmap))) return rc;
    for (uint32_t i = 0; i < n; i++) {
        if (prm[i].material & RT_PARAM_MEDIUM) {
            for (size_t j = 0; j < s->media_host.size(); j++)
                if (s->media_src[j] == (prm[i].material & ~RT_PARAM_MEDIUM)) { values[i] = s->media_host[j].density; break; }
            continue;
        }
        values[i] = s->mats_host[prm[i].material].p[prm[i].field];
    }
    return RT_OK;
}

extern "C" int rt_render_detached(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, const rt_sample_params* p, uint32_t spp,
                                  const rt_param* prm, uint32_t n, const float* eval_values, float* out) {
    if (!ctx || !s || !out || !eval_values) return RT_ERR_INVALID;
    int rc;
    if ((rc = render_args(ctx, s, cam, p, spp))) return rc;
    std::vector<int8_t> fmap;
    std::vector<MedPar> mpar;
    if ((rc = param_map(ctx, s, prm, n, fmap, &mpar, eval_values))) return rc;
    for (const MedPar& m : mpar) if (!(m.de > 0.0f)) return fail(ctx, RT_ERR_INVALID, "a density must be positive");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    std::vector<DMaterial> me = s->mats_host;
    for (uint32_t i = 0; i < n; i++) {
        if (prm[i].material & RT_PARAM_MEDIUM) continue;
        for (uint32_t f = prm[i].field; f < prm[i].field + prm[i].count; f++) me[prm[i].material].p[f] = eval_values[i];
    }
    void* dme = nullptr;
    size_t bme = (sizeof(DMaterial) * me.size() + 255) & ~(size_t)255;
    HIPCHECK(ctx, hipMalloc(&dme, bme + sizeof(MedPar) * (mpar.size() + 1)));
    std::unique_ptr<void, void (*)(void*)> guard(dme, [](void* q) { (void)hipFree(q); });
    HIPCHECK(ctx, hipMemcpy(dme, me.data(), sizeof(DMaterial) * me.size(), hipMemcpyHostToDevice));
    if (!mpar.empty()) HIPCHECK(ctx, hipMemcpy((char*)dme + bme, mpar.data(), sizeof(MedPar) * mpar.size(), hipMemcpyHostToDevice));
    const MedPar* dmp = (const MedPar*)((char*)dme + bme);
    uint32_t npix = p->width * p->height;
    Staged rout;
    if ((rc = stage_out(ctx, out, (size_t)npix * 12, rout, false))) return rc;
    DCamera C; DParams P;
    to_dev(cam, p, C, P);
    P.counters = nullptr;
    HIPCHECK(ctx, hipMemsetAsync(rout.dev, 0, (size_t)npix * 12, ctx->stream));
    if (ctx->timing) HIPCHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    const uint32_t chunk = 4;
    for (uint32_t k = 0; k < spp; k += chunk) {
        DParams Q = P;
        Q.sample_idx = P.sample_idx + k;
        hipLaunchKernelGGL(k_detached, dim3((npix + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, s->S, C, Q,
                           (const DMaterial*)dme, dmp, (int)mpar.size(), (float*)rout.dev, std::min(chunk, spp - k));
        HIPCHECK(ctx, hipGetLastError());
    }
    if (ctx->timing) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHECK(ctx, hipEventSynchronize(ctx->ev1));
        float ms = 0.0f;
        HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->kernel_ms += ms;
        ctx->launches += (spp + chunk - 1) / chunk;
    }
    return finish_out(ctx, out, (size_t)npix * 12, rout);
}

extern "C" int rt_backward(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, const rt_sample_params* p, uint32_t spp,
                           const rt_param* prm, uint32_t n, const float* dimg, float* grad) {
    if (!ctx || !s || !dimg || !grad) return RT_ERR_INVALID;
    int rc;
    if ((rc = render_args(ctx, s, cam, p, spp))) return rc;
    std::vector<int8_t> fmap;
    std::vector<MedPar> mpar;
    if ((rc = param_map(ctx, s, prm, n, fmap, &mpar))) return rc;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    uint32_t npix = p->width * p->height;
    Staged din, gout;
    if ((rc = stage_in(ctx, dimg, (size_t)npix * 12, din))) return rc;
    if ((rc = stage_out(ctx, grad, (size_t)n * 4, gout, false))) return rc;
    void* dmap = nullptr;
    size_t bmap = (fmap.size() + 255) & ~(size_t)255;
    HIPCHECK(ctx, hipMalloc(&dmap, bmap + sizeof(MedPar) * (mpar.size() + 1)));
    std::unique_ptr<void, void (*)(void*)> guard(dmap, [](void* q) { (void)hipFree(q); });
    HIPCHECK(ctx, hipMemcpy(dmap, fmap.data(), fmap.size(), hipMemcpyHostToDevice));
    if (!mpar.empty()) HIPCHECK(ctx, hipMemcpy((char*)dmap + bmap, mpar.data(), sizeof(MedPar) * mpar.size(), hipMemcpyHostToDevice));
    const MedPar* dmp = (const MedPar*)((char*)dmap + bmap);
    HIPCHECK(ctx, hipMemsetAsync(gout.dev, 0, (size_t)n * 4, ctx->stream));
    DCamera C; DParams P;
    to_dev(cam, p, C, P);
    P.counters = nullptr;
    if (ctx->timing) HIPCHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    const uint32_t chunk = 2;
    for (uint32_t k = 0; k < spp; k += chunk) {
        DParams Q = P;
        Q.sample_idx = P.sample_idx + k;
        hipLaunchKernelGGL(k_backward, dim3((npix + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, s->S, C, Q,
                           (const int8_t*)dmap, dmp, (int)mpar.size(), n, (const float*)din.dev, 1.0f / (float)spp,
                           std::min(chunk, spp - k), (float*)gout.dev);
        HIPCHECK(ctx, hipGetLastError());
    }
    if (ctx->timing) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHECK(ctx, hipEventSynchronize(ctx->ev1));
        float ms = 0.0f;
        HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->kernel_ms += ms;
        ctx->launches += (spp + chunk - 1) / chunk;
    }
    return finish_out(ctx, grad, (size_t)n * 4, gout);
}

extern "C" int rt_backward_wavefront(rt_ctx* ctx, rt_scene* s, const rt_camera* cam, const rt_sample_params* p, uint32_t spp,
                                     const rt_param* prm, uint32_t n, const float* dimg, float* grad,
                                     const rt_wavefront_opts* opts) {
    if (!ctx || !s || !dimg || !grad) return RT_ERR_INVALID;
    int rc;
    if ((rc = render_args(ctx, s, cam, p, spp))) return rc;
    std::vector<int8_t> fmap;
    std::vector<MedPar> mpar;
    if ((rc = param_map(ctx, s, prm, n, fmap, &mpar))) return rc;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    uint32_t npix = p->width * p->height;
    uint64_t total = (uint64_t)npix * spp;
    WfOpts o;
    if ((rc = wf_opts(ctx, opts, total, o))) return rc;
    WfQueues Q;
    if ((rc = wf_buffers(ctx, o.cap, Q))) return rc;
    Q.sort = o.sort;
    Q.knn_wave_max = o.knn_wave_max;
    Staged din, gout;
    if ((rc = stage_in(ctx, dimg, (size_t)npix * 12, din))) return rc;
    if ((rc = stage_out(ctx, grad, (size_t)n * 4, gout, false))) return rc;
    WfQueues Qs[RT_WF_MAX_STREAMS];
    Qs[0] = Q;
    if ((rc = wf_lanes(ctx, o, Qs))) return rc;
    const uint64_t nchunks = (total + o.cap - 1) / o.cap;
    const uint32_t L = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(o.streams, nchunks));
    // side buffers: per lane the per-slot adjoint state and per-sample primal
    // radiance; shared: the field map and the medium parameters
    size_t b_adj = (size_t)o.cap * sizeof(WfAdj), b_l = ((size_t)o.cap * 12 + 255) & ~(size_t)255;
    size_t b_lane = b_adj + b_l;
    size_t b_map = (fmap.size() + 255) & ~(size_t)255;
    char* side = nullptr;
    HIPCHECK(ctx, hipMalloc(&side, L * b_lane + b_map + sizeof(MedPar) * (mpar.size() + 1)));
    std::unique_ptr<void, void (*)(void*)> guard(side, [](void* q) { (void)hipFree(q); });
    char* shared = side + L * b_lane;
    HIPCHECK(ctx, hipMemcpy(shared, fmap.data(), fmap.size(), hipMemcpyHostToDevice));
    if (!mpar.empty()) HIPCHECK(ctx, hipMemcpy(shared + b_map, mpar.data(), sizeof(MedPar) * mpar.size(), hipMemcpyHostToDevice));
    WfAdjQ As[RT_WF_MAX_STREAMS];
    for (uint32_t k = 0; k < L; k++) {
        WfAdjQ& A = As[k];
        A.adj = (WfAdj*)(side + k * b_lane);
        A.lsamp = (const float*)(side + k * b_lane + b_adj);
        A.fmap = (const int8_t*)shared;
        A.mp = (const MedPar*)(shared + b_map);
        A.nmp = (int)mpar.size();
        A.dimg = (const float*)din.dev;
        A.inv_spp = 1.0f / (float)spp;
        A.nparams = n;
        A.grad = (float*)gout.dev;
    }
    HIPCHECK(ctx, hipMemsetAsync(gout.dev, 0, (size_t)n * 4, ctx->stream));
    DCamera C; DParams P;
    to_dev(cam, p, C, P);
    P.counters = nullptr;
    auto stream_of = [&](uint32_t k) { return k ? ctx->xstream[k - 1] : ctx->stream; };
    WfBatch none = {nullptr, nullptr, 0};
    if (ctx->timing) HIPCHECK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (L > 1) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
        for (uint32_t k = 1; k < L; k++) HIPCHECK(ctx, hipStreamWaitEvent(stream_of(k), ctx->ev_fork, 0));
    }
    uint32_t depth = P.max_depth;
    uint64_t chunks = 0;
    for (uint64_t first = 0; first < total; first += o.cap, chunks++) {
        const uint32_t k = (uint32_t)(chunks % L);
        hipStream_t st = stream_of(k);
        const WfQueues& Qk = Qs[k];
        const WfAdjQ& A = As[k];
        uint32_t m = (uint32_t)std::min<uint64_t>(o.cap, total - first);
        // primal: radiance of every sample of the chunk
        HIPCHECK(ctx, hipMemsetAsync((void*)A.lsamp, 0, (size_t)m * 12, st));
        WfQueues Qp = Qk;
        Qp.lsamp = (float*)A.lsamp;
        wf_chunk<false>(ctx, st, s->S, C, P, Qp, first, m, o.grid, o.occ, o.block, depth, none, false);   // no k-NN maps here (render_args)
        // replay from the same streams
        hipLaunchKernelGGL((k_wf_generate<false>), dim3((m + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, C, P, Qk, first, m, none);
        hipLaunchKernelGGL(k_wfb_init, dim3((m + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, Qk, A, m);
        hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(32), 0, st, Qk.counters, m, (const uint32_t*)nullptr, 0);
        WfQueues QA = Qk, QB = Qk;
        QB.active = Qk.next; QB.next = Qk.active;
        for (uint32_t d = 0; d < depth; d++) {
            const WfQueues& Qd = (d & 1) ? QB : QA;
            launch_trace<false>(false, o.occ, o.block, o.grid, st, s->S, Qd, none);
            hipLaunchKernelGGL(k_wfb_shade, dim3(o.grid), dim3(BLOCK), 0, st, s->S, P, Qd, A, (int)d);
            hipLaunchKernelGGL(k_wfb_shadow, dim3(o.grid), dim3(BLOCK), 0, st, s->S, Qd, A);
            hipLaunchKernelGGL(k_wf_advance, dim3(1), dim3(32), 0, st, Qd.counters);
        }
        HIPCHECK(ctx, hipGetLastError());
    }
    for (uint32_t k = 1; k < L; k++) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev_join[k - 1], stream_of(k)));
        HIPCHECK(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_join[k - 1], 0));
    }
    ctx->ext_used = 0;
    ctx->ext_knn.clear();
    if (ctx->timing) {
        HIPCHECK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHECK(ctx, hipEventSynchronize(ctx->ev1));
        float ms = 0.0f;
        HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->kernel_ms += ms;
        ctx->launches += chunks;
    }
    return finish_out(ctx, grad, (size_t)n * 4, gout);
}

extern "C" {
// ---- photon map --------------------------------------------------------------------------
// Left-balanced kd-tree in heap order (node i has children 2i+1, 2i+2, nodes
// 0..n-1): split on the widest axis of the subtree's bounds at the index that
// keeps the tree complete.  The gather (photon_irradiance) walks it without a
// stack.  The reference's tree (src/photon.rs:165-191) splits at n/2 of each
// range; both are exact range searches over the same photons.
struct PhRec { float4 a, b, c; };
static uint32_t lb_left_size(uint32_t n) {
    if (n <= 1) return 0;
    uint32_t m = 1;
    while (m <= n / 2) m *= 2;                 // largest power of two <= n
    uint32_t last = n - (m - 1);               // nodes on the partial last level
    return (m / 2 - 1) + std::min(last, m / 2);
}
static void heap_build(PhRec* src, uint32_t n, uint32_t node, float4* pos, float4* pw, float4* nr) {
    if (n == 0) return;
    float mn[3] = {FINF, FINF, FINF}, mx[3] = {-FINF, -FINF, -FINF};
    for (uint32_t i = 0; i < n; i++) {
        float p[3] = {src[i].a.x, src[i].a.y, src[i].a.z};
        for (int k = 0; k < 3; k++) { if (p[k] < mn[k]) mn[k] = p[k]; if (p[k] > mx[k]) mx[k] = p[k]; }
    }
    int axis = 0;
    for (int a = 1; a < 3; a++) if (!((mx[a] - mn[a]) < (mx[axis] - mn[axis]))) axis = a;
    uint32_t L = lb_left_size(n);
    auto key = [axis](const PhRec& p) { return axis == 0 ? p.a.x : (axis == 1 ? p.a.y : p.a.z); };
    std::nth_element(src, src + L, src + n, [&](const PhRec& x, const PhRec& y) { return key(x) < key(y); });
    pos[node] = make_float4(src[L].a.x, src[L].a.y, src[L].a.z, (float)axis);
    pw[node] = src[L].b;
    nr[node] = src[L].c;
    heap_build(src, L, 2 * node + 1, pos, pw, nr);
    heap_build(src + L + 1, n - L - 1, 2 * node + 2, pos, pw, nr);
}

// k-NN bucket tree: depth-D complete tree of planes over the photon range;
// each plane splits its range at the middle on the widest axis of the range's
// bounds (photons left of the middle have a coordinate <= the plane, the rest
// >=), and the 2^D leaves keep at most RT_PK_LEAF photons each, in place.
static void pk_build(PhRec* src, uint32_t n, uint32_t base, uint32_t node, uint32_t depth, uint32_t D,
                     float2* split, uint32_t* leaf_off) {
    if (depth == D) { leaf_off[node - ((1u << D) - 1u)] = base; return; }
    float mn[3] = {FINF, FINF, FINF}, mx[3] = {-FINF, -FINF, -FINF};
    for (uint32_t i = 0; i < n; i++) {
        float p[3] = {src[i].a.x, src[i].a.y, src[i].a.z};
        for (int k = 0; k < 3; k++) { if (p[k] < mn[k]) mn[k] = p[k]; if (p[k] > mx[k]) mx[k] = p[k]; }
    }
    int axis = 0;
    for (int a = 1; a < 3; a++) if (!((mx[a] - mn[a]) < (mx[axis] - mn[axis]))) axis = a;
    uint32_t mid = n / 2;
    auto key = [axis](const PhRec& p) { return axis == 0 ? p.a.x : (axis == 1 ? p.a.y : p.a.z); };
    float plane = 0.0f;
    if (n) {
        std::nth_element(src, src + mid, src + n, [&](const PhRec& x, const PhRec& y) { return key(x) < key(y); });
        plane = key(src[mid]);
    }
    int32_t ab = axis;
    float af;
    memcpy(&af, &ab, 4);
    split[node] = make_float2(plane, af);
    pk_build(src, mid, base, 2 * node + 1, depth + 1, D, split, leaf_off);
    pk_build(src + mid, n - mid, base + mid, 2 * node + 2, depth + 1, D, split, leaf_off);
}

// Heap-order the photons (position, power, normal per photon), lay the same
// photons out as the k-NN bucket tree, replace the scene's map on the device
// and keep the unordered set for export.
static int install_photons(rt_ctx* ctx, rt_scene* s, std::vector<float4>& kept, float r) {
    size_t m = kept.size() / 3;
    std::vector<PhRec> recs(m);
    for (size_t i = 0; i < m; i++) recs[i] = PhRec{kept[3 * i], kept[3 * i + 1], kept[3 * i + 2]};
    uint32_t D = 0;
    while ((size_t)RT_PK_LEAF << D < m) D++;
    if (D > 30) return fail(ctx, RT_ERR_UNSUPPORTED, "photon map too large");
    size_t nsplit = ((size_t)1 << D) - 1, nleaf = (size_t)1 << D;
    // one block: heap-order pos/pow/nrm (3m), bucket records (2m) and powers (m), planes, leaf offsets
    size_t f4 = 6 * m + 3;
    size_t bytes = f4 * sizeof(float4) + ((nsplit * sizeof(float2) + 15) & ~(size_t)15) + (nleaf + 1) * sizeof(uint32_t);
    std::vector<char> blob(bytes);
    float4* heap = (float4*)blob.data();
    {
        std::vector<PhRec> tmp(recs);
        heap_build(tmp.data(), (uint32_t)m, 0, heap, heap + m, heap + 2 * m);
    }
    float4* pk_rec = heap + 3 * m;
    float4* pk_pow = heap + 5 * m;
    float2* pk_split = (float2*)(blob.data() + f4 * sizeof(float4));
    uint32_t* pk_leaf = (uint32_t*)(blob.data() + f4 * sizeof(float4) + ((nsplit * sizeof(float2) + 15) & ~(size_t)15));
    pk_build(recs.data(), (uint32_t)m, 0, 0, 0, D, pk_split, pk_leaf);
    pk_leaf[nleaf] = (uint32_t)m;
    for (size_t i = 0; i < m; i++) {
        pk_rec[2 * i] = make_float4(recs[i].a.x, recs[i].a.y, recs[i].a.z, 0.0f);
        pk_rec[2 * i + 1] = make_float4(recs[i].c.x, recs[i].c.y, recs[i].c.z, 0.0f);
        pk_pow[i] = recs[i].b;
    }
    if (s->ph_mem) { (void)hipFree(s->ph_mem); s->ph_mem = nullptr; }
    s->S.num_photons = 0;
    HIPCHECK(ctx, hipMalloc(&s->ph_mem, bytes));
    HIPCHECK(ctx, hipMemcpy(s->ph_mem, blob.data(), bytes, hipMemcpyHostToDevice));
    const float4* dph = (const float4*)s->ph_mem;
    s->S.ph_pos = dph;
    s->S.ph_pow = dph + m;
    s->S.ph_nrm = dph + 2 * m;
    s->S.pk_rec = dph + 3 * m;
    s->S.pk_pow = dph + 5 * m;
    s->S.pk_split = (const float2*)((const char*)s->ph_mem + f4 * sizeof(float4));
    s->S.pk_leaf = (const uint32_t*)((const char*)s->ph_mem + f4 * sizeof(float4) + ((nsplit * sizeof(float2) + 15) & ~(size_t)15));
    s->S.pk_depth = D;
    s->S.num_photons = (uint32_t)m;
    s->S.ph_r2 = r * r;
    s->S.ph_norm = 2.0f / (PI * PI * r * r);
    s->S.ph_knn = 0;
    s->photons_host.swap(kept);
    return RT_OK;
}

int rt_scene_photon_bounds(rt_ctx* ctx, rt_scene* s, float* out6, int* has) {
    if (!ctx || !s || !out6 || !has) return RT_ERR_INVALID;
    *has = s->have_pb ? 1 : 0;
    for (int k = 0; k < 3; k++) { out6[k] = s->pb_min[k]; out6[3 + k] = s->pb_max[k]; }
    return RT_OK;
}

// RT_EMIT_SCENE -> the emitter SceneData::rebuild_caustics picks
// (src/scene.rs:35-74); the sky / env evaluation and the eight env draws run
// on the device (k_caustic_emitter).
int rt_photon_emitter(rt_ctx* ctx, rt_scene* s, const rt_photon_params* pp, rt_photon_params* out) {
    if (!ctx || !s || !pp || !out) return RT_ERR_INVALID;
    *out = *pp;
    if (pp->emitter == RT_EMIT_QUAD || pp->emitter == RT_EMIT_SUN_DISK) return RT_OK;
    if (pp->emitter != RT_EMIT_SCENE) return fail(ctx, RT_ERR_INVALID, "unknown photon emitter kind");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    float* d = nullptr;
    HIPCHECK(ctx, hipMalloc(&d, 8 * sizeof(float)));
    std::unique_ptr<void, void (*)(void*)> g(d, [](void* p) { (void)hipFree(p); });
    hipLaunchKernelGGL(k_caustic_emitter, dim3(1), dim3(64), 0, ctx->stream, s->S, d);
    HIPCHECK(ctx, hipGetLastError());
    float h[8];
    HIPCHECK(ctx, hipMemcpyAsync(h, d, 7 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (h[6] == 0.0f) {
        V3 u = mk(pp->u[0], pp->u[1], pp->u[2]), v = mk(pp->v[0], pp->v[1], pp->v[2]);
        if (!(len2(cross(u, v)) > 0.0f)) return fail(ctx, RT_ERR_INVALID, "no caustic emitter: no sun, no env map and no quad");
        out->emitter = RT_EMIT_QUAD;
        return RT_OK;
    }
    out->emitter = RT_EMIT_SUN_DISK;
    for (int k = 0; k < 3; k++) { out->sun_dir[k] = h[k]; out->color[k] = h[3 + k]; }
    return RT_OK;
}

int rt_photon_build(rt_ctx* ctx, rt_scene* s, const rt_photon_params* pp_in, uint32_t* num_stored) {
    if (!ctx || !s || !pp_in) return RT_ERR_INVALID;
    rt_photon_params res;
    int rc0 = rt_photon_emitter(ctx, s, pp_in, &res);
    if (rc0) return rc0;
    const rt_photon_params* pp = &res;
    if (pp->num_photons == 0) return fail(ctx, RT_ERR_INVALID, "num_photons = 0");
    if (pp->knn > RT_KNN_MAX) return fail(ctx, RT_ERR_UNSUPPORTED, "knn above RT_KNN_MAX");
    if (pp->wavelengths != 0 && pp->wavelengths != 1 && pp->wavelengths != 4) return fail(ctx, RT_ERR_INVALID, "wavelengths must be 0, 1 or 4");
    if (!(pp->gather_radius > 0.0f)) return fail(ctx, RT_ERR_INVALID, "gather_radius must be positive");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    DEmitter E;
    memset(&E, 0, sizeof E);
    E.wide = pp->wavelengths != 0 ? 1u : 0u;
    uint32_t n = pp->num_photons;
    if (pp->emitter == RT_EMIT_QUAD) {
        V3 lu = mk(pp->u[0], pp->u[1], pp->u[2]), lv = mk(pp->v[0], pp->v[1], pp->v[2]);
        V3 cr = cross(lu, lv);
        float area = std::sqrt(len2(cr));
        V3 nrm = cr / area;
        V3 up = std::fabs(nrm.y) < 0.9f ? mk(0, 1, 0) : mk(1, 0, 0);
        V3 cu = cross(nrm, up);
        V3 ta = cu / std::sqrt(len2(cu));
        V3 ba = cross(nrm, ta);
        E.kind = 0;
        for (int k = 0; k < 3; k++) { E.o[k] = pp->origin[k]; E.u[k] = pp->u[k]; E.v[k] = pp->v[k]; }
        E.n[0] = nrm.x; E.n[1] = nrm.y; E.n[2] = nrm.z; E.ta[0] = ta.x; E.ta[1] = ta.y; E.ta[2] = ta.z; E.ba[0] = ba.x; E.ba[1] = ba.y; E.ba[2] = ba.z;
        float f = area * PI / (float)n;
        for (int k = 0; k < 3; k++) E.power[k] = pp->color[k] * f;
        E.lum_clamp = 15.0f;
    } else {
        // PhotonMap::build (src/photon.rs:41-94): a disk facing the sun, sized
        // to the bounding sphere of the world's photon bounds and pushed out
        // along the sun direction; without finite geometry, the 20-unit disk
        // 30 units up the sun direction.
        V3 sd = mk(pp->sun_dir[0], pp->sun_dir[1], pp->sun_dir[2]);
        V3 sdu = sd / std::sqrt(len2(sd));
        V3 down = -sd;
        down = down / std::sqrt(len2(down));
        V3 up = std::fabs(down.x) < 0.999f ? mk(1, 0, 0) : mk(0, 1, 0);
        V3 tt = cross(down, up);
        tt = tt / std::sqrt(len2(tt));
        V3 bb = cross(down, tt);
        const float CLIP = 1.0e3f, DISK_R = 20.0f;
        auto clip = [&](float v) { return v < -CLIP ? -CLIP : (v > CLIP ? CLIP : v); };
        V3 dc;
        float r;
        if (s->have_pb) {
            V3 mn = mk(clip(s->pb_min[0]), clip(s->pb_min[1]), clip(s->pb_min[2]));
            V3 mx = mk(clip(s->pb_max[0]), clip(s->pb_max[1]), clip(s->pb_max[2]));
            V3 c = (mn + mx) * 0.5f;
            r = std::fmax(std::sqrt(len2(mx - mn)) * 0.5f, 1.0e-3f);
            dc = c + sdu * r;
        } else {
            dc = sdu * 30.0f;
            r = DISK_R;
        }
        float disk_area = PI * r * r;
        float f = disk_area / ((float)n * PI);
        E.kind = 1;
        for (int k = 0; k < 3; k++) E.power[k] = pp->color[k] * f;
        E.lum_clamp = 15.0f * disk_area / (PI * DISK_R * DISK_R);
        E.disk_c[0] = dc.x; E.disk_c[1] = dc.y; E.disk_c[2] = dc.z; E.disk_r = r;
        E.t[0] = tt.x; E.t[1] = tt.y; E.t[2] = tt.z; E.b[0] = bb.x; E.b[1] = bb.y; E.b[2] = bb.z;
        E.sun_down[0] = down.x; E.sun_down[1] = down.y; E.sun_down[2] = down.z;
    }
    float4* d_out = nullptr; uint32_t* d_ok = nullptr;
    HIPCHECK(ctx, hipMalloc(&d_out, (size_t)n * 3 * sizeof(float4)));
    std::unique_ptr<void, void (*)(void*)> g1(d_out, [](void* p) { (void)hipFree(p); });
    HIPCHECK(ctx, hipMalloc(&d_ok, (size_t)n * sizeof(uint32_t)));
    std::unique_ptr<void, void (*)(void*)> g2(d_ok, [](void* p) { (void)hipFree(p); });
    DScene S = s->S;
    S.num_photons = 0;
    S.pearl_sun_on = 0;       // photon tracing sees no pearl sun (src/material.rs:1052-1059)
    hipLaunchKernelGGL(k_photon_trace, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, S, E, n, d_out, d_ok);
    HIPCHECK(ctx, hipGetLastError());
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<float4> all((size_t)n * 3);
    std::vector<uint32_t> ok(n);
    HIPCHECK(ctx, hipMemcpy(all.data(), d_out, all.size() * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHECK(ctx, hipMemcpy(ok.data(), d_ok, ok.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::vector<float4> kept;
    for (uint32_t i = 0; i < n; i++) if (ok[i]) { kept.push_back(all[3 * (size_t)i]); kept.push_back(all[3 * (size_t)i + 1]); kept.push_back(all[3 * (size_t)i + 2]); }
    int rc;
    if ((rc = install_photons(ctx, s, kept, pp->gather_radius))) return rc;
    s->S.ph_knn = pp->knn;
    if (num_stored) *num_stored = (uint32_t)(s->photons_host.size() / 3);
    return RT_OK;
}

int rt_photon_set_gather(rt_ctx* ctx, rt_scene* s, float gather_radius, uint32_t knn) {
    if (!ctx || !s) return RT_ERR_INVALID;
    if (!(gather_radius > 0.0f)) return fail(ctx, RT_ERR_INVALID, "gather_radius must be positive");
    if (knn > RT_KNN_MAX) return fail(ctx, RT_ERR_UNSUPPORTED, "knn above RT_KNN_MAX");
    s->S.ph_r2 = gather_radius * gather_radius;
    s->S.ph_norm = 2.0f / (PI * PI * gather_radius * gather_radius);
    s->S.ph_knn = knn;
    return RT_OK;
}

int rt_photon_import(rt_ctx* ctx, rt_scene* s, const float* photons, uint32_t n, float gather_radius) {
    if (!ctx || !s || (!photons && n)) return RT_ERR_INVALID;
    if (!(gather_radius > 0.0f)) return fail(ctx, RT_ERR_INVALID, "gather_radius must be positive");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    std::vector<float4> kept((size_t)n * 3);
    for (size_t i = 0; i < n; i++) {
        const float* q = photons + 9 * i;
        for (int k = 0; k < 9; k++)
            if (!std::isfinite(q[k])) return fail(ctx, RT_ERR_INVALID, "photon " + std::to_string(i) + " is not finite");
        kept[3 * i] = make_float4(q[0], q[1], q[2], 0.0f);
        kept[3 * i + 1] = make_float4(q[3], q[4], q[5], 0.0f);
        kept[3 * i + 2] = make_float4(q[6], q[7], q[8], 0.0f);
    }
    return install_photons(ctx, s, kept, gather_radius);
}

int rt_photon_irradiance(rt_ctx* ctx, rt_scene* s, const float* pn, uint32_t n, float* irr) {
    if (!ctx || !s || (!pn && n) || (!irr && n)) return RT_ERR_INVALID;
    if (n == 0) return RT_OK;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    Staged pin, pout;
    int rc;
    if ((rc = stage_in(ctx, pn, (size_t)n * 24, pin))) return rc;
    if ((rc = stage_out(ctx, irr, (size_t)n * 12, pout, false))) return rc;
    if (s->S.ph_knn) hipLaunchKernelGGL(k_photon_irr<true>, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, s->S, (const float*)pin.dev, n, (float*)pout.dev);
    else hipLaunchKernelGGL(k_photon_irr<false>, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, s->S, (const float*)pin.dev, n, (float*)pout.dev);
    HIPCHECK(ctx, hipGetLastError());
    return finish_out(ctx, irr, (size_t)n * 12, pout);
}

int rt_photon_export(rt_ctx* ctx, rt_scene* s, float* out, uint32_t cap, uint32_t* count) {
    if (!ctx || !s || !count) return RT_ERR_INVALID;
    size_t m = s->photons_host.size() / 3;
    *count = (uint32_t)m;
    if (!out) return RT_OK;
    for (size_t i = 0; i < m && i < cap; i++) {
        const float4* p = &s->photons_host[3 * i];
        float* o = out + 9 * i;
        o[0] = p[0].x; o[1] = p[0].y; o[2] = p[0].z;
        o[3] = p[1].x; o[4] = p[1].y; o[5] = p[1].z;
        o[6] = p[2].x; o[7] = p[2].y; o[8] = p[2].z;
    }
    return RT_OK;
}

// ---- measurement ----------------------------------------------------------------------------
int rt_ctx_set_timing(rt_ctx* ctx, int enable) {
    if (!ctx) return RT_ERR_INVALID;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    if (enable && !ctx->ev0) {
        HIPCHECK(ctx, hipEventCreate(&ctx->ev0));
        HIPCHECK(ctx, hipEventCreate(&ctx->ev1));
    }
    ctx->timing = enable != 0;
    ctx->kernel_ms = 0.0;
    ctx->launches = 0;
    for (int k = 0; k < 4; k++) { ctx->stage_ms[k] = 0.0; ctx->stage_n[k] = 0; }
    return RT_OK;
}

int rt_ctx_get_stage_timing(rt_ctx* ctx, double* extend_ms, uint64_t* extend_launches) {
    if (!ctx || !extend_ms || !extend_launches) return RT_ERR_INVALID;
    *extend_ms = ctx->stage_ms[0];
    *extend_launches = ctx->stage_n[0];
    return RT_OK;
}
int rt_ctx_get_stage_times(rt_ctx* ctx, double ms[4], uint64_t launches[4]) {
    if (!ctx || !ms || !launches) return RT_ERR_INVALID;
    for (int k = 0; k < 4; k++) { ms[k] = ctx->stage_ms[k]; launches[k] = ctx->stage_n[k]; }
    return RT_OK;
}

int rt_ctx_get_timing(rt_ctx* ctx, double* kernel_ms, uint64_t* launches) {
    if (!ctx || !kernel_ms || !launches) return RT_ERR_INVALID;
    *kernel_ms = ctx->kernel_ms;
    *launches = ctx->launches;
    return RT_OK;
}

int rt_ctx_set_megakernel_occupancy(rt_ctx* ctx, uint32_t occupancy) {
    if (!ctx) return RT_ERR_INVALID;
    if (occupancy != 0 && occupancy != 2) return fail(ctx, RT_ERR_INVALID, "megakernel occupancy must be 0 or 2");
    ctx->mk_occ = (int)occupancy;
    return RT_OK;
}
int rt_ctx_set_counting(rt_ctx* ctx, int enable) {
    if (!ctx) return RT_ERR_INVALID;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    if (enable) {
        if (!ctx->counters) HIPCHECK(ctx, hipMalloc(&ctx->counters, 8 * sizeof(unsigned long long)));
        HIPCHECK(ctx, hipMemset(ctx->counters, 0, 8 * sizeof(unsigned long long)));
    } else if (ctx->counters) {
        HIPCHECK(ctx, hipFree(ctx->counters));
        ctx->counters = nullptr;
    }
    return RT_OK;
}

int rt_ctx_get_counters(rt_ctx* ctx, uint64_t out[8]) {
    if (!ctx || !out) return RT_ERR_INVALID;
    if (!ctx->counters) return fail(ctx, RT_ERR_INVALID, "counting is off");
    HIPCHECK(ctx, hipMemcpy(out, ctx->counters, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return RT_OK;
}

}  // extern "C"

