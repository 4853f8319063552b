// Device scene layout and the BVH / primitive queries of the radiance path.
//
// HBM layout (all arrays 256-B aligned by the allocator):
//   nodes   : one 256-B record per 8-wide BVH node, SoA inside the node —
//             minx[8] miny[8] minz[8] maxx[8] maxy[8] maxz[8] child[8] pad[8].
//             The world tree and every nested tree (mesh, box) share the
//             array; child >= 0 is a node index, child < 0 is leaf -child-1
//             (world: object slot; nested: primitive slot), INT32_MIN is empty.
//   objs    : world objects in leaf order (DObject.tag = input index).
//   tris    : per triangle 3 x float4 = v0, e1 = v1-v0, e2 = v2-v0 (48 B),
//             leaf order inside each mesh; shading data (normals, tangents,
//             uvs) in separate arrays, read only for the closest hit.
//   quads   : per quad q, u, v, normal, w, (d, area) (80 B).
//   planes  : convex half-spaces (n.xyz, d).
// Semantics follow rustracer's BvhTree (src/bvh.rs:229-575): the slab test
// replicates AVX2 minps/maxps (second operand on NaN) and the ordered <=
// compare, children are visited near-to-far from a descending sort of t_near,
// and every interval is closed.
#pragma once
#include "dev_math.h"

namespace rt {

enum { SH_SPHERE = 1, SH_QUAD = 2, SH_MESH = 3, SH_BOX = 4, SH_CONVEX = 5, SH_CYLINDER = 6, SH_DISK = 7, SH_CONE = 8,
       SH_PLANE = 9 };
enum { XF_TRANSLATE = 1, XF_ROTATE = 2, XF_SCALE = 3 };
enum { MED_NONE = 0, MED_CONSTANT = 1, MED_GRID = 2, MED_NOISE = 3 };

struct DXform { uint32_t kind; float v[9]; };

struct DObject {
    uint32_t shape, material, xf_first, xf_count;
    uint32_t data_first, data_count;   // MESH: tris / BOX: quads / QUAD: quad index / CONVEX: planes
    uint32_t root;                     // nested BVH root node (MESH, BOX)
    uint32_t medium;                   // 0 or media index + 1
    uint32_t tag;                      // input index
    uint32_t emissive;
    float light_area;
    float p[16];
};

struct DQuad { float q[3], u[3], v[3], n[3], w[3], d, area; float pad[3]; };
static_assert(sizeof(DQuad) == 80, "quad record is five 16-B words");

struct DMedium {
    uint32_t kind; float density; float albedo[3]; float g;
    uint32_t nx, ny, nz; uint32_t grid_off;     // GRID: dims and offset; NOISE: offset of the Perlin tables
    float bmin[3], bmax[3];
    float noise_scale, threshold;               // NOISE
};

// toff: RT_TEX_IMAGE byte offset of the image in DScene::tex_pix (tw x th
// texels), RT_TEX_NOISE / RT_TEX_MOTTLE float offset of the Perlin tables in
// DScene::grids.
// map_*: the PBR texture maps (albedo, roughness, metallic, ambient
// occlusion, normal) as byte offset / width / height in tex_pix, width 0 = none.
struct DMaterial { uint32_t kind, tex; float p[24]; uint32_t toff, tw, th, pad; uint32_t map_off[5], map_w[5], map_h[5], pad2; };

struct DScene {
    const float4* nodes;
    const DObject* objs; uint32_t num_objs;
    const DObject* lights; uint32_t num_lights;
    const DXform* xforms;
    const float4* tris; const int32_t* tri_tag;
    const float* tri_n; const float* tri_t; const float* tri_uv;
    const DQuad* quads;
    const float4* planes;
    const DMaterial* mats;
    const DMedium* media; const float* grids;
    const uint8_t* tex_pix; const float* srgb_lut;   // image texels (sRGB bytes) and their 256-entry decode table
    const float* nacre;                    // thin-film colour over optical path difference: 256 RGB entries
    float pearl_sun[3]; uint32_t pearl_sun_on;      // the renderer's pearl sun (a Physical background's)
    // background
    uint32_t bg_kind, bg_visible;
    float bg_color[3], bg_top[3], bg_bottom[3], sun_dir[3], turbidity;
    uint32_t env_w, env_h; float env_rot, env_inv_total;
    uint32_t env_sampling;                 // 0 row/column CDF, 1 alias table
    const float* env_aprob; const uint32_t* env_alias;   // alias table: acceptance probability, alias texel
    const float* env_rgb; const float* env_marg; const float* env_cond;
    const float* elut;
    float norms[3], fire_norms[3];
    float norms_w[3], fire_norms_w[3];     // the same over the wide table (380-780 nm)
    // photon map: left-balanced kd-tree in heap order, structure of arrays
    // (position + split axis in w, power, normal), 16 B per array per photon
    const float4* ph_pos; const float4* ph_pow; const float4* ph_nrm;
    uint32_t num_photons; float ph_r2, ph_norm;
    uint32_t ph_knn;                       // 0: fixed-radius gather; k: k-NN gather bounded by ph_r2
    // the same photons in the k-NN bucket tree: leaf order records {pos, 0}
    // {normal, 0} and powers, heap-order planes {value, axis bits}, leaf offsets
    const float4* pk_rec; const float4* pk_pow; const float2* pk_split; const uint32_t* pk_leaf;
    uint32_t pk_depth;
    uint32_t stack_bound;
};

// ---- per-lane traversal stack: LDS part + private overflow -------------------
#ifndef RT_LDS_STACK
#define RT_LDS_STACK 32
#endif
#define RT_PRIV_STACK 96
// Pointers into LDS: per-lane stack and heap accesses through them compile to
// ds_read / ds_write, not to flat accesses that wait on every memory counter.
typedef __attribute__((address_space(3))) int32_t lds_i32;
typedef __attribute__((address_space(3))) float lds_f32;
// The overflow stack through a private (address-space-5) pointer: stores to
// it cannot alias global memory, so loads of scene records stay hoisted out
// of the traversal loops.
typedef __attribute__((address_space(5))) int32_t priv_i32;
// The stack's fields are all scalars and pointers, so the compiler keeps the
// struct in registers; the overflow array is a separate private array of the
// kernel (dynamically indexed, hence in scratch), touched only past
// RT_LDS_STACK entries.
struct Stack {
    lds_i32* lds;      // base of this lane's column, stride = blockDim.x
    uint32_t stride;
    priv_i32* priv;    // this lane's overflow: RT_PRIV_STACK entries
    int32_t sp;
    lds_f32* heap;     // k-NN photon gather: this lane's LDS column (stride) of RT_KNN_MAX floats
    bool has_heap;     // heap is set (an LDS pointer may be 0, so its value cannot say)
    uint32_t n_nodes, n_prims, n_rays;   // traversal work, for the counting pass
    uint32_t s_nodes, s_prims, s_rays;   // the same for shadow (transmittance) queries
    uint32_t n_pbytes;                   // photon gathers: record bytes read
};
RT_D void st_push(Stack& s, int32_t v) {
    if (s.sp < RT_LDS_STACK) s.lds[s.sp * s.stride] = v;
    else s.priv[s.sp - RT_LDS_STACK] = v;
    s.sp++;
}
RT_D int32_t st_pop(Stack& s) {
    s.sp--;
    return s.sp < RT_LDS_STACK ? s.lds[s.sp * s.stride] : s.priv[s.sp - RT_LDS_STACK];
}

// ---- primitive tests ---------------------------------------------------------
RT_D bool sphere_test(const float* p, const Ray& r, float t_min, float t_max, float& t_out) {
    V3 c = mk(p[0], p[1], p[2]);
    if (p[7] != 0.0f) c = c + r.time * (mk(p[4], p[5], p[6]) - c);
    V3 oc = r.o - c;
    float a = len2(r.d);
    if (!(a > 0.0f)) return false;
    float half_b = dot(oc, r.d);
    float cc = len2(oc) - p[3] * p[3];
    float disc = half_b * half_b - a * cc;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float root = (-half_b - sq) / a;
    if (root < t_min || root > t_max) {
        root = (-half_b + sq) / a;
        if (root < t_min || root > t_max) return false;
    }
    t_out = root;
    return true;
}

RT_D bool quad_test(const DQuad& qm, const Ray& r, float t_min, float t_max, float& t_out, float& al, float& be) {
    // the 80-B record as five 16-B loads in flight together (quad arrays are
    // 16-B aligned: device allocations are, and the record is 5 x 16 B)
    const float4* p4 = reinterpret_cast<const float4*>(&qm);
    float4 x0 = p4[0], x1 = p4[1], x2 = p4[2], x3 = p4[3], x4 = p4[4];
    __builtin_amdgcn_sched_barrier(0);
    DQuad q;
    q.q[0] = x0.x; q.q[1] = x0.y; q.q[2] = x0.z; q.u[0] = x0.w;
    q.u[1] = x1.x; q.u[2] = x1.y; q.v[0] = x1.z; q.v[1] = x1.w;
    q.v[2] = x2.x; q.n[0] = x2.y; q.n[1] = x2.z; q.n[2] = x2.w;
    q.w[0] = x3.x; q.w[1] = x3.y; q.w[2] = x3.z; q.d = x3.w;
    q.area = x4.x;
    V3 n = mk(q.n[0], q.n[1], q.n[2]);
    float denom = dot(n, r.d);
    if (!(fabsf(denom) >= 1e-8f)) return false;
    float t = (q.d - dot(n, r.o)) / denom;
    if (t < t_min || t > t_max) return false;
    V3 P = at(r, t);
    V3 planar = P - mk(q.q[0], q.q[1], q.q[2]);
    V3 w = mk(q.w[0], q.w[1], q.w[2]);
    float alpha = dot(w, cross(planar, mk(q.v[0], q.v[1], q.v[2])));
    float beta = dot(w, cross(mk(q.u[0], q.u[1], q.u[2]), planar));
    if (!(alpha >= 0.0f && alpha <= 1.0f) || !(beta >= 0.0f && beta <= 1.0f)) return false;
    t_out = t; al = alpha; be = beta;
    return true;
}

// Moller-Trumbore (src/triangle.rs:75-126) on precomputed v0, e1, e2.
RT_D bool tri_test(const float4* tris, uint32_t i, const Ray& r, float t_min, float t_max, float& t_out, float& b1o, float& b2o) {
    float4 A = tris[3 * i], B = tris[3 * i + 1], Cc = tris[3 * i + 2];
    __builtin_amdgcn_sched_barrier(0);      // the record's three loads in flight together
    V3 v0 = mk(A.x, A.y, A.z), e1 = mk(B.x, B.y, B.z), e2 = mk(Cc.x, Cc.y, Cc.z);
    V3 h = cross(r.d, e2);
    float a = dot(e1, h);
    if (fabsf(a) < 1e-8f) return false;
    float f = 1.0f / a;
    V3 s = r.o - v0;
    float b1 = f * dot(s, h);
    if (!(b1 >= 0.0f && b1 <= 1.0f)) return false;
    V3 qv = cross(s, e1);
    float b2 = f * dot(r.d, qv);
    if (b2 < 0.0f || b1 + b2 > 1.0f) return false;
    float t = f * dot(e2, qv);
    if (t < t_min || t > t_max) return false;
    t_out = t; b1o = b1; b2o = b2;
    return true;
}

// convex_hit (src/diamond.rs:203-253); idx_out = plane index of the normal.
RT_D bool convex_test(const float4* pl, uint32_t first, uint32_t count, const Ray& r, float t_min, float t_max, float& t_out, int& idx_out) {
    float t_enter = t_min, t_exit = t_max;
    int ei = -1, xi = -1;
    bool entry = false;
    for (uint32_t k = 0; k < count; k++) {
        float4 P = pl[first + k];
        V3 n = mk(P.x, P.y, P.z);
        float nd = dot(n, r.d);
        float ndo = dot(n, r.o);
        float t_face = (P.w - ndo) / nd;
        if (fabsf(nd) < 1e-8f) {
            if (ndo > P.w) return false;
            continue;
        }
        if (nd < 0.0f) { if (t_face > t_enter) { t_enter = t_face; ei = (int)k; entry = true; } }
        else { if (t_face < t_exit) { t_exit = t_face; xi = (int)k; } }
        if (t_enter > t_exit) return false;
    }
    if (t_enter > t_exit) return false;
    if (t_enter == t_exit && t_enter < t_max) return false;
    if (entry) { t_out = t_enter; idx_out = ei; return true; }
    if (t_exit > t_max) return false;
    t_out = t_exit; idx_out = xi;
    return true;
}

// ---- transforms ---------------------------------------------------------------
RT_HD V3 rot_apply(const float* m, V3 v) {
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z, m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
RT_HD V3 rot_apply_t(const float* m, V3 v) {
    return mk(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z, m[2] * v.x + m[5] * v.y + m[8] * v.z);
}
// World ray -> object space: outermost wrapper first.
RT_D Ray to_object(const DScene& S, const DObject& ob, const Ray& r) {
    if (ob.xf_count == 0) return r;
    V3 o = r.o, d = r.d;
    for (int k = (int)ob.xf_count - 1; k >= 0; k--) {
        const DXform& x = S.xforms[ob.xf_first + k];
        if (x.kind == XF_TRANSLATE) o = o - mk(x.v[0], x.v[1], x.v[2]);
        else if (x.kind == XF_ROTATE) { o = rot_apply_t(x.v, o); d = rot_apply_t(x.v, d); }
        else { float inv = 1.0f / x.v[0]; o = inv * o; d = inv * d; }
    }
    Ray q = make_ray(o, d, r.time);
    q.wavelength = r.wavelength; q.spectral = r.spectral; q.wl = r.wl; q.tint = r.tint;
    return q;
}

// ---- node test (test_children_avx2, src/bvh.rs:330-364) ------------------------
RT_D float mnps(float a, float b) { return a < b ? a : b; }
RT_D float mxps(float a, float b) { return a > b ? a : b; }

// f32::total_cmp key
RT_D int32_t tkey(float f) { int32_t i = __float_as_int(f); return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1); }

// Returns the hit mask; key[c] = the total_cmp key of child c's t_near.
// Children read as ints.  The node's 14 float4 loads are all issued before
// the slab tests (a scheduling barrier keeps the compiler from spreading them
// over three dependent rounds to save registers), so a node costs one memory
// round trip.  Each lane loads an axis's planes near side first: the upper
// bounds when its ray's inverse direction component is negative.
RT_D uint32_t node_test(const float4* node, const Ray& r, float t_min, float t_max, int32_t key[8], int32_t ch[8]) {
    const bool neg[3] = {r.inv.x < 0.0f, r.inv.y < 0.0f, r.inv.z < 0.0f};
    float4 q[14];
#pragma unroll
    for (int a = 0; a < 3; a++) {                      // planes of axis a: float4 2a, 2a+1 lower, 6+2a, 7+2a upper
        const int s = neg[a] ? 6 : 0;
        q[2 * a] = node[2 * a + s]; q[2 * a + 1] = node[2 * a + 1 + s];
        q[6 + 2 * a] = node[6 + 2 * a - s]; q[7 + 2 * a] = node[7 + 2 * a - s];
    }
    q[12] = node[12]; q[13] = node[13];
    __builtin_amdgcn_sched_barrier(0);      // all 14 loads in flight before the slab tests
    float nr[3][8], fr[3][8];                 // near and far planes
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float4 n0 = q[2 * a], n1 = q[2 * a + 1], f0 = q[6 + 2 * a], f1 = q[7 + 2 * a];
        nr[a][0] = n0.x; nr[a][1] = n0.y; nr[a][2] = n0.z; nr[a][3] = n0.w;
        nr[a][4] = n1.x; nr[a][5] = n1.y; nr[a][6] = n1.z; nr[a][7] = n1.w;
        fr[a][0] = f0.x; fr[a][1] = f0.y; fr[a][2] = f0.z; fr[a][3] = f0.w;
        fr[a][4] = f1.x; fr[a][5] = f1.y; fr[a][6] = f1.z; fr[a][7] = f1.w;
    }
    float4 c0 = q[12], c1 = q[13];
    ch[0] = __float_as_int(c0.x); ch[1] = __float_as_int(c0.y); ch[2] = __float_as_int(c0.z); ch[3] = __float_as_int(c0.w);
    ch[4] = __float_as_int(c1.x); ch[5] = __float_as_int(c1.y); ch[6] = __float_as_int(c1.z); ch[7] = __float_as_int(c1.w);
    const float o[3] = {r.o.x, r.o.y, r.o.z}, id[3] = {r.inv.x, r.inv.y, r.inv.z};
    uint32_t mask = 0;
    // The slab bounds are NaN only when a direction component's inverse is
    // infinite (0 x inf).  With all three inverses finite, a child's box
    // lo <= hi (an empty slot's box is a point) and t_min > 0: the near
    // plane's distance is the smaller of the pair bit for bit (rounded
    // subtraction, and multiplication by a number of fixed sign, keep the
    // order), so the compare-selects of the reference reduce to the hardware
    // max / min (three-operand forms) over the axes' near and far distances:
    // t_near >= t_min > 0, so no signed zero reaches t_near or the mask, and
    // t_near's bits are its total_cmp key.  Otherwise the reference's
    // compare-selects on the (lower, upper) pairs, axis by axis.
    const bool fast = t_min > 0.0f && fabsf(r.inv.x) < FINF && fabsf(r.inv.y) < FINF && fabsf(r.inv.z) < FINF;
    if (fast) {
#pragma unroll
        for (int c = 0; c < 8; c++) {
            float dn[3], df[3];
#pragma unroll
            for (int a = 0; a < 3; a++) { dn[a] = (nr[a][c] - o[a]) * id[a]; df[a] = (fr[a][c] - o[a]) * id[a]; }
            float tn = fmaxf(fmaxf(fmaxf(t_min, dn[0]), dn[1]), dn[2]);
            float tf = fminf(fminf(fminf(t_max, df[0]), df[1]), df[2]);
            key[c] = __float_as_int(tn);
            if (tn <= tf && ch[c] != INT32_MIN) mask |= 1u << c;
        }
    } else {
#pragma unroll
        for (int c = 0; c < 8; c++) {
            float t0 = t_min, t1 = t_max;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float lo = neg[a] ? fr[a][c] : nr[a][c], hi = neg[a] ? nr[a][c] : fr[a][c];
                float da = (lo - o[a]) * id[a];
                float db = (hi - o[a]) * id[a];
                t0 = mxps(t0, mnps(da, db));
                t1 = mnps(t1, mxps(da, db));
            }
            key[c] = tkey(t0);
            if (t0 <= t1 && ch[c] != INT32_MIN) mask |= 1u << c;
        }
    }
    return mask;
}

// Push hit children far-to-near (stable for equal keys) so the nearest pops
// first.  Each child's slot is its rank in that order: the hit children with
// a larger key, or an equal key and a lower index, go below it.  One integer
// comparison per pair of slots decides both directions (for c < d, c goes
// below d exactly when key[c] >= key[d]); the sort itself indexes no private
// array dynamically (the stack's overflow past RT_LDS_STACK entries, a
// private array, is written only on deep stacks).
RT_D void push_sorted(Stack& st, uint32_t mask, const int32_t key[8], const int32_t ch[8]) {
    int32_t rank[8];
#pragma unroll
    for (int c = 0; c < 8; c++) rank[c] = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const bool hc = (mask >> c) & 1u;
#pragma unroll
        for (int d = c + 1; d < 8; d++) {
            const bool hd = (mask >> d) & 1u;
            const bool c_below = key[c] >= key[d];
            rank[d] += (hc && c_below) ? 1 : 0;
            rank[c] += (hd && !c_below) ? 1 : 0;
        }
    }
    const int32_t sp = st.sp;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        if (!(mask & (1u << c))) continue;
        int32_t at = sp + rank[c];
        if (at < RT_LDS_STACK) st.lds[at * st.stride] = ch[c];
        else st.priv[at - RT_LDS_STACK] = ch[c];
    }
    st.sp = sp + __popc(mask);
}

// ---- cylinder / cone / disk / infinite plane ----------------------------------------------
// Object-space inline parameters as uploaded: CYLINDER and CONE p0-2 base
// centre, p3 radius, p4 height; DISK p0-2 centre, p3-5 unit normal, p6 radius,
// p7-9 / p10-12 the tangent frame of Vec3::onb, p13 area; PLANE p0-2 point,
// p3-5 unit normal, p6 wave amplitude, p7 wave scale, p8-10 / p11-13 frame.
// Closed [t_min, t_max]; `kind` names the surface hit (0 lateral / disk /
// plane, 1 bottom cap or cone base, 2 top cap).
static constexpr float TAU_F = 6.28318530717958647692f;

// Cylinder::hit (src/cylinder.rs:133-185)
RT_D bool cylinder_test(const float* p, const Ray& r, float t_min, float t_max, float& t_out, int& kind) {
    V3 c = mk(p[0], p[1], p[2]);
    float rad = p[3], hgt = p[4];
    V3 oc = r.o - c;
    float dx = r.d.x, dz = r.d.z, dy = r.d.y;
    float best_t = t_max;
    kind = -1;
    float a = dx * dx + dz * dz;
    if (a > 1e-8f) {
        float half_b = oc.x * dx + oc.z * dz;
        float cc = oc.x * oc.x + oc.z * oc.z - rad * rad;
        float disc = half_b * half_b - a * cc;
        if (disc >= 0.0f) {
            float sq = sqrtf(disc);
            float roots[2] = {(-half_b - sq) / a, (-half_b + sq) / a};
            for (int k = 0; k < 2; k++) {
                float root = roots[k];
                if (root < t_min || root > best_t) continue;
                float ly = (r.o.y + r.d.y * root) - c.y;
                if (ly < 0.0f || ly > hgt) continue;
                best_t = root; kind = 0;
            }
        }
    }
    if (fabsf(dy) > 1e-8f) {
        for (int k = 0; k < 2; k++) {
            float cap_y = k == 0 ? 0.0f : hgt;
            float t = (cap_y - oc.y) / dy;
            if (t < t_min || t > best_t) continue;
            float lx = (r.o.x + r.d.x * t) - c.x, lz = (r.o.z + r.d.z * t) - c.z;
            if (lx * lx + lz * lz > rad * rad) continue;
            best_t = t; kind = 1 + k;
        }
    }
    t_out = best_t;
    return kind >= 0;
}

// Cone::hit (src/cone.rs:19-82): lateral x^2 + z^2 = k (h - y)^2, base disk at y = 0.
RT_D bool cone_test(const float* p, const Ray& r, float t_min, float t_max, float& t_out, int& kind) {
    V3 c = mk(p[0], p[1], p[2]);
    float rad = p[3], h = p[4];
    V3 oc = r.o - c;
    float k = (rad / h) * (rad / h);
    float dx = r.d.x, dy = r.d.y, dz = r.d.z;
    float hy = h - oc.y;
    float best_t = t_max;
    kind = -1;
    float a = dx * dx + dz * dz - k * dy * dy;
    float half_b = oc.x * dx + oc.z * dz + k * hy * dy;
    float cc = oc.x * oc.x + oc.z * oc.z - k * hy * hy;
    if (fabsf(a) > 1e-8f) {
        float disc = half_b * half_b - a * cc;
        if (disc >= 0.0f) {
            float sq = sqrtf(disc);
            float roots[2] = {(-half_b - sq) / a, (-half_b + sq) / a};
            for (int j = 0; j < 2; j++) {
                float root = roots[j];
                if (root < t_min || root > best_t) continue;
                float ly = (r.o.y + r.d.y * root) - c.y;
                if (ly < 0.0f || ly > h) continue;
                best_t = root; kind = 0;
            }
        }
    }
    if (fabsf(dy) > 1e-8f) {
        float t = -oc.y / dy;
        if (t >= t_min && t <= best_t) {
            float lx = (r.o.x + r.d.x * t) - c.x, lz = (r.o.z + r.d.z * t) - c.z;
            if (lx * lx + lz * lz <= rad * rad) { best_t = t; kind = 1; }
        }
    }
    t_out = best_t;
    return kind >= 0;
}

// Disk::hit (src/disk.rs:147-160)
RT_D bool disk_test(const float* p, const Ray& r, float t_min, float t_max, float& t) {
    V3 n = mk(p[3], p[4], p[5]);
    float denom = dot(n, r.d);
    if (!(fabsf(denom) >= 1e-8f)) return false;
    t = dot(n, mk(p[0], p[1], p[2]) - r.o) / denom;
    if (t < t_min || t > t_max) return false;
    V3 off = at(r, t) - mk(p[0], p[1], p[2]);
    return !(len2(off) > p[6] * p[6]);
}

// InfinitePlane::hit (src/plane.rs:40-46)
RT_D bool plane_test(const float* p, const Ray& r, float t_min, float t_max, float& t) {
    V3 n = mk(p[3], p[4], p[5]);
    float denom = dot(n, r.d);
    if (!(fabsf(denom) >= 1e-8f)) return false;
    t = dot(n, mk(p[0], p[1], p[2]) - r.o) / denom;
    return !(t < t_min || t > t_max);
}

// Closest-hit test of an analytic shape with inline parameters.
RT_D bool analytic_test(const DObject& ob, const Ray& r, float t_min, float t_max, float& t, int& kind) {
    kind = 0;
    switch (ob.shape) {
        case SH_CYLINDER: return cylinder_test(ob.p, r, t_min, t_max, t, kind);
        case SH_CONE: return cone_test(ob.p, r, t_min, t_max, t, kind);
        case SH_DISK: return disk_test(ob.p, r, t_min, t_max, t);
        default: return plane_test(ob.p, r, t_min, t_max, t);
    }
}
RT_HD bool is_analytic(uint32_t shape) { return shape >= SH_CYLINDER && shape <= SH_PLANE; }

// ---- media --------------------------------------------------------------------------
RT_D uint64_t fbits(float f) { return (uint64_t)(uint32_t)__float_as_uint(f); }
RT_D Rng medium_rng(const Ray& r) {
    return rng_seed((fbits(r.o.x) * 6364136223846793005ULL) ^ (fbits(r.o.y) * 0x9E3779B97F4A7C15ULL)
                    ^ (fbits(r.o.z) * 0xBF58476D1CE4E5B9ULL) ^ (fbits(r.d.x) * 0x94D049BB133111EBULL)
                    ^ fbits(r.wavelength));
}

// Boundary hit of a medium's shape (sphere or box) in object space.
RT_D bool boundary_t(const DScene& S, const DObject& ob, const Ray& r, float t_min, float t_max, float& t) {
    if (ob.shape == SH_SPHERE) return sphere_test(ob.p, r, t_min, t_max, t);
    // SH_BOX: closest of its six quads (a nested BvhTree in the reference;
    // the closest hit is order independent up to exact ties)
    bool found = false; float closest = t_max;
    for (uint32_t k = 0; k < 6; k++) {
        float tt, a, b;
        if (quad_test(S.quads[ob.data_first + k], r, t_min, closest, tt, a, b)) { closest = tt; found = true; }
    }
    t = closest;
    return found;
}

// Perlin::noise / turb (src/perlin.rs:35-72) over tables at S.grids + off:
// ranvec[256][3], then perm_x, perm_y, perm_z stored as exact floats.
RT_D float perlin_noise(const float* tab, V3 p) {
    float fx = floorf(p.x), fy = floorf(p.y), fz = floorf(p.z);
    float u = p.x - fx, v = p.y - fy, w = p.z - fz;
    u = u * u * (3.0f - 2.0f * u);
    v = v * v * (3.0f - 2.0f * v);
    w = w * w * (3.0f - 2.0f * w);
    uint32_t i = (uint32_t)f2i(fx), j = (uint32_t)f2i(fy), k = (uint32_t)f2i(fz);
    const float* px = tab + 768; const float* py = tab + 1024; const float* pz = tab + 1280;
    float accum = 0.0f;
    for (uint32_t di = 0; di < 2; di++)
        for (uint32_t dj = 0; dj < 2; dj++)
            for (uint32_t dk = 0; dk < 2; dk++) {
                uint32_t idx = (uint32_t)px[(i + di) & 255u] ^ (uint32_t)py[(j + dj) & 255u] ^ (uint32_t)pz[(k + dk) & 255u];
                V3 c = mk(tab[3 * idx], tab[3 * idx + 1], tab[3 * idx + 2]);
                float fi = (float)di, fj = (float)dj, fk = (float)dk;
                V3 wt = mk(u - fi, v - fj, w - fk);
                accum += (fi * u + (1.0f - fi) * (1.0f - u)) * (fj * v + (1.0f - fj) * (1.0f - v)) *
                         (fk * w + (1.0f - fk) * (1.0f - w)) * dot(c, wt);
            }
    return accum;
}
RT_D float perlin_turb(const float* tab, V3 p, int depth) {
    float accum = 0.0f, weight = 1.0f;
    V3 q = p;
    for (int d = 0; d < depth; d++) {
        accum += weight * perlin_noise(tab, q);
        weight *= 0.5f;
        q = mk(q.x * 2.0f, q.y * 2.0f, q.z * 2.0f);
    }
    return fabsf(accum);
}
// NoiseMedium::density_at (src/volume.rs:190-193)
RT_D float noise_density(const DScene& S, const DMedium& m, V3 p) {
    float n = perlin_turb(S.grids + m.grid_off, p * m.noise_scale, 7) * 0.5f;
    return fmx(n - m.threshold, 0.0f) * m.density;
}

RT_D float grid_density(const DScene& S, const DMedium& m, V3 p) {
    float fx = (p.x - m.bmin[0]) / (m.bmax[0] - m.bmin[0]) * (float)m.nx - 0.5f;
    float fy = (p.y - m.bmin[1]) / (m.bmax[1] - m.bmin[1]) * (float)m.ny - 0.5f;
    float fz = (p.z - m.bmin[2]) / (m.bmax[2] - m.bmin[2]) * (float)m.nz - 0.5f;
    fx = fclamp(fx, 0.0f, (float)(m.nx - 1)); fy = fclamp(fy, 0.0f, (float)(m.ny - 1)); fz = fclamp(fz, 0.0f, (float)(m.nz - 1));
    uint32_t x0 = f2u(fx), y0 = f2u(fy), z0 = f2u(fz);
    uint32_t x1 = min(x0 + 1, m.nx - 1), y1 = min(y0 + 1, m.ny - 1), z1 = min(z0 + 1, m.nz - 1);
    float tx = fx - (float)x0, ty = fy - (float)y0, tz = fz - (float)z0;
    const float* g = S.grids + m.grid_off;
#define GV(x, y, z) g[((size_t)(z) * m.ny + (y)) * m.nx + (x)]
    float c00 = GV(x0, y0, z0) * (1.0f - tx) + GV(x1, y0, z0) * tx;
    float c10 = GV(x0, y1, z0) * (1.0f - tx) + GV(x1, y1, z0) * tx;
    float c01 = GV(x0, y0, z1) * (1.0f - tx) + GV(x1, y0, z1) * tx;
    float c11 = GV(x0, y1, z1) * (1.0f - tx) + GV(x1, y1, z1) * tx;
#undef GV
    float c0 = c00 * (1.0f - ty) + c10 * ty;
    float c1 = c01 * (1.0f - ty) + c11 * ty;
    return (c0 * (1.0f - tz) + c1 * tz) * m.density;
}

// ConstantMedium::hit / NoiseMedium::hit (src/volume.rs:97-139, 197-236).
RT_D bool medium_hit(const DScene& S, const DObject& ob, const Ray& r, float t_min, float t_max, float& t_out) {
    float t1, t2;
    if (!boundary_t(S, ob, r, -FINF, FINF, t1)) return false;
    if (!boundary_t(S, ob, r, t1 + 0.0001f, FINF, t2)) return false;
    if (t1 < t_min) t1 = t_min;
    if (t2 > t_max) t2 = t_max;
    if (t1 >= t2) return false;
    if (t1 < 0.0f) t1 = 0.0f;
    const DMedium& m = S.media[ob.medium - 1];
    float ray_length = vlen(r.d);
    Rng rng = medium_rng(r);
    if (m.kind == MED_CONSTANT) {
        float dist_inside = (t2 - t1) * ray_length;
        float random = fmx(rng_f(rng), 1.1920929e-7f);
        float hit_dist = (-1.0f / m.density) * logf(random);
        if (hit_dist > dist_inside) return false;
        t_out = t1 + hit_dist / ray_length;
        return true;
    }
    float inv_maj = 1.0f / (m.density * ray_length);
    float t = t1;
    for (int it = 0; it < 1 << 20; it++) {
        t -= inv_maj * logf(fmx(rng_f(rng), 1.1920929e-7f));
        if (t >= t2) return false;
        float ld = m.kind == MED_NOISE ? noise_density(S, m, at(r, t)) : grid_density(S, m, at(r, t));
        if (rng_f(rng) < ld / m.density) { t_out = t; return true; }
    }
    return false;
}

RT_D float medium_transmittance(const DScene& S, const DObject& ob, const Ray& r, float t_min, float t_max) {
    float a, b;
    if (!boundary_t(S, ob, r, -FINF, FINF, a)) return 1.0f;
    if (!boundary_t(S, ob, r, a + 0.0001f, FINF, b)) return 1.0f;
    float t1 = fmx(fmx(a, 0.0f), t_min);
    float t2 = fmn(b, t_max);
    if (t1 >= t2) return 1.0f;
    float l = (t2 - t1) * vlen(r.d);
    const DMedium& m = S.media[ob.medium - 1];
    if (m.kind == MED_CONSTANT) return expf(-m.density * l);
    const int K = 16;
    float seg_t = l / vlen(r.d);
    float sum = 0.0f;
    for (int i = 0; i < K; i++) {
        float t = t1 + seg_t * ((float)i + 0.5f) / (float)K;
        sum += m.kind == MED_NOISE ? noise_density(S, m, at(r, t)) : grid_density(S, m, at(r, t));
    }
    return expf(-sum / (float)K * l);
}

// ---- closest hit ------------------------------------------------------------------------
struct HitId {
    float t;
    int32_t slot;    // world object slot (leaf order), -1 = miss
    int32_t prim;    // nested primitive slot / convex plane index / -2 = medium event
};

// Closest hit inside a nested tree (mesh or box); updates closest / prim.
RT_D void nested_closest(const DScene& S, const DObject& ob, const Ray& r, float t_min, float& closest, int32_t& prim, bool& any, Stack& st) {
    // the object's fields read once, before the loop (a load inside the leaf
    // branch cannot be hoisted by the compiler: it is conditional)
    const bool mesh = ob.shape == SH_MESH;
    const uint32_t first = ob.data_first;
    int32_t base = st.sp;
    st_push(st, (int32_t)ob.root);
    while (st.sp > base) {
        int32_t e = st_pop(st);
        if (e < 0) {
            uint32_t k = (uint32_t)(-e - 1);
            float t, a, b;
            st.n_prims++;
            bool h = mesh ? tri_test(S.tris, first + k, r, t_min, closest, t, a, b)
                          : quad_test(S.quads[first + k], r, t_min, closest, t, a, b);
            if (h) { closest = t; prim = (int32_t)k; any = true; }
            continue;
        }
        int32_t key[8]; int32_t ch[8];
        uint32_t mask = node_test(S.nodes + (size_t)e * 16, r, t_min, closest, key, ch);
        st.n_nodes++;
        if (mask) push_sorted(st, mask, key, ch);
    }
}

// Test one world object (any shape, any wrapper) for a closer hit.
RT_D void object_closest(const DScene& S, const DObject& ob, int32_t slot, const Ray& wr, float t_min, float& closest, HitId& best, Stack& st) {
    Ray r = to_object(S, ob, wr);
    if (ob.shape != SH_MESH && ob.shape != SH_BOX) st.n_prims++;
    if (ob.medium) {
        float t;
        if (medium_hit(S, ob, r, t_min, closest, t)) { closest = t; best.t = t; best.slot = slot; best.prim = -2; }
        return;
    }
    switch (ob.shape) {
        case SH_SPHERE: {
            float t;
            if (sphere_test(ob.p, r, t_min, closest, t)) { closest = t; best.t = t; best.slot = slot; best.prim = 0; }
            break;
        }
        case SH_QUAD: {
            float t, a, b;
            if (quad_test(S.quads[ob.data_first], r, t_min, closest, t, a, b)) { closest = t; best.t = t; best.slot = slot; best.prim = 0; }
            break;
        }
        case SH_CONVEX: {
            float t; int idx;
            if (convex_test(S.planes, ob.data_first, ob.data_count, r, t_min, closest, t, idx)) { closest = t; best.t = t; best.slot = slot; best.prim = idx; }
            break;
        }
        case SH_CYLINDER: case SH_CONE: case SH_DISK: case SH_PLANE: {
            float t; int kind;
            if (analytic_test(ob, r, t_min, closest, t, kind)) { closest = t; best.t = t; best.slot = slot; best.prim = kind; }
            break;
        }
        default: {
            int32_t prim = -1; bool any = false;
            nested_closest(S, ob, r, t_min, closest, prim, any, st);
            if (any) { best.t = closest; best.slot = slot; best.prim = prim; }
        }
    }
}

// BvhTree::hit_index over the world (src/bvh.rs:229-288).
RT_D HitId trace_closest(const DScene& S, const Ray& r, float t_min, float t_max, Stack& st) {
    HitId best; best.t = t_max; best.slot = -1; best.prim = -1;
    st.n_rays++;
    float closest = t_max;
    st.sp = 0;
    st_push(st, 0);
    while (st.sp > 0) {
        int32_t e = st_pop(st);
        if (e < 0) {
            int32_t slot = -e - 1;
            object_closest(S, S.objs[slot], slot, r, t_min, closest, best, st);
            continue;
        }
        int32_t key[8]; int32_t ch[8];
        uint32_t mask = node_test(S.nodes + (size_t)e * 16, r, t_min, closest, key, ch);
        st.n_nodes++;
        if (mask) push_sorted(st, mask, key, ch);
    }
    return best;
}

// ---- occlusion / transmittance ---------------------------------------------------------
RT_D bool nested_any(const DScene& S, const DObject& ob, const Ray& r, float t_min, float t_max, Stack& st) {
    const bool mesh = ob.shape == SH_MESH;
    const uint32_t first = ob.data_first;
    int32_t base = st.sp;
    st_push(st, (int32_t)ob.root);
    while (st.sp > base) {
        int32_t e = st_pop(st);
        if (e < 0) {
            uint32_t k = (uint32_t)(-e - 1);
            float t, a, b;
            st.n_prims++;
            bool h = mesh ? tri_test(S.tris, first + k, r, t_min, t_max, t, a, b)
                          : quad_test(S.quads[first + k], r, t_min, t_max, t, a, b);
            if (h) { st.sp = base; return true; }
            continue;
        }
        int32_t key[8]; int32_t ch[8];
        uint32_t mask = node_test(S.nodes + (size_t)e * 16, r, t_min, t_max, key, ch);
        st.n_nodes++;
#pragma unroll
        for (int c = 0; c < 8; c++) if (mask & (1u << c)) st_push(st, ch[c]);
    }
    return false;
}

// Hittable::transmittance of one world object (opaque surfaces: 0 or 1).
RT_D float object_transmittance(const DScene& S, const DObject& ob, const Ray& wr, float t_min, float t_max, Stack& st) {
    Ray r = to_object(S, ob, wr);
    if (ob.shape != SH_MESH && ob.shape != SH_BOX) st.n_prims++;
    float tr;
    if (ob.medium) tr = medium_transmittance(S, ob, r, t_min, t_max);
    else {
        bool h;
        switch (ob.shape) {
            case SH_SPHERE: { float t; h = sphere_test(ob.p, r, t_min, t_max, t); break; }
            case SH_QUAD: { float t, a, b; h = quad_test(S.quads[ob.data_first], r, t_min, t_max, t, a, b); break; }
            case SH_CONVEX: { float t; int idx; h = convex_test(S.planes, ob.data_first, ob.data_count, r, t_min, t_max, t, idx); break; }
            case SH_CYLINDER: case SH_CONE: case SH_DISK: case SH_PLANE: { float t; int kind; h = analytic_test(ob, r, t_min, t_max, t, kind); break; }
            default: h = nested_any(S, ob, r, t_min, t_max, st);
        }
        tr = h ? 0.0f : 1.0f;
    }
    // UniformScale: T_world = T_obj^factor (src/transform.rs:47-55)
    if (tr > 0.0f && tr < 1.0f)
        for (uint32_t k = 0; k < ob.xf_count; k++)
            if (S.xforms[ob.xf_first + k].kind == XF_SCALE) tr = powf(tr, S.xforms[ob.xf_first + k].v[0]);
    return tr;
}

// BvhTree::transmittance (src/bvh.rs:532-575); any_hit when `any_only`.
RT_D float trace_transmittance_impl(const DScene& S, const Ray& r, float t_min, float t_max, Stack& st, bool any_only);
// Shadow queries count their work separately from closest-hit queries.
RT_D float trace_transmittance(const DScene& S, const Ray& r, float t_min, float t_max, Stack& st, bool any_only) {
    uint32_t n0 = st.n_nodes, p0 = st.n_prims;
    float tr = trace_transmittance_impl(S, r, t_min, t_max, st, any_only);
    st.s_nodes += st.n_nodes - n0; st.s_prims += st.n_prims - p0; st.s_rays++;
    st.n_nodes = n0; st.n_prims = p0;
    return tr;
}
RT_D float trace_transmittance_impl(const DScene& S, const Ray& r, float t_min, float t_max, Stack& st, bool any_only) {
    float tr = 1.0f;
    st.sp = 0;
    st_push(st, 0);
    while (st.sp > 0) {
        int32_t e = st_pop(st);
        if (e < 0) {
            const DObject& ob = S.objs[-e - 1];
            if (any_only && ob.medium) continue;   // media report no any_hit (src/volume.rs:141)
            tr *= object_transmittance(S, ob, r, t_min, t_max, st);
            if (tr <= 0.0f) return 0.0f;
            continue;
        }
        int32_t key[8]; int32_t ch[8];
        uint32_t mask = node_test(S.nodes + (size_t)e * 16, r, t_min, t_max, key, ch);
        st.n_nodes++;
#pragma unroll
        for (int c = 0; c < 8; c++) if (mask & (1u << c)) st_push(st, ch[c]);
    }
    return tr;
}

// ---- full hit record for the winner ------------------------------------------------------
struct Rec {
    V3 p, n, tangent;
    float t, u, v;
    bool front;
    uint32_t mat;        // material index; medium events use mat = 0xffffffff and medium id
    uint32_t medium;
    int32_t obj_id, prim_id;
};

RT_D Rec make_rec(V3 p, float t, const Ray& r, V3 outward) {
    Rec h; h.p = p; h.t = t;
    h.front = dot(r.d, outward) < 0.0f;
    h.n = h.front ? outward : -outward;
    h.u = 0; h.v = 0; h.tangent = mk(0, 0, 0); h.medium = 0;
    return h;
}

RT_D V3 micro_tilt(uint32_t idx, float amp, V3 n) {
    if (amp <= 0.0f) return n;
    uint64_t h = ((uint64_t)idx * 0x9E3779B97F4A7C15ULL) ^ 0xD1B54A32D192ED03ULL;
    float a = ((float)(uint32_t)(h >> 32) / 4294967295.0f) * 2.0f - 1.0f;
    float b = ((float)(uint32_t)h / 4294967295.0f) * 2.0f - 1.0f;
    V3 up = fabsf(n.x) < 0.9f ? mk(1, 0, 0) : mk(0, 1, 0);
    V3 t = unit(cross(n, up));
    V3 b2 = cross(n, t);
    return unit(n + t * (a * amp) + b2 * (b * amp));
}

// Object-space record of a known winner (re-runs only the winning test).
RT_D Rec object_record(const DScene& S, const DObject& ob, const Ray& r, const HitId& id, float t_min) {
    Rec h;
    if (id.prim == -2) {
        h = make_rec(at(r, id.t), id.t, r, mk(1, 0, 0));
        h.n = mk(1, 0, 0); h.front = true; h.medium = ob.medium;
        return h;
    }
    switch (ob.shape) {
        case SH_SPHERE: {
            V3 c = mk(ob.p[0], ob.p[1], ob.p[2]);
            if (ob.p[7] != 0.0f) c = c + r.time * (mk(ob.p[4], ob.p[5], ob.p[6]) - c);
            V3 P = at(r, id.t);
            V3 outward = (P - c) / ob.p[3];
            h = make_rec(P, id.t, r, outward);
            float theta = acosf(fclamp(-outward.y, -1.0f, 1.0f));
            float phi = atan2f(-outward.z, outward.x) + PI;
            h.u = phi / (2.0f * PI); h.v = theta / PI;
            break;
        }
        case SH_QUAD: case SH_BOX: {
            const DQuad& q = S.quads[ob.data_first + (ob.shape == SH_BOX ? id.prim : 0)];
            float t, a, b;
            quad_test(q, r, t_min, id.t, t, a, b);
            h = make_rec(at(r, id.t), id.t, r, mk(q.n[0], q.n[1], q.n[2]));
            h.u = a; h.v = b;
            break;
        }
        case SH_CONVEX: {
            // no bounding plane recorded: the reference's default normal (0, 1, 0)
            V3 on = mk(0.0f, 1.0f, 0.0f);
            if (id.prim >= 0) { float4 P = S.planes[ob.data_first + id.prim]; on = mk(P.x, P.y, P.z); }
            V3 p = at(r, id.t);
            V3 rel = p - mk(ob.p[0], ob.p[1], ob.p[2]);
            float u = (atan2f(-rel.z, rel.x) + PI) / (2.0f * PI);
            float v = fclamp((rel.y + 0.5f * ob.p[5] + ob.p[4]) / (ob.p[3] + ob.p[4] + ob.p[5]), 0.0f, 1.0f);
            h = make_rec(p, id.t, r, on);
            if (ob.p[6] > 0.0f) {
                uint32_t facet = 0;
                for (uint32_t k = 0; k < ob.data_count; k++) {
                    float4 Q = S.planes[ob.data_first + k];
                    if (fabsf(Q.x - on.x) < 1e-6f && fabsf(Q.y - on.y) < 1e-6f && fabsf(Q.z - on.z) < 1e-6f) { facet = k; break; }
                }
                h.n = micro_tilt(facet, ob.p[6], h.n);
            }
            h.u = u; h.v = v;
            break;
        }
        case SH_CYLINDER: case SH_CONE: {
            // Cylinder / Cone records (src/cylinder.rs:154-183, src/cone.rs:44-78)
            V3 c = mk(ob.p[0], ob.p[1], ob.p[2]);
            float rad = ob.p[3], hgt = ob.p[4];
            V3 p = at(r, id.t);
            float lx = p.x - c.x, lz = p.z - c.z;
            if (id.prim == 0) {
                float ly = p.y - c.y;
                V3 outward;
                if (ob.shape == SH_CYLINDER) outward = mk((p.x - c.x) / rad, 0.0f, (p.z - c.z) / rad);
                else {
                    float k = (rad / hgt) * (rad / hgt);
                    V3 grad = mk(lx, k * (hgt - ly), lz);
                    outward = len2(grad) > 1e-12f ? unit(grad) : mk(0.0f, 1.0f, 0.0f);
                }
                h = make_rec(p, id.t, r, outward);
                float phi = atan2f(lz, lx);
                float uu = phi / TAU_F + 0.5f;
                h.u = uu - truncf(uu);
                h.v = ly / hgt;
            } else {
                float ny = (id.prim == 1) ? -1.0f : 1.0f;
                h = make_rec(p, id.t, r, mk(0.0f, ny, 0.0f));
                h.u = (lx / rad + 1.0f) * 0.5f;
                h.v = (lz / rad + 1.0f) * 0.5f;
            }
            break;
        }
        case SH_DISK: {
            // Disk::hit record (src/disk.rs:154-158)
            V3 p = at(r, id.t);
            V3 off = p - mk(ob.p[0], ob.p[1], ob.p[2]);
            h = make_rec(p, id.t, r, mk(ob.p[3], ob.p[4], ob.p[5]));
            h.u = (dot(off, mk(ob.p[7], ob.p[8], ob.p[9])) / ob.p[6] + 1.0f) * 0.5f;
            h.v = (dot(off, mk(ob.p[10], ob.p[11], ob.p[12])) / ob.p[6] + 1.0f) * 0.5f;
            break;
        }
        case SH_PLANE: {
            // InfinitePlane::hit record with the optional ripple normal (src/plane.rs:47-70)
            V3 n = mk(ob.p[3], ob.p[4], ob.p[5]);
            V3 ua = mk(ob.p[8], ob.p[9], ob.p[10]), va = mk(ob.p[11], ob.p[12], ob.p[13]);
            V3 p = at(r, id.t);
            V3 off = p - mk(ob.p[0], ob.p[1], ob.p[2]);
            V3 sn = n;
            float amp = ob.p[6];
            if (amp > 1e-6f) {
                float u = dot(off, ua), v = dot(off, va);
                float sc = ob.p[7];
                float phi1 = sc * u + 0.7f * sc * v;
                float phi2 = 1.3f * sc * v - 0.5f * sc * u;
                float dh_du = amp * sc * cosf(phi1) - 0.25f * amp * sc * cosf(phi2);
                float dh_dv = 0.7f * amp * sc * cosf(phi1) + 0.65f * amp * sc * cosf(phi2);
                sn = unit((n - dh_du * ua) - dh_dv * va);
            }
            h = make_rec(p, id.t, r, sn);
            h.u = rem1(dot(off, ua));
            h.v = rem1(dot(off, va));
            break;
        }
        default: {  // mesh triangle
            uint32_t i = ob.data_first + (uint32_t)id.prim;
            float t, b1, b2;
            tri_test(S.tris, i, r, t_min, id.t, t, b1, b2);
            float4 A = S.tris[3 * i], B = S.tris[3 * i + 1], Cc = S.tris[3 * i + 2];
            V3 e1 = mk(B.x, B.y, B.z), e2 = mk(Cc.x, Cc.y, Cc.z);
            float b0 = 1.0f - b1 - b2;
            V3 P = at(r, id.t);
            V3 sn; V3 n0, n1, n2;
            if (S.tri_n) {
                const float* n = S.tri_n + (size_t)i * 9;
                n0 = mk(n[0], n[1], n[2]); n1 = mk(n[3], n[4], n[5]); n2 = mk(n[6], n[7], n[8]);
            } else {
                V3 fn = unit(cross(e1, e2));   // Triangle::flat face normal
                n0 = n1 = n2 = fn;
            }
            V3 sn_raw = b0 * n0 + b1 * n1 + b2 * n2;
            sn = len2(sn_raw) > 1e-12f ? unit(sn_raw) : unit(cross(e1, e2));
            float uu, vv;
            if (S.tri_uv) {
                const float* uv = S.tri_uv + (size_t)i * 6;
                uu = b0 * uv[0] + b1 * uv[2] + b2 * uv[4];
                vv = b0 * uv[1] + b1 * uv[3] + b2 * uv[5];
            } else {
                uu = b0 * 0.0f + b1 * 1.0f + b2 * 0.5f;
                vv = b0 * 0.0f + b1 * 0.0f + b2 * 1.0f;
            }
            V3 t0, t1, t2;
            if (S.tri_t) {
                const float* tg = S.tri_t + (size_t)i * 9;
                t0 = mk(tg[0], tg[1], tg[2]); t1 = mk(tg[3], tg[4], tg[5]); t2 = mk(tg[6], tg[7], tg[8]);
            } else {
                t0 = t1 = t2 = unit(e1);
            }
            V3 raw_t = b0 * t0 + b1 * t1 + b2 * t2;
            raw_t = raw_t - sn * dot(raw_t, sn);
            V3 tangent;
            if (len2(raw_t) > 1e-8f) tangent = unit(raw_t); else { V3 a_, b_; vonb(sn, a_, b_); tangent = a_; }
            h = make_rec(P, id.t, r, sn);
            h.u = uu; h.v = vv; h.tangent = tangent;
        }
    }
    return h;
}

// Object-space record -> world (innermost wrapper first).
RT_D void record_to_world(const DScene& S, const DObject& ob, Rec& h) {
    for (uint32_t k = 0; k < ob.xf_count; k++) {
        const DXform& x = S.xforms[ob.xf_first + k];
        if (x.kind == XF_TRANSLATE) h.p += mk(x.v[0], x.v[1], x.v[2]);
        else if (x.kind == XF_ROTATE) { h.p = rot_apply(x.v, h.p); h.n = rot_apply(x.v, h.n); h.tangent = rot_apply(x.v, h.tangent); }
        else h.p = x.v[0] * h.p;
    }
}

RT_D Rec finalize(const DScene& S, const Ray& wr, const HitId& id, float t_min) {
    const DObject& ob = S.objs[id.slot];
    Ray r = to_object(S, ob, wr);
    Rec h = object_record(S, ob, r, id, t_min);
    record_to_world(S, ob, h);
    h.mat = ob.material;
    h.obj_id = (int32_t)ob.tag;
    h.prim_id = 0;
    if (ob.shape == SH_MESH && id.prim >= 0) h.prim_id = S.tri_tag[ob.data_first + id.prim];
    else if (ob.shape == SH_BOX && id.prim >= 0) h.prim_id = (int32_t)S.quads[ob.data_first + id.prim].pad[0];
    else if (ob.shape == SH_CONVEX && id.prim >= 0) h.prim_id = id.prim;    // bounding half-space (facet) index
    return h;
}

// ---- light list (HittableList, src/hittable.rs:156-225) -------------------------------------
RT_D bool light_hit(const DScene& S, const DObject& ob, const Ray& wr, float t_min, float t_max, float& t, Rec& rec, bool want_rec) {
    Ray r = to_object(S, ob, wr);
    HitId id; id.slot = 0; id.prim = 0;
    switch (ob.shape) {
        case SH_SPHERE: if (!sphere_test(ob.p, r, t_min, t_max, t)) return false; break;
        case SH_QUAD: { float a, b; if (!quad_test(S.quads[ob.data_first], r, t_min, t_max, t, a, b)) return false; break; }
        case SH_CONVEX: { int idx; if (!convex_test(S.planes, ob.data_first, ob.data_count, r, t_min, t_max, t, idx)) return false; id.prim = idx; break; }
        case SH_DISK: if (!disk_test(ob.p, r, t_min, t_max, t)) return false; break;
        default: return false;
    }
    if (want_rec) {
        id.t = t;
        rec = object_record(S, ob, r, id, t_min);
        record_to_world(S, ob, rec);
        rec.mat = ob.material;
    }
    return true;
}

RT_D bool lights_closest(const DScene& S, const Ray& r, float t_min, float t_max, Rec& rec) {
    float closest = t_max; int32_t best = -1;
    for (uint32_t i = 0; i < S.num_lights; i++) {
        float t; Rec dummy;
        if (light_hit(S, S.lights[i], r, t_min, closest, t, dummy, false)) { closest = t; best = (int32_t)i; }
    }
    if (best < 0) return false;
    float t;
    light_hit(S, S.lights[best], r, t_min, closest, t, rec, true);
    return true;
}

}  // namespace rt
