// CPU ORACLE - test infrastructure only.
//
// A scalar restatement of the reference's wavefront path tracer, used to
// check the gfx950 kernels. It is not part of the product: only tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it.
//
// Each path is traced to completion in program order (a "megakernel" loop),
// which yields the same per-path result as the reference's stage-by-stage
// wavefront because every random number is a function of (pixel, sample
// index, dimension) only. Function headers cite the reference file:line they
// restate. Traversal is brute force over all triangles or a simple
// median-split binary BVH with a conservative slab test; both return the
// lexicographic minimum of (t, flattened triangle index).
//
// Deterministic FP32 transcendentals come from the product header
// btrc_amd/csrc/core/detmath.h (pinned on its own against libm): sharing it
// is what lets host and GPU follow identical sampling decisions.
//
// Parity pinning of each part is listed in oracle/README.md.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/btrc_hip.h"
#include "../btrc_amd/csrc/core/detmath.h"

#pragma STDC FP_CONTRACT OFF

namespace orc {

const float PI = 3.1415926535f;          // utils/math/scalar.h:8
const float FMAX = 3.40282347e+38f;      // utils/math/scalar.h:7
namespace dm = btrc_detmath;

static inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

struct Vec
{
    float x, y, z;
};
static inline Vec mk(float x, float y, float z) { return Vec{ x, y, z }; }
static inline Vec add(Vec a, Vec b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline Vec sub(Vec a, Vec b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline Vec mul(Vec a, Vec b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline Vec scl(Vec a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline Vec sclf(float s, Vec a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline Vec dvs(Vec a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline Vec neg(Vec a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(Vec a, Vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline Vec cross(Vec a, Vec b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float len(Vec a) { return std::sqrt(dot(a, a)); }
static inline Vec nrm(Vec a) { const float inv = 1.0f / std::sqrt(dot(a, a)); return scl(a, inv); }
static inline bool zero(Vec a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
static inline float fmx(float a, float b) { return a > b ? a : b; }
static inline float fmn(float a, float b) { return a < b ? a : b; }
static inline float sq(float x) { return x * x; }
static inline float lum(Vec c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; } // core/spectrum.cpp:84
static inline Vec lerpv(Vec a, Vec b, float t) { return add(scl(a, 1.0f - t), scl(b, t)); }
static inline bool finite3(Vec a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
static inline float el(Vec a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

struct Frm
{
    Vec x, y, z;
};
static inline Frm frm(Vec x, Vec y, Vec z) { return Frm{ nrm(x), nrm(y), nrm(z) }; } // cframe.cpp:18-23
static inline Vec l2g(const Frm &f, Vec l) { return add(add(scl(f.x, l.x), scl(f.y, l.y)), scl(f.z, l.z)); }
static inline Vec g2l(const Frm &f, Vec g) { return mk(dot(g, f.x), dot(g, f.y), dot(g, f.z)); }
static inline Frm rot_z(const Frm &f, Vec z) // cframe.cpp:79-84
{
    const Vec nx = cross(f.y, z);
    const Vec ny = cross(z, nx);
    return frm(nx, ny, z);
}
static inline Frm from_z(Vec z) // utils/math/frame.cpp:45-57
{
    const Vec nz = nrm(z);
    const Vec ny = 1.0f - std::fabs(nz.x) < 0.1f ? cross(nz, mk(0, 1, 0)) : cross(nz, mk(1, 0, 0));
    return frm(cross(nz, ny), ny, nz);
}

struct Xf
{
    float m[12], n[12];
};
static inline Vec xpoint(const Xf &t, Vec p) // ctransform.cpp:31-37
{
    return mk(t.m[0] * p.x + t.m[1] * p.y + t.m[2] * p.z + t.m[3],
              t.m[4] * p.x + t.m[5] * p.y + t.m[6] * p.z + t.m[7],
              t.m[8] * p.x + t.m[9] * p.y + t.m[10] * p.z + t.m[11]);
}
static inline Vec xvec(const Xf &t, Vec v) // ctransform.cpp:39-47
{
    return mk(t.m[0] * v.x + t.m[1] * v.y + t.m[2] * v.z,
              t.m[4] * v.x + t.m[5] * v.y + t.m[6] * v.z,
              t.m[8] * v.x + t.m[9] * v.y + t.m[10] * v.z);
}
static inline Vec xnor(const Xf &t, Vec n) // ctransform.cpp:49-55
{
    return mk(t.n[0] * n.x + t.n[4] * n.y + t.n[8] * n.z,
              t.n[1] * n.x + t.n[5] * n.y + t.n[9] * n.z,
              t.n[2] * n.x + t.n[6] * n.y + t.n[10] * n.z);
}

// utils/intersection.cpp:5-37
static Vec offset_pos(Vec p, Vec n)
{
    const float ORIGIN = 1.0f / 32.0f, FS = 1.0f / 65536.0f, IS = 256.0f;
    const int32_t ox = (int32_t)(2.0f * IS * n.x), oy = (int32_t)(2.0f * IS * n.y), oz = (int32_t)(2.0f * IS * n.z);
    auto bump = [](float v, int32_t o) {
        int32_t b; std::memcpy(&b, &v, 4);
        b += v < 0.0f ? -o : o;
        float r; std::memcpy(&r, &b, 4);
        return r;
    };
    return mk(std::fabs(p.x) < ORIGIN ? p.x + FS * n.x : bump(p.x, ox),
              std::fabs(p.y) < ORIGIN ? p.y + FS * n.y : bump(p.y, oy),
              std::fabs(p.z) < ORIGIN ? p.z + FS * n.z : bump(p.z, oz));
}
static Vec offset_dir(Vec p, Vec n, Vec d) { return offset_pos(p, dot(n, d) > 0.0f ? n : neg(n)); }

// utils/cmath/cdistribution.cpp
static Vec sphere_uniform(float u1, float u2)
{
    const float z = 1.0f - (u1 + u1);
    const float phi = 2.0f * PI * u2;
    const float r = std::sqrt(fmx(0.0f, 1.0f - z * z));
    return mk(r * dm::cos(phi), r * dm::sin(phi), z);
}
static Vec hemisphere_cos(float u1, float u2)
{
    float sx = 0, sy = 0;
    u1 = 2.0f * u1 - 1.0f;
    u2 = 2.0f * u2 - 1.0f;
    if(u1 != 0.0f || u2 != 0.0f)
    {
        float th, r;
        if(std::fabs(u1) > std::fabs(u2)) { r = u1; th = 0.25f * PI * (u2 / u1); }
        else { r = u2; th = 0.5f * PI - 0.25f * PI * (u1 / u2); }
        sx = r * dm::cos(th);
        sy = r * dm::sin(th);
    }
    return mk(sx, sy, std::sqrt(fmx(0.0f, 1.0f - (sx * sx + sy * sy))));
}
static float hemisphere_cos_pdf(Vec v) { return v.z >= 0.0f ? v.z / PI : 0.0f; }

// ------------------------------------------------------------------ Halton

// first 1000 primes (utils/prime.cpp) by a sieve
static const std::vector<uint32_t> &primes()
{
    static std::vector<uint32_t> p = [] {
        std::vector<uint32_t> r;
        std::vector<bool> comp(8000, false);
        for(uint32_t i = 2; r.size() < 1000; ++i)
        {
            if(comp[i])
                continue;
            r.push_back(i);
            for(uint32_t j = i * i; j < 8000; j += i)
                comp[j] = true;
        }
        return r;
    }();
    return p;
}

// utils/low_discrepancy.cpp:49-57
static uint64_t mixbits(uint64_t v)
{
    v ^= v >> 31;
    v *= 0x7fb5d329728ea185ull;
    v ^= v >> 27;
    v *= 0x81dadef4bc2dd44dull;
    v ^= v >> 33;
    return v;
}

// utils/low_discrepancy.cpp:12-45
static uint32_t perm_elem(uint32_t i, uint32_t l, uint32_t p)
{
    uint32_t w = l - 1;
    w |= w >> 1; w |= w >> 2; w |= w >> 4; w |= w >> 8; w |= w >> 16;
    do
    {
        i ^= p; i *= 0xe170893du; i ^= p >> 16; i ^= (i & w) >> 4;
        i ^= p >> 8; i *= 0x0929eb3fu; i ^= p >> 23; i ^= (i & w) >> 1;
        i *= 1u | p >> 27; i *= 0x6935fa69u; i ^= (i & w) >> 11; i *= 0x74dcb303u;
        i ^= (i & w) >> 2; i *= 0x9e501cc3u; i ^= (i & w) >> 2; i *= 0xc860a3dfu;
        i &= w; i ^= i >> 5;
    } while(i >= l);
    return (i + p) % l;
}

// utils/low_discrepancy.cpp:83-101
static float owen_ri(uint32_t base, uint64_t a, uint64_t hash)
{
    const float inv_base = 1.0f / (float)base;
    float inv_base_m = 1.0f;
    uint64_t rev = 0;
    while(1.0f - inv_base_m < 1.0f)
    {
        const uint64_t next = a / base;
        uint64_t digit = a - next * base;
        digit = perm_elem((uint32_t)digit, base, (uint32_t)mixbits(hash ^ rev));
        rev = rev * base + digit;
        inv_base_m *= inv_base;
        a = next;
    }
    return fmn(inv_base_m * (float)rev, 0x1.fffffep-1f);
}

struct Halton
{
    int32_t scales[2], exps[2], minv[2];
    uint32_t dim;
    uint64_t idx;

    // sampler/halton.cpp:118-133 and 46-66
    void init(int32_t w, int32_t h, uint32_t px, uint32_t py, int32_t sample)
    {
        const int32_t res[2] = { w, h };
        for(int i = 0; i < 2; ++i)
        {
            const int32_t base = i == 0 ? 2 : 3;
            int32_t s = 1, e = 0;
            while(s < std::min(res[i], 128)) { s *= base; ++e; }
            scales[i] = s;
            exps[i] = e;
        }
        auto inv = [](int64_t a, int64_t n) {
            // extended Euclid (halton.cpp:18-37)
            int64_t r0 = a, r1 = n, s0 = 1, s1 = 0;
            while(r1) { const int64_t q = r0 / r1; int64_t t = r0 - q * r1; r0 = r1; r1 = t; t = s0 - q * s1; s0 = s1; s1 = t; }
            const int64_t m = s0 % n;
            return m < 0 ? m + n : m;
        };
        minv[0] = (int32_t)inv(scales[1], scales[0]);
        minv[1] = (int32_t)inv(scales[0], scales[1]);
        idx = 0;
        const int32_t stride = scales[0] * scales[1];
        if(stride > 1)
        {
            const uint64_t pm[2] = { px % 128u, py % 128u };
            for(int i = 0; i < 2; ++i)
            {
                uint64_t v = pm[i], off = 0;
                const uint64_t b = i == 0 ? 2 : 3;
                for(int k = 0; k < exps[i]; ++k) { off = off * b + v % b; v /= b; }
                idx += off * (uint64_t)(stride / scales[i]) * (uint64_t)minv[i];
            }
            idx %= (uint64_t)stride;
        }
        idx += (uint64_t)(sample * stride);
        dim = 2;
    }
    float sample(uint32_t d) const
    {
        return owen_ri(primes()[d], idx, mixbits((uint64_t)(1u + (d << 4))));
    }
    float get1d()
    {
        if(dim >= 1000) dim = 2;
        const float r = sample(dim);
        dim += 1;
        return r;
    }
    void get3d(float &a, float &b, float &c)
    {
        if(dim + 2 >= 1000) dim = 2;
        a = sample(dim); b = sample(dim + 1); c = sample(dim + 2);
        dim += 3;
    }
};

// MurmurHash64A (Austin Appleby, public domain; utils/hash.h:12-71 follows
// it) over an arbitrary byte string: 8-byte little-endian blocks, then the
// 1..7 tail bytes, then the final avalanche.
static uint64_t murmur64a(const uint8_t *key, size_t len, uint64_t seed)
{
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    const int r = 47;
    uint64_t h = seed ^ (len * m);
    const size_t blocks = len / 8;
    for(size_t b = 0; b < blocks; ++b)
    {
        uint64_t k = 0;
        for(int j = 7; j >= 0; --j)
            k = (k << 8) | key[8 * b + j];
        k *= m;
        k ^= k >> r;
        k *= m;
        h ^= k;
        h *= m;
    }
    const uint8_t *tail = key + 8 * blocks;
    const size_t rest = len & 7;
    if(rest)
    {
        for(size_t j = rest; j-- > 0;)
            h ^= (uint64_t)tail[j] << (8 * j);
        h *= m;
    }
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}

// hash::hash(pixel.x, pixel.y) (utils/hash.h:85-97): the two uint32 values'
// bytes back to back, seed 0
static uint64_t hash_pixel(uint32_t x, uint32_t y)
{
    uint8_t buf[8];
    std::memcpy(buf, &x, 4);
    std::memcpy(buf + 4, &y, 4);
    return murmur64a(buf, 8, 0);
}

// PCG32 (O'Neill 2014): LCG x -> a*x + inc mod 2^64, output XSH-RR of the
// old state. Jump-ahead composes the affine step map by binary powering.
struct Pcg
{
    static constexpr uint64_t A = 6364136223846793005ull;
    uint64_t state = 0, inc = 1;

    uint32_t next()
    {
        const uint64_t x = state;
        state = A * x + inc;
        const uint32_t v = (uint32_t)((x ^ (x >> 18)) >> 27);
        const unsigned rot = (unsigned)(x >> 59);
        return rot ? (v >> rot) | (v << (32 - rot)) : v;
    }
    // pcg32_srandom_r(initstate = offset, initseq = seq)
    void seed(uint64_t seq, uint64_t offset)
    {
        state = 0;
        inc = 2 * seq + 1;
        next();
        state += offset;
        next();
    }
    void advance(uint64_t n)
    {
        // (a, c) is the map x -> a x + c; result = step^n
        uint64_t ra = 1, rc = 0, sa = A, sc = inc;
        for(; n; n >>= 1)
        {
            if(n & 1)
            {
                rc = sa * rc + sc; // step o result
                ra = sa * ra;
            }
            sc = sa * sc + sc;     // step o step
            sa = sa * sa;
        }
        state = ra * state + rc;
    }
    float uniform()
    {
        const float f = (float)next() / 4294967296.0f;
        return std::min(f, 0x1.fffffep-1f);
    }
    float get1d() { return uniform(); }
};

// IndependentSampler (sampler/independent.cpp:11-25), pbrt-v4 RNG(seq) seeding
struct Indep
{
    Pcg rng;
    void init(uint32_t px, uint32_t py, int32_t sample)
    {
        const uint64_t seq = hash_pixel(px, py);
        rng.seed(seq, mixbits(seq));
        rng.advance((uint64_t)((int64_t)sample * 65536));
    }
    float get1d() { return rng.uniform(); }
    void get3d(float &a, float &b, float &c)
    {
        a = get1d();
        b = get1d();
        c = get1d();
    }
};

// the path's sampler: Halton or independent
struct Smp
{
    bool indep = false;
    Halton h;
    Indep p;
    float get1d() { return indep ? p.get1d() : h.get1d(); }
    void get3d(float &a, float &b, float &c)
    {
        if(indep)
            p.get3d(a, b, c);
        else
            h.get3d(a, b, c);
    }
    // the four words a path's sampler keeps between stages (pixel first)
    void pack(uint32_t pixel, uint32_t w[4]) const
    {
        w[0] = pixel;
        if(indep)
        {
            w[1] = (uint32_t)p.rng.state;
            w[2] = (uint32_t)(p.rng.state >> 32);
            w[3] = 0;
        }
        else
        {
            w[1] = h.dim;
            w[2] = (uint32_t)h.idx;
            w[3] = (uint32_t)(h.idx >> 32);
        }
    }
    struct Packed { uint32_t w[4]; };
    Packed pack(uint32_t pixel) const
    {
        Packed r;
        pack(pixel, r.w);
        return r;
    }
};

// seed of a shadow ray's transmittance stream: a hash of the path's sampler
// words when the ray is made (this build's choice; the reference draws from
// per-slot shadow samplers, shadow.cpp:45-49)
static uint64_t shadow_seed(const Smp::Packed &p)
{
    const uint64_t a = ((uint64_t)p.w[0] << 32) | p.w[1];
    const uint64_t b = ((uint64_t)p.w[2] << 32) | p.w[3];
    return mixbits(a ^ mixbits(b));
}

// ------------------------------------------------------------------ scene

struct Light
{
    int type; // 0 mesh, 1 sky, 2 IBL
    int inst, geo;
    float scale;
    Vec intensity;
    Vec up, lower, upper;
    // IBL (light/ibl.cpp): frame, texture, importance table
    Frm fr;
    int tex = -1, lw = 0, lh = 0;
    std::vector<float> probs, accept;
    std::vector<uint32_t> alias;
};

// AliasTable::initialize, utils/math/alias.cpp:13-58 (sum in index order);
// returns the sum of the weights
static float alias_build(const std::vector<float> &w, std::vector<float> &accept, std::vector<uint32_t> &alias)
{
    float sum = 0.0f;
    for(float a : w)
        sum += a;
    const uint32_t n = (uint32_t)w.size();
    accept.resize(n);
    alias.resize(n);
    if(n == 0)
        return sum;
    const float ratio = (float)n / sum;
    std::vector<uint32_t> ov, un;
    for(uint32_t i = 0; i < n; ++i)
    {
        accept[i] = w[i] * ratio;
        alias[i] = i;
        if(accept[i] > 1) ov.push_back(i);
        else if(accept[i] < 1) un.push_back(i);
    }
    while(!ov.empty() && !un.empty())
    {
        const uint32_t o = ov.back(), u = un.back();
        ov.pop_back();
        un.pop_back();
        accept[o] -= 1 - accept[u];
        alias[u] = o;
        if(accept[o] > 1) ov.push_back(o);
        else if(accept[o] < 1) un.push_back(o);
    }
    for(auto i : ov) accept[i] = 1;
    for(auto i : un) accept[i] = 1;
    return sum;
}

struct Geo
{
    uint32_t off, n;
    float total_area;
    std::vector<float> accept;
    std::vector<uint32_t> alias;
};

struct Inst
{
    int geo, mat, light;
    uint32_t inner, outer; // medium ids
    Xf xf;
};

// a participating medium (medium/homogeneous.cpp, medium/hetergeneous.cpp)
struct Med
{
    int type;
    float sigma_t, inv_maj, g;
    Vec albedo;
    int nx, ny, nz;
    std::vector<float> dens, alb;
    bool has_xf;
    float w2l[12];
};

// an Array2D texture (texture2d/array2d.cpp:19-30 over utils/cuda/texture.cpp:63-105)
struct Tex
{
    int w, h, format, au, av, filter, srgb;
    float border[4];
    std::vector<uint32_t> t8;
    std::vector<float> t32;
};

struct Scene
{
    std::vector<Geo> geos;
    std::vector<btrc_material_desc> mats;
    std::vector<Inst> insts;
    std::vector<Light> lights;
    int env = -1;
    std::vector<Med> media; // by descending priority
    int R = 0;
    struct Vol
    {
        Vec o, x, y, z;
        float ix2, iy2, iz2;  // 1 / |edge|^2
        Vec xd, yd, zd;       // edge / |edge|^2
        float maxd, sigma;
        int nx = 0, ny = 0, nz = 0, ax = 0, ay = 0, az = 0;
        std::vector<float> dens, alb;
        Vec albedo;
    };
    std::vector<Vol> vols;    // the volume-primitive medium
    Vec bb_lo{ FMAX, FMAX, FMAX }, bb_hi{ -FMAX, -FMAX, -FMAX }; // Scene::commit's bbox
    // per triangle of all geometries: frame records (triangle_mesh.cpp:40-64)
    std::vector<Vec> gx, gy, gz, sa, sba, sca, la, lba, lca;
    // texture coordinates (a, b - a, c - a) per triangle, u then v
    std::vector<float> tuv;
    std::vector<Tex> texs;
    float decode[512];
    // flattened world triangles
    std::vector<Vec> w0, w1, w2;
    std::vector<uint32_t> tinst, tprim;
    // two levels: local triangles per geometry (off + prim), and per
    // instance its first flattened index and padded world box
    bool two_level = false;
    float extent = 0.0f; // largest |coordinate| of the world-space triangles
    std::vector<Vec> l0, l1, l2;
    std::vector<uint32_t> ibase;
    std::vector<float> ibox; // 6 per instance
    // binary BVH
    struct Node { float lo[3], hi[3]; int32_t left, right; uint32_t first, count; };
    std::vector<Node> nodes;
    std::vector<uint32_t> order;
};

static void build_bvh(Scene &s)
{
    const uint32_t n = (uint32_t)s.w0.size();
    s.order.resize(n);
    for(uint32_t i = 0; i < n; ++i)
        s.order[i] = i;
    s.nodes.clear();
    if(n == 0)
        return;
    struct T { int32_t node; uint32_t b, e; };
    auto bounds = [&](uint32_t b, uint32_t e, Scene::Node &nd) {
        for(int a = 0; a < 3; ++a) { nd.lo[a] = INFINITY; nd.hi[a] = -INFINITY; }
        for(uint32_t i = b; i < e; ++i)
        {
            const uint32_t t = s.order[i];
            for(const Vec &v : { s.w0[t], s.w1[t], s.w2[t] })
                for(int a = 0; a < 3; ++a)
                {
                    nd.lo[a] = std::min(nd.lo[a], el(v, a));
                    nd.hi[a] = std::max(nd.hi[a], el(v, a));
                }
        }
    };
    s.nodes.push_back(Scene::Node{});
    std::vector<T> st{ { 0, 0, n } };
    while(!st.empty())
    {
        const T t = st.back();
        st.pop_back();
        Scene::Node &nd = s.nodes[t.node];
        bounds(t.b, t.e, nd);
        nd.left = nd.right = -1;
        if(t.e - t.b <= 2)
        {
            nd.first = t.b;
            nd.count = t.e - t.b;
            continue;
        }
        int axis = 0;
        float ext = -1;
        for(int a = 0; a < 3; ++a)
            if(nd.hi[a] - nd.lo[a] > ext) { ext = nd.hi[a] - nd.lo[a]; axis = a; }
        const uint32_t mid = (t.b + t.e) / 2;
        auto key = [&](uint32_t i) {
            return el(s.w0[i], axis) + el(s.w1[i], axis) + el(s.w2[i], axis);
        };
        std::nth_element(s.order.begin() + t.b, s.order.begin() + mid, s.order.begin() + t.e,
                         [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
        const int32_t l = (int32_t)s.nodes.size();
        s.nodes.push_back(Scene::Node{});
        s.nodes.push_back(Scene::Node{});
        s.nodes[t.node].left = l;
        s.nodes[t.node].right = l + 1;
        st.push_back({ l, t.b, mid });
        st.push_back({ l + 1, mid, t.e });
    }
}

static void ibl_table(const Scene &s, Light &l);

static Scene *make_scene(const btrc_scene_desc *sd)
{
    Scene *s = new Scene;
    uint32_t off = 0;
    for(int gi = 0; gi < sd->geometry_count; ++gi)
    {
        const btrc_geometry_desc &g = sd->geometries[gi];
        Geo geo;
        geo.off = off;
        geo.n = (uint32_t)g.triangle_count;
        std::vector<float> area(g.triangle_count);
        for(int t = 0; t < g.triangle_count; ++t)
        {
            const float *p = g.positions + 9 * (size_t)t;
            const Vec a = mk(p[0], p[1], p[2]), b = mk(p[3], p[4], p[5]), c = mk(p[6], p[7], p[8]);
            // utils/triangle_mesh_loader.cpp:84-127
            const Vec gzv = nrm(cross(sub(b, a), sub(c, a)));
            Vec nz[3];
            for(int k = 0; k < 3; ++k)
            {
                nz[k] = gzv;
                if(g.normals)
                {
                    const float *q = g.normals + 9 * (size_t)t + 3 * k;
                    const Vec v = mk(q[0], q[1], q[2]);
                    if(!zero(v))
                        nz[k] = v;
                }
            }
            float uv[6] = { 0, 0, 0, 0, 0, 0 };
            if(g.tex_coords)
                std::memcpy(uv, g.tex_coords + 6 * (size_t)t, sizeof(uv));
            // triangle_dpdu, utils/math/triangle.h:8-23
            Vec ex;
            {
                const Vec BA = sub(b, a), CA = sub(c, a);
                const float m00 = uv[2] - uv[0], m01 = uv[3] - uv[1], m10 = uv[4] - uv[0], m11 = uv[5] - uv[1];
                const float det = m00 * m11 - m01 * m10;
                if(det == 0.0f)
                    ex = from_z(gzv).x;
                else
                {
                    const float id = 1.0f / det;
                    ex = nrm(sub(sclf(m11 * id, BA), sclf(m01 * id, CA)));
                }
            }
            for(float x : { uv[0], uv[2] - uv[0], uv[4] - uv[0], uv[1], uv[3] - uv[1], uv[5] - uv[1] })
                s->tuv.push_back(x);
            s->gx.push_back(ex);
            s->gy.push_back(nrm(cross(gzv, ex)));
            s->gz.push_back(gzv);
            s->sa.push_back(nz[0]);
            s->sba.push_back(sub(nz[1], nz[0]));
            s->sca.push_back(sub(nz[2], nz[0]));
            const Vec ba = sub(b, a), ca = sub(c, a);
            s->l0.push_back(a);
            s->l1.push_back(b);
            s->l2.push_back(c);
            s->la.push_back(a);
            s->lba.push_back(ba);
            s->lca.push_back(ca);
            area[t] = 0.5f * len(cross(ba, ca)); // utils/math/triangle.h:25-28
        }
        geo.total_area = alias_build(area, geo.accept, geo.alias);
        s->geos.push_back(std::move(geo));
        off += (uint32_t)g.triangle_count;
    }
    s->mats.assign(sd->materials, sd->materials + sd->material_count);
    for(int i = 0; i < sd->texture_count; ++i)
    {
        const btrc_texture_desc &t = sd->textures[i];
        Tex x;
        x.w = t.width; x.h = t.height; x.format = t.format;
        x.au = t.address_u; x.av = t.address_v; x.filter = t.filter; x.srgb = t.srgb_to_linear != 0;
        for(int k = 0; k < 4; ++k)
            x.border[k] = t.border[k];
        const size_t n = (size_t)t.width * t.height;
        if(t.format == BTRC_TEXTURE_RGBA8)
        {
            const uint8_t *p = static_cast<const uint8_t *>(t.texels);
            for(size_t k = 0; k < n; ++k)
                x.t8.push_back((uint32_t)p[4 * k] | ((uint32_t)p[4 * k + 1] << 8) | ((uint32_t)p[4 * k + 2] << 16) |
                               ((uint32_t)p[4 * k + 3] << 24));
        }
        else
            x.t32.assign(static_cast<const float *>(t.texels), static_cast<const float *>(t.texels) + 4 * n);
        s->texs.push_back(std::move(x));
    }
    // UNorm8 value / 255 and the inverse sRGB transfer function (IEC 61966-2-1)
    for(int v = 0; v < 256; ++v)
    {
        const double c = v / 255.0;
        s->decode[v] = (float)c;
        s->decode[256 + v] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
    }
    // Scene::commit's medium ids (core/scene.cpp:144-161): descending
    // priority, the volume-primitive medium next, then "none"
    s->R = sd->medium_count;
    std::vector<int> by_priority(s->R);
    for(int i = 0; i < s->R; ++i)
        by_priority[i] = i;
    std::stable_sort(by_priority.begin(), by_priority.end(),
                     [&](int a, int b) { return sd->media[a].priority > sd->media[b].priority; });
    std::vector<uint32_t> medium_rank(s->R);
    for(int r = 0; r < s->R; ++r)
    {
        const btrc_medium_desc &md = sd->media[by_priority[r]];
        medium_rank[by_priority[r]] = (uint32_t)r;
        Med m{};
        m.type = md.type;
        m.sigma_t = md.sigma_t;
        m.g = md.g;
        m.albedo = mk(md.albedo[0], md.albedo[1], md.albedo[2]);
        if(md.type == BTRC_MEDIUM_HETEROGENEOUS)
        {
            m.nx = md.grid_size[0];
            m.ny = md.grid_size[1];
            m.nz = md.grid_size[2];
            const size_t nv = (size_t)m.nx * m.ny * m.nz;
            m.dens.assign(md.sigma_t_grid, md.sigma_t_grid + nv);
            float mx = m.dens[0];
            for(float v : m.dens)
                mx = std::max(mx, v);
            m.inv_maj = 1.0f / std::max(mx, 0.001f); // hetergeneous.cpp:34-35
            if(md.albedo_grid)
                m.alb.assign(md.albedo_grid, md.albedo_grid + 3 * nv);
        }
        m.has_xf = md.has_transform != 0;
        for(int k = 0; k < 12; ++k)
            m.w2l[k] = md.world_to_local[k];
        s->media.push_back(std::move(m));
    }
    for(int i = 0; i < sd->instance_count; ++i)
    {
        const btrc_instance_desc &d = sd->instances[i];
        Inst in;
        in.geo = d.geometry;
        in.mat = d.material;
        in.light = -1;
        in.inner = d.inner_medium < 0 ? (uint32_t)s->R + 1 : medium_rank[d.inner_medium];
        in.outer = d.outer_medium < 0 ? (uint32_t)s->R + 1 : medium_rank[d.outer_medium];
        for(int k = 0; k < 12; ++k)
        {
            in.xf.m[k] = d.local_to_world[k];
            in.xf.n[k] = d.world_to_local[k];
        }
        if(d.has_light)
        {
            // MeshLight::set_geometry, mesh_light.cpp:17-22
            Light l{};
            l.type = 0;
            l.inst = i;
            l.geo = d.geometry;
            l.scale = len(sub(xpoint(in.xf, mk(1, 0, 0)), xpoint(in.xf, mk(0, 0, 0))));
            l.intensity = mk(d.light_intensity[0], d.light_intensity[1], d.light_intensity[2]);
            in.light = (int)s->lights.size();
            s->lights.push_back(l);
        }
        s->insts.push_back(in);
        s->ibase.push_back((uint32_t)s->w0.size());
        const btrc_geometry_desc &g = sd->geometries[d.geometry];
        for(int t = 0; t < g.triangle_count; ++t)
        {
            const float *p = g.positions + 9 * (size_t)t;
            s->w0.push_back(xpoint(in.xf, mk(p[0], p[1], p[2])));
            s->w1.push_back(xpoint(in.xf, mk(p[3], p[4], p[5])));
            s->w2.push_back(xpoint(in.xf, mk(p[6], p[7], p[8])));
            s->tinst.push_back((uint32_t)i);
            s->tprim.push_back((uint32_t)t);
        }
    }
    if(sd->envir_light.type == BTRC_ENVIR_GRADIENT_SKY)
    {
        // Scene::commit puts the environment light last (core/scene.cpp:92-99)
        Light l{};
        l.type = 1;
        l.up = nrm(mk(sd->envir_light.up[0], sd->envir_light.up[1], sd->envir_light.up[2]));
        l.lower = mk(sd->envir_light.lower[0], sd->envir_light.lower[1], sd->envir_light.lower[2]);
        l.upper = mk(sd->envir_light.upper[0], sd->envir_light.upper[1], sd->envir_light.upper[2]);
        s->env = (int)s->lights.size();
        s->lights.push_back(l);
    }
    else if(sd->envir_light.type == BTRC_ENVIR_IBL)
    {
        const btrc_envir_light_desc &e = sd->envir_light;
        Light l{};
        l.type = 2;
        // Frame::from_z (utils/math/frame.cpp:45-57), normalized again by CFrame(Frame) (cframe.cpp:11-22)
        const Frm f0 = from_z(mk(e.up[0], e.up[1], e.up[2]));
        l.fr = frm(f0.x, f0.y, f0.z);
        l.tex = e.texture;
        l.lw = e.lut_width;
        l.lh = e.lut_height;
        s->env = (int)s->lights.size();
        s->lights.push_back(l);
        ibl_table(*s, s->lights.back());
    }
    build_bvh(*s);
    for(const std::vector<Vec> *w : { &s->w0, &s->w1, &s->w2 })
        for(const Vec &v : *w)
            s->extent = std::max(s->extent, std::max(std::fabs(v.x), std::max(std::fabs(v.y), std::fabs(v.z))));
    s->two_level = sd->accel == BTRC_ACCEL_TWO_LEVEL;
    for(const std::vector<Vec> *w : { &s->w0, &s->w1, &s->w2 })
        for(const Vec &v : *w)
        {
            s->bb_lo = mk(fmn(s->bb_lo.x, v.x), fmn(s->bb_lo.y, v.y), fmn(s->bb_lo.z, v.z));
            s->bb_hi = mk(fmx(s->bb_hi.x, v.x), fmx(s->bb_hi.y, v.y), fmx(s->bb_hi.z, v.z));
        }
    for(int i = 0; i < sd->volume_count; ++i)
    {
        // VolumePrimitive (core/volume.cpp:21-79) and the BVH's primitive
        // record (core/volume/bvh.cpp:83-92)
        const btrc_volume_desc &vd = sd->volumes[i];
        Scene::Vol v;
        v.o = mk(vd.o[0], vd.o[1], vd.o[2]);
        v.x = mk(vd.x[0], vd.x[1], vd.x[2]);
        v.y = mk(vd.y[0], vd.y[1], vd.y[2]);
        v.z = mk(vd.z[0], vd.z[1], vd.z[2]);
        v.ix2 = 1.0f / dot(v.x, v.x);
        v.iy2 = 1.0f / dot(v.y, v.y);
        v.iz2 = 1.0f / dot(v.z, v.z);
        v.xd = dvs(v.x, dot(v.x, v.x));
        v.yd = dvs(v.y, dot(v.y, v.y));
        v.zd = dvs(v.z, dot(v.z, v.z));
        v.sigma = vd.sigma_t;
        v.maxd = vd.sigma_t;
        if(vd.sigma_t_size[0] > 0)
        {
            v.nx = vd.sigma_t_size[0]; v.ny = vd.sigma_t_size[1]; v.nz = vd.sigma_t_size[2];
            v.dens.assign(vd.sigma_t_grid, vd.sigma_t_grid + (size_t)v.nx * v.ny * v.nz);
            v.maxd = v.dens[0];
            for(float x : v.dens)
                v.maxd = fmx(v.maxd, x);
        }
        v.albedo = mk(vd.albedo[0], vd.albedo[1], vd.albedo[2]);
        if(vd.albedo_size[0] > 0)
        {
            v.ax = vd.albedo_size[0]; v.ay = vd.albedo_size[1]; v.az = vd.albedo_size[2];
            v.alb.assign(vd.albedo_grid, vd.albedo_grid + 3 * (size_t)v.ax * v.ay * v.az);
        }
        // VolumePrimitive::get_bounding_box: o and the three far corners
        for(const Vec &c : { v.o, add(v.o, v.x), add(v.o, v.y), add(v.o, v.z) })
        {
            s->bb_lo = mk(fmn(s->bb_lo.x, c.x), fmn(s->bb_lo.y, c.y), fmn(s->bb_lo.z, c.z));
            s->bb_hi = mk(fmx(s->bb_hi.x, c.x), fmx(s->bb_hi.y, c.y), fmx(s->bb_hi.z, c.z));
        }
        s->vols.push_back(std::move(v));
    }
    if(s->two_level)
    {
        // instance world boxes: the geometry box's transformed corners,
        // padded by 2^-14 of the largest coordinate (a culling aid only:
        // the result is the minimum over all instances either way)
        float mag = 0.0f;
        s->ibox.assign(6 * s->insts.size(), 0.0f);
        for(size_t i = 0; i < s->insts.size(); ++i)
        {
            const Inst &in = s->insts[i];
            const Geo &g = s->geos[in.geo];
            float *w = &s->ibox[6 * i];
            w[0] = w[1] = w[2] = FMAX;
            w[3] = w[4] = w[5] = -FMAX;
            for(uint32_t t = g.off; t < g.off + g.n; ++t)
                for(const Vec &v : { s->l0[t], s->l1[t], s->l2[t] })
                {
                    const Vec p = xpoint(in.xf, v);
                    for(int a = 0; a < 3; ++a)
                    {
                        w[a] = std::min(w[a], el(p, a));
                        w[3 + a] = std::max(w[3 + a], el(p, a));
                    }
                }
            if(g.n)
                for(int a = 0; a < 6; ++a)
                    mag = std::max(mag, std::fabs(w[a]));
        }
        for(size_t i = 0; i < s->insts.size(); ++i)
            if(s->geos[s->insts[i].geo].n)
                for(int a = 0; a < 3; ++a)
                {
                    s->ibox[6 * i + a] -= mag / 16384.0f;
                    s->ibox[6 * i + 3 + a] += mag / 16384.0f;
                }
    }
    return s;
}

// ------------------------------------------------------------------ intersection

struct Hit
{
    float t, u, v;
    uint32_t gid; // 0xffffffff = miss
};

// Watertight test (Woop, Benthin, Wald, JCGT 2013), single precision.
static bool tri_watertight(Vec o, Vec d, Vec v0, Vec v1, Vec v2, float tmin, float tmax, float &t, float &b1, float &b2)
{
    const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    const int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if(el(d, kz) < 0.0f)
        std::swap(kx, ky);
    const float Sx = el(d, kx) / el(d, kz), Sy = el(d, ky) / el(d, kz), Sz = 1.0f / el(d, kz);
    const Vec A = sub(v0, o), B = sub(v1, o), C = sub(v2, o);
    const float Ax = el(A, kx) - Sx * el(A, kz), Ay = el(A, ky) - Sy * el(A, kz);
    const float Bx = el(B, kx) - Sx * el(B, kz), By = el(B, ky) - Sy * el(B, kz);
    const float Cx = el(C, kx) - Sx * el(C, kz), Cy = el(C, ky) - Sy * el(C, kz);
    const float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0))
        return false;
    const float det = U + V + W;
    if(det == 0.0f)
        return false;
    const float T = U * (Sz * el(A, kz)) + V * (Sz * el(B, kz)) + W * (Sz * el(C, kz));
    const float r = 1.0f / det;
    const float tt = T * r;
    if(!(tt >= tmin && tt <= tmax))
        return false;
    t = tt;
    b1 = V * r;
    b2 = W * r;
    return true;
}

// lib/bvh Triangle::intersect, lib/bvh/include/bvh/triangle.hpp:95-115
static bool tri_mt(Vec o, Vec d, Vec p0, Vec p1, Vec p2, float tmin, float tmax, float &t, float &u, float &v)
{
    const Vec e1 = sub(p0, p1), e2 = sub(p2, p0), n = cross(e1, e2);
    const Vec c = sub(p0, o), r = cross(d, c);
    const float inv_det = 1.0f / dot(n, d);
    const float uu = dot(r, e2) * inv_det, vv = dot(r, e1) * inv_det, w = 1.0f - uu - vv;
    if(uu >= 0 && vv >= 0 && w >= 0)
    {
        const float tt = dot(n, c) * inv_det;
        if(tt >= tmin && tt <= tmax)
        {
            t = tt; u = uu; v = vv;
            return true;
        }
    }
    return false;
}

static inline bool test_tri(const Scene &s, int mode, uint32_t i, Vec o, Vec d, float tmin, Hit &h)
{
    float t, u, v;
    const bool ok = mode == BTRC_TRI_WATERTIGHT ? tri_watertight(o, d, s.w0[i], s.w1[i], s.w2[i], tmin, h.t, t, u, v)
                                                : tri_mt(o, d, s.w0[i], s.w1[i], s.w2[i], tmin, h.t, t, u, v);
    if(ok && (t < h.t || i < h.gid))
    {
        h.t = t; h.u = u; h.v = v; h.gid = i;
        return true;
    }
    return false;
}

// Ize's conservative slab test (JCGT 2013)
static bool box_hit(const Scene::Node &nd, Vec o, Vec inv, Vec invp, float tmin, float tmax)
{
    float te = tmin, tx = tmax;
    for(int a = 0; a < 3; ++a)
    {
        const float ia = el(inv, a);
        const float lo = nd.lo[a], hi = nd.hi[a];
        const bool neg = std::signbit(ia);
        const float nearp = neg ? hi : lo, farp = neg ? lo : hi;
        const float t0 = (nearp - el(o, a)) * ia;
        const float t1 = (farp - el(o, a)) * el(invp, a);
        te = t0 > te ? t0 : te;
        tx = t1 < tx ? t1 : tx;
    }
    return te <= tx;
}


// Box culling keeps every box entered within t (1 + 2^-20) + margin of the
// current hit: the slab entry and the triangle's t are rounded differently
// (the watertight t by up to ~2^-24 times the origin-to-vertex distance),
// and without the scene-scaled margin a triangle at the same t as the
// current hit can be culled, making the (t, index) minimum depend on visit
// order.
static inline float cull_margin(const Scene &s, Vec o)
{
    return (s.extent + fmx(std::fabs(o.x), fmx(std::fabs(o.y), std::fabs(o.z)))) * (1.0f / 262144.0f);
}
static inline float cull_dist(float t, float margin) { return t < FMAX ? t * 1.00000095f + margin : t; }

// world_to_local rows applied to a point / vector, in the operation order of
// the traversal kernel
static inline Vec xinv_point(const Xf &t, Vec p)
{
    return mk(t.n[0] * p.x + t.n[1] * p.y + t.n[2] * p.z + t.n[3],
              t.n[4] * p.x + t.n[5] * p.y + t.n[6] * p.z + t.n[7],
              t.n[8] * p.x + t.n[9] * p.y + t.n[10] * p.z + t.n[11]);
}
static inline Vec xinv_vec(const Xf &t, Vec v)
{
    return mk(t.n[0] * v.x + t.n[1] * v.y + t.n[2] * v.z,
              t.n[4] * v.x + t.n[5] * v.y + t.n[6] * v.z,
              t.n[8] * v.x + t.n[9] * v.y + t.n[10] * v.z);
}

// Two levels (Scene::commit's instance structure, core/scene.cpp:92-239):
// every instance's triangles are tested in its local space against the ray
// moved there; the closest hit is still the minimum of (t, flattened index).
static Hit trace_two_level(const Scene &s, int mode, bool any, bool use_bvh, Vec o, Vec d, float tmin, float tmax)
{
    Hit h{ tmax, 0.0f, 0.0f, 0xffffffffu };
    const Vec inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    auto pad = [](float x) { return std::isfinite(x) ? bitsf(fbits(x) + 2u) : x; };
    const Vec invp = mk(pad(inv.x), pad(inv.y), pad(inv.z));
    for(size_t i = 0; i < s.insts.size(); ++i)
    {
        const Inst &in = s.insts[i];
        const Geo &g = s.geos[in.geo];
        if(use_bvh)
        {
            Scene::Node nd{};
            for(int a = 0; a < 3; ++a)
            {
                nd.lo[a] = s.ibox[6 * i + a];
                nd.hi[a] = s.ibox[6 * i + 3 + a];
            }
            const float cull = cull_dist(h.t, cull_margin(s, o));
            if(!g.n || !box_hit(nd, o, inv, invp, tmin, cull))
                continue;
        }
        const Vec lo = xinv_point(in.xf, o), ld = xinv_vec(in.xf, d);
        for(uint32_t p = 0; p < g.n; ++p)
        {
            const uint32_t t = g.off + p, gid = s.ibase[i] + p;
            float tt, u, v;
            const bool ok = mode == BTRC_TRI_WATERTIGHT ? tri_watertight(lo, ld, s.l0[t], s.l1[t], s.l2[t], tmin, h.t, tt, u, v)
                                                        : tri_mt(lo, ld, s.l0[t], s.l1[t], s.l2[t], tmin, h.t, tt, u, v);
            if(ok && (tt < h.t || gid < h.gid))
            {
                h.t = tt; h.u = u; h.v = v; h.gid = gid;
                if(any)
                    return h;
            }
        }
    }
    return h;
}

static Hit trace(const Scene &s, int mode, bool any, bool use_bvh, Vec o, Vec d, float tmin, float tmax)
{
    if(s.two_level)
        return trace_two_level(s, mode, any, use_bvh, o, d, tmin, tmax);
    Hit h{ tmax, 0.0f, 0.0f, 0xffffffffu };
    const uint32_t n = (uint32_t)s.w0.size();
    if(!use_bvh || s.nodes.empty())
    {
        for(uint32_t i = 0; i < n; ++i)
            if(test_tri(s, mode, i, o, d, tmin, h) && any)
                return h;
        return h;
    }
    const Vec inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    auto pad = [](float x) { return std::isfinite(x) ? bitsf(fbits(x) + 2u) : x; };
    const Vec invp = mk(pad(inv.x), pad(inv.y), pad(inv.z));
    std::vector<int32_t> st{ 0 };
    while(!st.empty())
    {
        const Scene::Node &nd = s.nodes[st.back()];
        st.pop_back();
        const float cull = cull_dist(h.t, cull_margin(s, o));
        if(!box_hit(nd, o, inv, invp, tmin, cull))
            continue;
        if(nd.left < 0)
        {
            for(uint32_t k = nd.first; k < nd.first + nd.count; ++k)
                if(test_tri(s, mode, s.order[k], o, d, tmin, h) && any)
                    return h;
            continue;
        }
        st.push_back(nd.right);
        st.push_back(nd.left);
    }
    return h;
}

// ------------------------------------------------------------------ textures

struct F4 { float x, y, z, w; };

// texel index under an address mode (clamp to the edge, wrap modulo the
// size, mirror with period 2n, border: -1 outside)
static int tex_addr(int i, int n, int mode)
{
    if(mode == BTRC_ADDRESS_WRAP)
    {
        const int m = i % n;
        return m < 0 ? m + n : m;
    }
    if(mode == BTRC_ADDRESS_MIRROR)
    {
        int m = i % (2 * n);
        if(m < 0)
            m += 2 * n;
        return m < n ? m : 2 * n - 1 - m;
    }
    if(mode == BTRC_ADDRESS_BORDER)
        return (i < 0 || i >= n) ? -1 : i;
    return std::min(std::max(i, 0), n - 1);
}

static F4 tex_texel(const Scene &s, const Tex &t, int x, int y)
{
    if(x < 0 || y < 0)
        return F4{ t.border[0], t.border[1], t.border[2], t.border[3] };
    const size_t i = (size_t)y * t.w + x;
    if(t.format == BTRC_TEXTURE_RGBA32F)
        return F4{ t.t32[4 * i], t.t32[4 * i + 1], t.t32[4 * i + 2], t.t32[4 * i + 3] };
    const uint32_t p = t.t8[i];
    const float *lut = s.decode + (t.srgb ? 256 : 0);
    return F4{ lut[p & 255u], lut[(p >> 8) & 255u], lut[(p >> 16) & 255u], s.decode[p >> 24] };
}

static int tex_floor(float x, float &frac)
{
    const float f = std::floor(x);
    frac = x - f;
    return (int)std::fmin(std::fmax(f, -1073741824.0f), 1073741824.0f);
}

// a normalised-coordinate texture fetch with point or bilinear filtering
// (CUDA texture semantics restated with FP32 weights)
static F4 tex_sample(const Scene &s, int id, float u, float v)
{
    const Tex &t = s.texs[id];
    if(t.filter == BTRC_TEX_FILTER_POINT)
    {
        float fa, fb;
        const int x = tex_addr(tex_floor(u * (float)t.w, fa), t.w, t.au);
        const int y = tex_addr(tex_floor(v * (float)t.h, fb), t.h, t.av);
        return tex_texel(s, t, x, y);
    }
    float a, b;
    const int x0 = tex_floor(u * (float)t.w - 0.5f, a);
    const int y0 = tex_floor(v * (float)t.h - 0.5f, b);
    if(!(a >= 0.0f && a < 1.0f))
        a = 0.0f;
    if(!(b >= 0.0f && b < 1.0f))
        b = 0.0f;
    const int xa = tex_addr(x0, t.w, t.au), xb = tex_addr(x0 + 1, t.w, t.au);
    const int ya = tex_addr(y0, t.h, t.av), yb = tex_addr(y0 + 1, t.h, t.av);
    const F4 p = tex_texel(s, t, xa, ya), q = tex_texel(s, t, xb, ya);
    const F4 r = tex_texel(s, t, xa, yb), w = tex_texel(s, t, xb, yb);
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    return F4{ ((w00 * p.x + w10 * q.x) + w01 * r.x) + w11 * w.x, ((w00 * p.y + w10 * q.y) + w01 * r.y) + w11 * w.y,
               ((w00 * p.z + w10 * q.z) + w01 * r.z) + w11 * w.z, ((w00 * p.w + w10 * q.w) + w01 * r.w) + w11 * w.w };
}

// ------------------------------------------------------------------ IBL

static float phi_of(Vec w) // local_angle::phi, utils/local_angle.h:63-72
{
    float r = 0.0f;
    if(w.x != 0.0f || w.y != 0.0f)
    {
        r = dm::atan2(w.y, w.x);
        if(r < 0.0f)
            r = r + 2.0f * PI;
    }
    return r;
}

static float theta_of(Vec w) { return dm::acos(fmn(fmx(w.z, -1.0f), 1.0f)); } // local_angle.h:45-49

static float hammersley_ri(uint32_t bits) // utils/math/hammersley.h:7-16
{
    bits = (bits << 16u) | (bits >> 16u);
    bits = ((bits & 0x55555555u) << 1u) | ((bits & 0xAAAAAAAAu) >> 1u);
    bits = ((bits & 0x33333333u) << 2u) | ((bits & 0xCCCCCCCCu) >> 2u);
    bits = ((bits & 0x0F0F0F0Fu) << 4u) | ((bits & 0xF0F0F0F0u) >> 4u);
    bits = ((bits & 0x00FF00FFu) << 8u) | ((bits & 0xFF00FF00u) >> 8u);
    return (float)bits * 2.3283064365386963e-10f;
}

static float lerpf(float a, float b, float t) { return a * (1.0f - t) + b * t; } // utils/cmath/cscalar.h:41-44

// EnvirLightSampler::preprocess, light/env_sampler.cpp:16-122, with
// n_samples = 256 (IBL::commit, ibl.cpp:21-25)
static void ibl_table(const Scene &s, Light &l)
{
    const int w = l.lw, h = l.lh, n = w * h, ns = 256;
    std::vector<float> lm(n), area(n);
    for(int yi = 0; yi < h; ++yi)
        for(int xi = 0; xi < w; ++xi)
        {
            const float x0 = (float)xi / (float)w, x1 = (float)(xi + 1) / (float)w;
            const float y0 = (float)yi / (float)h, y1 = (float)(yi + 1) / (float)h;
            float sum = 0.0f;
            for(int k = 0; k < ns; ++k)
            {
                // hammersley2d(k, ns) stretched to [-0.1, 1.1)
                const float lx = 1.2f * ((float)k / (float)ns) - 0.1f;
                const float ly = 1.2f * hammersley_ri((uint32_t)k) - 0.1f;
                const F4 c = tex_sample(s, l.tex, lerpf(x0, x1, lx), lerpf(y0, y1, ly));
                sum = sum + lum(mk(c.x, c.y, c.z));
            }
            lm[yi * w + xi] = sum / (float)ns;
            area[yi * w + xi] = std::fabs(2.0f * PI * (x1 - x0) * (dm::cos(PI * y1) - dm::cos(PI * y0)));
        }
    float la = 0.0f, as = 0.0f;
    for(int i = 0; i < n; ++i)
    {
        la += lm[i] * area[i];
        as += area[i];
    }
    const float avg = la / as;
    for(int i = 0; i < n; ++i)
        lm[i] = std::max(0.0f, area[i] * (lm[i] - avg));
    float ts = 0.0f;
    for(float v : lm)
        ts += v;
    if(ts > 1e-3f)
    {
        const float ratio = 1 / ts;
        for(float &v : lm)
            v *= ratio;
    }
    alias_build(lm, l.accept, l.alias);
    l.probs = lm;
}

static Vec ibl_local_le(const Scene &s, const Light &l, Vec w) // IBL::eval_local, ibl.cpp:71-78
{
    const float u = fmn(fmx((1.0f / (2.0f * PI)) * phi_of(w), 0.0f), 1.0f);
    const float v = fmn(fmx((1.0f / PI) * theta_of(w), 0.0f), 1.0f);
    const F4 c = tex_sample(s, l.tex, u, v);
    return mk(c.x, c.y, c.z);
}

static int tile_of(float x, int n) // i32(floor(x)) with NaN -> 0, clamped to [0, n - 1]
{
    const float f = std::floor(x);
    if(!(f >= 0.0f))
        return 0;
    return f >= (float)(n - 1) ? n - 1 : (int)f;
}

static float ibl_pdf_local(const Light &l, Vec to) // EnvirLightSampler::pdf, env_sampler.cpp:161-187
{
    const Vec d = nrm(to);
    const float theta = dm::acos(d.z);
    const float u = phi_of(d) / (2.0f * PI), v = theta / PI;
    const int tx = tile_of(u * (float)l.lw, l.lw), ty = tile_of(v * (float)l.lh, l.lh);
    const float u0 = (float)tx / (float)l.lw, u1 = (float)(tx + 1) / (float)l.lw;
    const float v0 = (float)ty / (float)l.lh, v1 = (float)(ty + 1) / (float)l.lh;
    const float c0 = dm::cos(PI * v0), c1 = dm::cos(PI * v1);
    return l.probs[ty * l.lw + tx] * (1.0f / std::fabs(2.0f * PI * (u1 - u0) * (c1 - c0)));
}

// EnvirLightSampler::sample, env_sampler.cpp:124-159
static void ibl_sample_local(const Light &l, float s0, float s1, float s2, Vec &dir, float &pdf)
{
    const uint32_t n = (uint32_t)(l.lw * l.lh);
    const float nu = (float)n * s0;
    uint32_t k = (uint32_t)nu;
    k = k < n - 1 ? k : n - 1;
    const uint32_t tile = (nu - (float)k) <= l.accept[k] ? k : l.alias[k];
    const int ty = (int)tile / l.lw, tx = (int)tile % l.lw;
    const float u0 = (float)tx / (float)l.lw, u1 = (float)(tx + 1) / (float)l.lw;
    const float v0 = (float)ty / (float)l.lh, v1 = (float)(ty + 1) / (float)l.lh;
    float cv0 = dm::cos(PI * v0), cv1 = dm::cos(PI * v1);
    if(cv0 > cv1)
        std::swap(cv0, cv1);
    const float ct = cv0 + s2 * (cv1 - cv0);
    const float st = std::sqrt(fmx(0.0f, 1.0f - ct * ct));
    const float phi = 2.0f * PI * lerpf(u0, u1, s1);
    dir = mk(st * dm::cos(phi), st * dm::sin(phi), ct);
    pdf = l.probs[tile] * (1.0f / (2.0f * PI * (u1 - u0) * (cv1 - cv0)));
}

// ------------------------------------------------------------------ surfaces

struct SP
{
    Vec pos;
    Frm frame;
    Vec iz;
    float tu, tv; // texture coordinate
};

// the material at a surface point: every textured parameter sampled at its
// texture coordinate, colour as RGB, the others from the first channel
// (disney.cpp:813-842)
static btrc_material_desc material_at(const Scene &s, int id, const SP &sp)
{
    btrc_material_desc m = s.mats[id];
    if(m.color_texture > 0)
    {
        const F4 c = tex_sample(s, m.color_texture - 1, sp.tu, sp.tv);
        m.color[0] = c.x;
        m.color[1] = c.y;
        m.color[2] = c.z;
    }
    const struct { int32_t tex; float *value; } scalars[] = {
        { m.metallic_texture, &m.metallic }, { m.roughness_texture, &m.roughness },
        { m.specular_texture, &m.specular }, { m.specular_tint_texture, &m.specular_tint },
        { m.anisotropic_texture, &m.anisotropic }, { m.sheen_texture, &m.sheen },
        { m.sheen_tint_texture, &m.sheen_tint }, { m.clearcoat_texture, &m.clearcoat },
        { m.clearcoat_gloss_texture, &m.clearcoat_gloss }, { m.transmission_texture, &m.transmission },
        { m.transmission_roughness_texture, &m.transmission_roughness }, { m.ior_texture, &m.ior },
    };
    for(const auto &p : scalars)
        if(p.tex > 0)
            *p.value = tex_sample(s, p.tex - 1, sp.tu, sp.tv).x;
    return m;
}

// NormalMap::adjust_frame (material/utils/normal_map.cpp:13-21): the map's
// RGB as 2 rgb - 1 in the shading frame is the new shading z
static Frm normal_mapped(const Scene &s, const btrc_material_desc &m, const SP &sp, const Frm &shading)
{
    if(m.normal_texture <= 0 || m.type == BTRC_MATERIAL_INVISIBLE)
        return shading;
    const F4 c = tex_sample(s, m.normal_texture - 1, sp.tu, sp.tv);
    const Vec local = mk(2.0f * c.x - 1.0f, 2.0f * c.y - 1.0f, 2.0f * c.z - 1.0f);
    return rot_z(shading, nrm(l2g(shading, local)));
}

// get_hitinfo, core/scene.cpp:9-64
static SP hitinfo(const Scene &s, Vec o, Vec d, const Inst &in, uint32_t prim, float t, float u, float v)
{
    const uint32_t i = s.geos[in.geo].off + prim;
    SP sp;
    sp.pos = add(o, sclf(t, d));
    const Frm gf = frm(s.gx[i], s.gy[i], s.gz[i]);
    sp.frame.x = nrm(xvec(in.xf, gf.x));
    sp.frame.y = nrm(xvec(in.xf, gf.y));
    sp.frame.z = nrm(xnor(in.xf, gf.z));
    const Vec n = add(add(s.sa[i], scl(s.sba[i], u)), scl(s.sca[i], v));
    sp.iz = nrm(xnor(in.xf, n));
    const float *c = &s.tuv[6 * (size_t)i];
    sp.tu = (c[0] + c[1] * u) + c[2] * v;
    sp.tv = (c[3] + c[4] * u) + c[5] * v;
    return sp;
}

// ---------------------------------------------------------- BSDF components

static float cos2sin(float c) { return std::sqrt(fmx(0.0f, 1.0f - c * c)); } // utils/local_angle.h:22-25
static float tan_th(Vec w) // local_angle.h:28-36
{
    const float t = 1.0f - w.z * w.z;
    return t > 0.0f ? std::sqrt(t) / w.z : 0.0f;
}
// sin and cos of local_angle::phi (local_angle.h:55-76) from x / r and y / r
static void phi_sc(Vec w, float &s, float &c)
{
    if(w.x != 0.0f || w.y != 0.0f)
    {
        const float r = std::sqrt(w.x * w.x + w.y * w.y);
        s = w.y / r;
        c = w.x / r;
    }
    else
    {
        s = 0.0f;
        c = 1.0f;
    }
}

// material/utils/fresnel.cpp:5-33
static float fresnel_d(float ei, float eo, float ci)
{
    if(ci < 0.0f) { std::swap(ei, eo); ci = -ci; }
    const float si = std::sqrt(fmx(0.0f, 1.0f - ci * ci));
    const float st = eo / ei * si;
    if(st >= 1.0f)
        return 1.0f;
    const float ct = std::sqrt(fmx(0.0f, 1.0f - st * st));
    const float pa = (ei * ci - eo * ct) / (ei * ci + eo * ct);
    const float pe = (eo * ci - ei * ct) / (eo * ci + ei * ct);
    return 0.5f * (pa * pa + pe * pe);
}
static float schlick(float r0, float c) // fresnel.cpp:35-41
{
    const float t = 1.0f - c, t2 = t * t, t5 = t2 * t2 * t;
    return r0 + (1.0f - r0) * t5;
}
static Vec schlick3(Vec r0, float c) // fresnel.cpp:43-49
{
    const float t = 1.0f - c, t2 = t * t, t5 = t2 * t2 * t;
    return add(r0, scl(sub(mk(1, 1, 1), r0), t5));
}

// material/utils/microfacet.cpp
static float ca_(float sp, float cp, float ax, float ay) { return sq(cp / ax) + sq(sp / ay); }
static float smith_gtr2(float tn, float a)
{
    if(tn == 0.0f)
        return 1.0f;
    const float r = a * tn;
    return 2.0f / (1.0f + std::sqrt(1.0f + r * r));
}
static float aniso_gtr2(float sp, float cp, float st, float ct, float ax, float ay)
{
    const float A = ca_(sp, cp, ax, ay);
    const float RD = sq(st) * A + sq(ct);
    return 1.0f / (PI * ax * ay * sq(RD));
}
static float smith_aniso(float cp, float sp, float ax, float ay, float tn)
{
    const float t = sq(ax * cp) + sq(ay * sp);
    const float lam = -0.5f + 0.5f * std::sqrt(1.0f + t * sq(tn));
    return 1.0f / (1.0f + lam);
}
static Vec sample_aniso_gtr2(float ax, float ay, float s0, float s1)
{
    const float ang = 2.0f * PI * s0;
    float sp = ay * dm::sin(ang), cp = ax * dm::cos(ang);
    const float nr = 1.0f / std::sqrt(sq(sp) + sq(cp));
    sp = sp * nr;
    cp = cp * nr;
    const float A = ca_(sp, cp, ax, ay);
    const float ct = std::sqrt(A * (1.0f - s1) / ((1.0f - A) * s1 + A));
    const float st = cos2sin(ct);
    return nrm(mk(st * cp, st * sp, ct));
}
static float gtr1(float st, float ct, float a)
{
    return (sq(a) - 1.0f) / (2.0f * PI * dm::log(a) * (sq(a * ct) + sq(st)));
}
static Vec sample_gtr1(float a, float s0, float s1)
{
    const float phi = 2.0f * PI * s0;
    const float ct = std::sqrt((dm::pow(a, 2.0f - 2.0f * s1) - 1.0f) / (sq(a) - 1.0f));
    const float st = cos2sin(ct);
    return nrm(mk(st * dm::cos(phi), st * dm::sin(phi), ct));
}
static Vec sample_vndf(Vec ve, float ax, float ay, float s0, float s1)
{
    const Vec vh = nrm(mk(ax * ve.x, ay * ve.y, ve.z));
    const float lsq = vh.x * vh.x + vh.y * vh.y;
    const Vec t1 = lsq > 3e-4f ? dvs(mk(-vh.y, vh.x, 0.0f), std::sqrt(lsq)) : mk(1, 0, 0);
    const Vec t2 = cross(vh, t1);
    const float r = std::sqrt(s0);
    const float phi = 2.0f * PI * s1;
    const float a1 = r * dm::cos(phi), a2r = r * dm::sin(phi);
    const float s = 0.5f * (1.0f + vh.z);
    const float a2 = (1.0f - s) * std::sqrt(1.0f - a1 * a1) + s * a2r;
    const Vec nh = add(add(sclf(a1, t1), sclf(a2, t2)), sclf(std::sqrt(fmx(0.0f, 1.0f - a1 * a1 - a2 * a2)), vh));
    return nrm(mk(ax * nh.x, ay * nh.y, fmx(0.0f, nh.z)));
}

struct BS
{
    Vec f, dir;
    float pdf;
    bool delta = false;
};

struct Dis
{
    Vec C, Ct;
    float met, rough, spec, stint, sheen, shtint, cc, ccr, trans, ior, tax, tay, ax, ay, wd, ws, wc, wt;
};

// DisneyBSDF::DisneyBSDF, builtin/material/disney.cpp:434-475
static Dis dis_make(const btrc_material_desc &m)
{
    Dis d;
    d.C = mk(m.color[0], m.color[1], m.color[2]);
    const float l = lum(d.C);
    d.Ct = l > 0.0f ? dvs(d.C, l) : mk(1, 1, 1);
    d.met = m.metallic;
    d.rough = m.roughness;
    d.spec = m.specular;
    d.stint = m.specular_tint;
    d.sheen = m.sheen;
    d.shtint = m.sheen_tint;
    d.trans = m.transmission;
    d.ior = fmx(1.01f, m.ior);
    const float aspect = m.anisotropic > 0.0f ? std::sqrt(1.0f - 0.9f * m.anisotropic) : 1.0f;
    d.ax = fmx(0.001f, sq(m.roughness) / aspect);
    d.ay = fmx(0.001f, sq(m.roughness) * aspect);
    d.tax = fmx(0.001f, sq(m.transmission_roughness) / aspect);
    d.tay = fmx(0.001f, sq(m.transmission_roughness) * aspect);
    d.cc = m.clearcoat;
    d.ccr = fmx(0.0001f, sq(0.1f * (1.0f - m.clearcoat_gloss) + 0.0f * m.clearcoat_gloss));
    const float A = fmn(fmx(l * (1.0f - m.metallic), 0.3f), 0.7f);
    const float B = 1.0f - A;
    d.wd = A * (1.0f - m.transmission);
    d.wt = A * m.transmission;
    d.ws = B * 2.0f / (2.0f + m.clearcoat);
    d.wc = B * m.clearcoat / (2.0f + m.clearcoat);
    return d;
}

static float om5(float x) { const float t = 1.0f - x, t2 = t * t; return t2 * t2 * t; }

// D (returned through D) and G of an anisotropic GTR2 lobe
static float dg(Vec h, Vec i, Vec o, float ax, float ay, float &D)
{
    float sph, cph, spi, cpi, spo, cpo;
    phi_sc(h, sph, cph);
    D = aniso_gtr2(sph, cph, cos2sin(h.z), h.z, ax, ay);
    phi_sc(i, spi, cpi);
    phi_sc(o, spo, cpo);
    return smith_aniso(cpi, spi, ax, ay, tan_th(i)) * smith_aniso(cpo, spo, ax, ay, tan_th(o));
}

static Vec dis_eval(const Dis &d, Vec i, Vec o) // disney.cpp:477-558
{
    if(std::fabs(i.z) < 1e-4f || std::fabs(o.z) < 1e-4f)
        return mk(0, 0, 0);
    if(i.z * o.z < 0.0f)
    {
        // eval_trans, disney.cpp:100-145
        if(d.trans == 0.0f)
            return mk(0, 0, 0);
        const float eta = o.z > 0.0f ? d.ior : 1.0f / d.ior;
        Vec h = nrm(add(o, sclf(eta, i)));
        if(h.z < 0.0f)
            h = neg(h);
        const float cd = dot(o, h);
        const float F = fresnel_d(d.ior, 1.0f, cd);
        float D;
        const float G = dg(h, i, o, d.tax, d.tay, D);
        const float sdem = cd + eta * dot(i, h);
        const float corr = 1.0f / eta;
        const Vec sC = mk(std::sqrt(d.C.x), std::sqrt(d.C.y), std::sqrt(d.C.z));
        const float val = (1.0f - F) * D * G * eta * eta * dot(i, h) * dot(o, h) * corr * corr
                        / (i.z * o.z * sdem * sdem);
        const float tf = o.z > 0.0f ? d.trans : 1.0f;
        return scl(sclf((1.0f - d.met) * tf, sC), std::fabs(val));
    }
    if(i.z < 0.0f && o.z < 0.0f)
    {
        // eval_inner_refl, disney.cpp:147-180
        if(d.trans == 0.0f)
            return mk(0, 0, 0);
        const Vec h = neg(nrm(add(i, o)));
        const float F = fresnel_d(d.ior, 1.0f, dot(o, h));
        float D;
        const float G = dg(h, i, o, d.tax, d.tay, D);
        return scl(sclf(d.trans, d.C), std::fabs(F * D * G / (4.0f * i.z * o.z)));
    }
    if(i.z <= 0.0f || o.z <= 0.0f)
        return mk(0, 0, 0);
    const Vec h = nrm(add(i, o));
    const float cd = dot(i, h);
    Vec diffuse = mk(0, 0, 0), sheen = mk(0, 0, 0);
    if(d.met < 1.0f)
    {
        // eval_diffuse, disney.cpp:76-86
        const Vec lam = dvs(d.C, PI);
        const float fl = om5(i.z), fv = om5(o.z);
        const float rr = 2.0f * d.rough * cd * cd;
        const Vec retro = scl(scl(dvs(d.C, PI), rr), fl + fv + fl * fv * (rr - 1.0f));
        diffuse = add(scl(scl(lam, 1.0f - 0.5f * fl), 1.0f - 0.5f * fv), retro);
        if(d.sheen > 0.0f) // eval_sheen, disney.cpp:88-92
            sheen = scl(sclf(4.0f * d.sheen, lerpv(mk(1, 1, 1), d.Ct, d.shtint)), om5(cd));
    }
    // eval_specular, disney.cpp:182-220
    Vec spec;
    {
        const Vec Cs = lerpv(lerpv(mk(1, 1, 1), d.Ct, d.stint), d.C, d.met);
        const Vec Fd = scl(Cs, fresnel_d(d.ior, 1.0f, cd));
        const Vec Fc = schlick3(Cs, cd);
        const Vec F = lerpv(sclf(d.spec, Fd), Fc, d.met);
        float D;
        const float G = dg(h, i, o, d.ax, d.ay, D);
        spec = dvs(scl(scl(F, D), G), std::fabs(4.0f * i.z * o.z));
    }
    Vec cc = mk(0, 0, 0);
    if(d.cc > 0.0f)
    {
        // eval_clearcoat, disney.cpp:94-98
        const float Dc = gtr1(cos2sin(h.z), h.z, d.ccr);
        const float Fc = schlick(0.04f, cd);
        const float Gc = smith_gtr2(tan_th(i), 0.25f) * smith_gtr2(tan_th(o), 0.25f);
        const float v = d.cc * Dc * Fc * Gc / std::fabs(4.0f * i.z * o.z);
        cc = mk(v, v, v);
    }
    return add(add(sclf((1.0f - d.met) * (1.0f - d.trans), add(diffuse, sheen)), spec), cc);
}

static float pdf_trans(const Dis &d, Vec i, Vec o) // disney.cpp:369-402
{
    const float eta = o.z > 0.0f ? d.ior : 1.0f / d.ior;
    Vec h = nrm(add(o, sclf(eta, i)));
    if(h.z < 0.0f)
        h = neg(h);
    if(((o.z > 0.0f) != (dot(h, o) > 0.0f)) || ((i.z > 0.0f) != (dot(h, i) > 0.0f)))
        return 0.0f;
    const float sdem = dot(o, h) + eta * dot(i, h);
    const float j = eta * eta * dot(i, h) / (sdem * sdem);
    float sph, cph;
    phi_sc(h, sph, cph);
    const float D = aniso_gtr2(sph, cph, cos2sin(h.z), h.z, d.tax, d.tay);
    return std::fabs(dot(i, h) * D * j);
}

static float pdf_inner(const Dis &d, Vec i, Vec o) // disney.cpp:404-420
{
    const Vec h = neg(nrm(add(i, o)));
    float sph, cph;
    phi_sc(h, sph, cph);
    const float D = aniso_gtr2(sph, cph, cos2sin(h.z), h.z, d.tax, d.tay);
    return std::fabs(h.z * D / (4.0f * dot(i, h)));
}

static float dis_pdf(const Dis &d, Vec i, Vec o) // disney.cpp:682-730
{
    if(std::fabs(i.z) < 1e-4f || std::fabs(o.z) < 1e-4f)
        return 0.0f;
    if(o.z < 0.0f)
    {
        if(d.trans == 0.0f)
            return 0.0f;
        const float mF = fmn(fmx(fresnel_d(d.ior, 1.0f, o.z), 0.1f), 0.9f);
        if(i.z > 0.0f)
            return (1.0f - mF) * pdf_trans(d, i, o);
        return mF * pdf_inner(d, i, o);
    }
    if(i.z < 0.0f)
        return d.wt * pdf_trans(d, i, o);
    const float pd = hemisphere_cos_pdf(i);
    // pdf_specular_clearcoat, disney.cpp:339-367
    const Vec h = nrm(add(i, o));
    float sph, cph, spo_unused, cpo;
    phi_sc(h, sph, cph);
    phi_sc(o, spo_unused, cpo);
    const float spo = cos2sin(cpo);
    const float Ds = aniso_gtr2(sph, cph, cos2sin(h.z), h.z, d.ax, d.ay);
    const float ps = smith_aniso(cpo, spo, d.ax, d.ay, tan_th(o)) * Ds / (4.0f * o.z);
    const float pc = h.z * gtr1(cos2sin(h.z), h.z, d.ccr) / (4.0f * dot(i, h));
    return d.wd * pd + d.ws * ps + d.wc * pc;
}

static Vec refl(Vec o, Vec h) { return sub(sclf(2.0f * dot(o, h), h), o); }

static Vec samp_trans(const Dis &d, Vec o, float s0, float s1) // disney.cpp:277-308
{
    const Vec h = sample_aniso_gtr2(d.tax, d.tay, s0, s1);
    if(h.z <= 0.0f || (o.z > 0.0f) != (dot(h, o) > 0.0f))
        return mk(0, 0, 0);
    const float eta = o.z > 0.0f ? 1.0f / d.ior : d.ior;
    const Vec oh = dot(h, o) > 0.0f ? h : neg(h);
    // refract, disney.cpp:55-74
    const float ci = dot(o, oh);
    const float s2t = sq(eta) * (1.0f - sq(ci));
    const float c2t = 1.0f - s2t;
    if(c2t <= 0.0f)
        return mk(0, 0, 0);
    const float ct = std::sqrt(c2t);
    const Vec i = nrm(sub(sclf(eta * ci - ct, oh), sclf(eta, o)));
    if(i.z * o.z > 0.0f || ((i.z > 0.0f) != (dot(h, i) > 0.0f)))
        return mk(0, 0, 0);
    return nrm(i);
}

static BS dis_sample(const Dis &d, Vec o, float s0, float s1, float s2) // disney.cpp:560-658
{
    BS r{ mk(0, 0, 0), mk(0, 0, 0), 0.0f };
    if(std::fabs(o.z) < 1e-4f)
        return r;
    Vec i = mk(0, 0, 0);
    if(o.z < 0.0f)
    {
        if(d.trans == 0.0f)
            return r;
        const float mF = fmn(fmx(fresnel_d(d.ior, 1.0f, o.z), 0.1f), 0.9f);
        if(s0 >= mF)
            i = samp_trans(d, o, s1, s2);
        else
        {
            // sample_inner_refl, disney.cpp:310-331
            const Vec h = sample_aniso_gtr2(d.tax, d.tay, s1, s2);
            if(h.z > 0.0f)
            {
                const Vec t = refl(o, h);
                if(!(t.z > 0.0f))
                    i = nrm(t);
            }
        }
    }
    else
    {
        float sel = s0;
        if(sel < d.wd)
            i = hemisphere_cos(s1, s2);
        else if((sel = sel - d.wd) < d.wt)
            i = samp_trans(d, o, s1, s2);
        else if((sel = sel - d.wt) < d.ws)
        {
            const Vec h = sample_vndf(o, d.ax, d.ay, s1, s2); // sample_specular, disney.cpp:227-252
            if(h.z > 0.0f)
            {
                const Vec t = refl(o, h);
                if(t.z > 0.0f)
                    i = nrm(t);
            }
        }
        else
        {
            const Vec h = sample_gtr1(d.ccr, s1, s2); // sample_clearcoat, disney.cpp:254-275
            if(h.z > 0.0f)
            {
                const Vec t = refl(o, h);
                if(t.z > 0.0f)
                    i = nrm(t);
            }
        }
    }
    if(zero(i))
        return r;
    r.dir = i;
    r.f = dis_eval(d, i, o);
    r.pdf = dis_pdf(d, i, o);
    const Vec v = dvs(scl(r.f, std::fabs(r.dir.z)), r.pdf); // handle_bad_sample, disney.cpp:566-573
    if(fmx(fmx(v.x, v.y), v.z) > 10.0f)
        r = BS{ mk(0, 0, 0), mk(0, 0, 0), 0.0f };
    return r;
}

// Lambert component, builtin/material/component/diffuse.cpp
static BS lam_sample(Vec albedo, Vec o, float s0, float s1)
{
    if(o.z <= 0.0f)
        return BS{ mk(0, 0, 0), mk(0, 0, 0), 0.0f };
    const Vec i = hemisphere_cos(s0, s1);
    return BS{ dvs(albedo, PI), i, hemisphere_cos_pdf(i) };
}

// BSDFAggregate with one component of weight 1 plus ShaderFrame
// (material/utils/aggregate.cpp:32-229, shader_frame.cpp)
struct Shd
{
    Frm geo, sh;
    bool st;
    bool disney;
    bool invisible = false;
    Vec albedo;
    Dis d;
};

static bool upper(const Frm &f, Vec w) { return dot(w, f.z) > 0.0f; }
static bool fringe(const Shd &s, Vec w) { return upper(s.geo, w) != upper(s.sh, w); }
static Frm flip(const Shd &s, Vec wo) { return fringe(s, wo) ? s.geo : s.sh; }
static float corr(const Shd &s, const Frm &sh, Vec wi)
{
    const float down = std::fabs(dot(s.geo.z, wi));
    return down > 1e-4f ? std::fabs(dot(sh.z, wi)) / down : 1.0f;
}
static float term(const Shd &s, Vec wi)
{
    if(!s.st)
        return 1.0f;
    const float g = fmn(fmx(dot(s.geo.z, wi) / (dot(s.sh.z, wi) * dot(s.geo.z, s.sh.z)), 0.0f), 1.0f);
    return -g * g * g + g * g + g;
}

static BS shd_sample(const Shd &s, Vec wo, float s0, float s1, float s2)
{
    if(s.invisible)
    {
        // InvisibleShader::sample_bidir, material/invisible.cpp:18-27
        BS r{ dvs(mk(1, 1, 1), std::fabs(dot(s.geo.z, nrm(wo)))), neg(wo), 1.0f, true };
        return r;
    }
    const Frm sh = flip(s, wo);
    const Vec lo = nrm(g2l(sh, wo));
    const BS c = s.disney ? dis_sample(s.d, lo, s0, s1, s2) : lam_sample(s.albedo, lo, s0, s1);
    BS r{ mk(0, 0, 0), mk(0, 0, 0), 0.0f };
    if(c.pdf <= 0.0f)
        return r;
    const Vec dir = l2g(sh, c.dir);
    if(fringe(s, dir))
        return r;
    r.dir = dir;
    r.pdf = c.pdf;
    r.f = scl(scl(c.f, corr(s, sh, dir)), term(s, dir));
    return r;
}

static Vec shd_eval(const Shd &s, Vec wi, Vec wo)
{
    if(s.invisible)
        return mk(0, 0, 0);
    const Frm sh = flip(s, wo);
    const Vec li = nrm(g2l(sh, wi)), lo = nrm(g2l(sh, wo));
    Vec f;
    if(s.disney)
        f = dis_eval(s.d, li, lo);
    else
        f = (li.z <= 0.0f || lo.z <= 0.0f) ? mk(0, 0, 0) : dvs(s.albedo, PI);
    return scl(scl(f, corr(s, sh, wi)), term(s, wi));
}

static float shd_pdf(const Shd &s, Vec wi, Vec wo)
{
    if(s.invisible)
        return 0.0f;
    const Frm sh = flip(s, wo);
    const Vec li = nrm(g2l(sh, wi)), lo = nrm(g2l(sh, wo));
    if(s.disney)
        return dis_pdf(s.d, li, lo);
    return (li.z <= 0.0f || lo.z <= 0.0f) ? 0.0f : hemisphere_cos_pdf(li);
}

// ------------------------------------------------------------------ lights

static Vec sky_le(const Light &l, Vec to) // gradient_sky.cpp:20-25
{
    const float c = dot(l.up, nrm(to));
    const float s = fmn(fmx(0.5f * (c + 1.0f), 0.0f), 1.0f);
    return add(scl(l.lower, 1.0f - s), scl(l.upper, s));
}

// EnvirLight::eval_le / pdf_li / sample_li: GradientSky (gradient_sky.cpp:20-40)
// or IBL (ibl.cpp:27-49)
static Vec env_le(const Scene &s, const Light &l, Vec to)
{
    return l.type == 2 ? ibl_local_le(s, l, nrm(g2l(l.fr, to))) : sky_le(l, to);
}

static float env_pdf(const Light &l, Vec to)
{
    return l.type == 2 ? ibl_pdf_local(l, g2l(l.fr, to)) : 1.0f / (4.0f * PI);
}

static void env_sample(const Scene &s, const Light &l, float s0, float s1, float s2, Vec &dir, Vec &li, float &pdf)
{
    if(l.type == 2)
    {
        Vec local;
        ibl_sample_local(l, s0, s1, s2, local, pdf);
        dir = l2g(l.fr, local);
        li = ibl_local_le(s, l, local);
        return;
    }
    dir = sphere_uniform(s0, s1);
    li = sky_le(l, dir);
    pdf = 1.0f / (4.0f * PI);
}

struct Shadow
{
    bool ok;
    Vec o, d, li;
    float t1, pdf;
};

// sample_light_li_impl + sample_surface_light_li, renderer/wavefront/helper.cpp:10-65,127-142;
// UniformLightSampler::sample, core/light_sampler.cpp:41-58;
// MeshLight::sample_li, light/mesh_light.cpp:38-58; TriangleMesh::sample_inline,
// geometry/triangle_mesh.cpp:188-228; CAliasTable::sample, utils/cmath/calias.cpp:15-21
static Shadow sample_light(const Scene &s, Vec ref, Vec nor, bool offset, float l0, float l1, float l2, float lsel)
{
    Shadow r{};
    const int n = (int)s.lights.size();
    if(n == 0)
        return r;
    int idx = (int)(lsel * (float)n);
    idx = idx < n - 1 ? idx : n - 1;
    const float sel_pdf = 1.0f / (float)n;
    const Light &l = s.lights[idx];
    r.ok = true;
    if(l.type == 0)
    {
        const Geo &g = s.geos[l.geo];
        const float nu = (float)g.n * l0;
        uint32_t k = (uint32_t)nu;
        k = k < g.n - 1 ? k : g.n - 1;
        const uint32_t prim = (nu - (float)k) <= g.accept[k] ? k : g.alias[k];
        const float t = std::sqrt(l1);
        const float b1 = 1.0f - t, b2 = t * l2;
        const uint32_t i = g.off + prim;
        const Vec lp = add(add(s.la[i], scl(s.lba[i], b1)), scl(s.lca[i], b2));
        const Inst &in = s.insts[l.inst];
        const Vec spos = xpoint(in.xf, lp);
        const Vec snor = nrm(xnor(in.xf, nrm(s.gz[i])));
        const Vec p2r = sub(ref, spos);
        const float dist2 = dot(p2r, p2r);
        const float pdf = 1.0f / (l.scale * l.scale) * (1.0f / g.total_area) * dist2 / std::fabs(dot(snor, nrm(p2r)));
        const Vec rad = dot(p2r, snor) > 0.0f ? l.intensity : mk(0, 0, 0);
        const Vec diff = sub(spos, ref);
        const Vec dst = offset_dir(spos, snor, neg(diff));
        r.o = offset ? offset_dir(ref, nor, diff) : ref; // helper.cpp:115-142
        r.d = sub(dst, r.o);
        r.t1 = 1.0f;
        r.li = rad;
        r.pdf = sel_pdf * pdf;
    }
    else
    {
        Vec dir, li;
        float pdf;
        env_sample(s, l, l0, l1, l2, dir, li, pdf);
        r.o = offset ? offset_dir(ref, nor, dir) : ref;
        r.d = dir;
        r.t1 = FMAX;
        r.li = li;
        r.pdf = sel_pdf * pdf;
    }
    return r;
}

// ------------------------------------------------------------------ media

// henyey_greenstein, utils/henyey_greenstein.h:8-16
static float hg(float g, float u)
{
    const float g2 = g * g;
    const float dem = 1.0f + g2 - 2.0f * (g * u);
    return (1.0f - g2) / (4.0f * PI * dem * std::sqrt(dem));
}

// HenyeyGreensteinPhaseShader::sample, core/medium.cpp:36-66; returns the
// phase value (colour times the phase function)
static Vec hg_sample(float g, Vec color, Vec wo, float s0, float s1, Vec &dir, float &pdf)
{
    const float sv = s0 + s0 - 1.0f;
    float u = sv;
    if(!(std::fabs(g) < 0.001f))
    {
        const float g2 = g * g;
        const float q = (1.0f - g2) / (1.0f + g * sv);
        u = (1.0f + g2 - q * q) / (g + g);
    }
    const float ct = -u;
    const float st = std::sqrt(fmx(0.0f, 1.0f - ct * ct));
    const float phi = 2.0f * PI * s1;
    const Vec local = mk(st * dm::sin(phi), st * dm::cos(phi), ct);
    const Frm f = from_z(wo);
    dir = add(add(sclf(local.x, f.x), sclf(local.y, f.y)), sclf(local.z, f.z));
    pdf = hg(g, u);
    return scl(color, pdf);
}

// linear filtering with clamp addressing at normalised coordinates: texel i
// sits at (i + 0.5) / n; lerp along x, then y, then z
static float tex3(const std::vector<float> &v, int nx, int ny, int nz, int ch, int c, Vec uvw)
{
    const float q[3] = { uvw.x * (float)nx - 0.5f, uvw.y * (float)ny - 0.5f, uvw.z * (float)nz - 0.5f };
    const int n[3] = { nx, ny, nz };
    int i0[3], i1[3];
    float t[3];
    for(int a = 0; a < 3; ++a)
    {
        float x = q[a];
        if(!(x >= -1.0f)) x = -1.0f;
        if(x > (float)n[a]) x = (float)n[a];
        const float fl = std::floor(x);
        t[a] = x - fl;
        const int i = (int)fl;
        i0[a] = std::min(std::max(i, 0), n[a] - 1);
        i1[a] = std::min(std::max(i + 1, 0), n[a] - 1);
    }
    auto at = [&](int x, int y, int z) { return v[(((size_t)z * ny + y) * nx + x) * ch + c]; };
    auto lerp = [](float a, float b, float w) { return a * (1.0f - w) + b * w; };
    const float c00 = lerp(at(i0[0], i0[1], i0[2]), at(i1[0], i0[1], i0[2]), t[0]);
    const float c10 = lerp(at(i0[0], i1[1], i0[2]), at(i1[0], i1[1], i0[2]), t[0]);
    const float c01 = lerp(at(i0[0], i0[1], i1[2]), at(i1[0], i0[1], i1[2]), t[0]);
    const float c11 = lerp(at(i0[0], i1[1], i1[2]), at(i1[0], i1[1], i1[2]), t[0]);
    return lerp(lerp(c00, c10, t[1]), lerp(c01, c11, t[1]), t[2]);
}

static Vec med_uvw(const Med &m, Vec p) // TransformMedium, core/medium.cpp:98-129
{
    if(!m.has_xf)
        return p;
    const float *w = m.w2l;
    return mk(w[0] * p.x + w[1] * p.y + w[2] * p.z + w[3], w[4] * p.x + w[5] * p.y + w[6] * p.z + w[7],
              w[8] * p.x + w[9] * p.y + w[10] * p.z + w[11]);
}

const int MAX_TRACKING = 1 << 20;

// work statistics of the medium stage for the benchmark's work model:
// delta-tracking segments and density lookups (each a trilinear fetch of
// eight voxels)
static std::atomic<uint64_t> g_track_segments{ 0 }, g_track_lookups{ 0 };

// Medium::sample along a -> b: HomogeneousMedium (homogeneous.cpp:25-58),
// delta tracking of HetergeneousMedium (hetergeneous.cpp:25-72)
template<class S>
static bool med_sample(const Med &m, Vec a, Vec b, S &smp, Vec &pos, Vec &color)
{
    const float tmax = len(sub(b, a));
    if(m.type == BTRC_MEDIUM_HOMOGENEOUS)
    {
        const float st = -dm::log(1.0f - smp.get1d()) / m.sigma_t;
        if(!(st < tmax))
            return false;
        const float t = st / tmax;
        pos = add(scl(a, 1.0f - t), scl(b, t));
        color = m.albedo;
        return true;
    }
    const Vec la = add(mk(0.5f, 0.5f, 0.5f), sclf(0.5f, med_uvw(m, a)));
    const Vec lb = add(mk(0.5f, 0.5f, 0.5f), sclf(0.5f, med_uvw(m, b)));
    const Vec lstep = dvs(sub(lb, la), tmax), wstep = dvs(sub(b, a), tmax);
    float t = 0.0f;
    g_track_segments.fetch_add(1, std::memory_order_relaxed);
    for(int it = 0; it < MAX_TRACKING; ++it)
    {
        t = t + -dm::log(1.0f - smp.get1d()) * m.inv_maj;
        if(t >= tmax)
            return false;
        const Vec uvw = add(la, sclf(t, lstep));
        g_track_lookups.fetch_add(1, std::memory_order_relaxed);
        const float dens = tex3(m.dens, m.nx, m.ny, m.nz, 1, 0, uvw);
        if(smp.get1d() < dens * m.inv_maj)
        {
            pos = add(a, sclf(t, wstep));
            color = m.alb.empty() ? m.albedo
                                  : mk(tex3(m.alb, m.nx, m.ny, m.nz, 3, 0, uvw), tex3(m.alb, m.nx, m.ny, m.nz, 3, 1, uvw),
                                       tex3(m.alb, m.nx, m.ny, m.nz, 3, 2, uvw));
            return true;
        }
    }
    return false;
}

// Medium::tr along a -> b: sigma_a-only closed form (homogeneous.cpp:60-73),
// ratio tracking (hetergeneous.cpp:74-106)
template<class S>
static Vec med_tr(const Med &m, Vec a, Vec b, S &smp)
{
    if(m.type == BTRC_MEDIUM_HOMOGENEOUS)
    {
        const Vec sa = sclf(m.sigma_t, sub(mk(1, 1, 1), m.albedo));
        const float l = -len(sub(a, b));
        return mk(dm::exp(l * sa.x), dm::exp(l * sa.y), dm::exp(l * sa.z));
    }
    const float tmax = len(sub(b, a));
    const Vec la = add(mk(0.5f, 0.5f, 0.5f), sclf(0.5f, med_uvw(m, a)));
    const Vec lb = add(mk(0.5f, 0.5f, 0.5f), sclf(0.5f, med_uvw(m, b)));
    const Vec lstep = dvs(sub(lb, la), tmax);
    float tr = 1.0f, t = 0.0f;
    for(int it = 0; it < MAX_TRACKING; ++it)
    {
        t = t + -dm::log(1.0f - smp.get1d()) * m.inv_maj;
        if(t >= tmax)
            break;
        const Vec uvw = add(la, sclf(t, lstep));
        tr = tr * (1.0f - tex3(m.dens, m.nx, m.ny, m.nz, 1, 0, uvw) * m.inv_maj);
    }
    return mk(tr, tr, tr);
}

// ------------------------------------------------------------------ volume primitives
//
// VolumePrimitiveMedium::sample / tr (core/volume.cpp:108-197) over pieces
// between primitive boundaries; on each piece volume::Aggregate's delta /
// ratio tracking over the primitives holding its midpoint
// (core/volume/aggregate.cpp:128-254). Boundaries and overlap sets by
// scanning every primitive (the reference's BVH finds the same minimum and
// the same set; MAX_OVERLAP_COUNT = 4, core/volume/bvh.h:10).

static Vec vol_uvw(const Scene::Vol &v, Vec p)
{
    const Vec op = sub(p, v.o);
    return mk(dot(v.x, op) * v.ix2, dot(v.y, op) * v.iy2, dot(v.z, op) * v.iz2);
}

static bool vol_inside(const Scene::Vol &v, Vec p)
{
    const Vec op = sub(p, v.o);
    const float a = dot(v.xd, op), b = dot(v.yd, op), c = dot(v.zd, op);
    return 0.0f <= a && a <= 1.0f && 0.0f <= b && b <= 1.0f && 0.0f <= c && c <= 1.0f;
}

static float vol_cross(const Scene::Vol &v, Vec p, Vec d, float tmax)
{
    const Vec op = sub(p, v.o);
    float lo = 0.0f, hi = tmax;
    for(const Vec *ax : { &v.xd, &v.yd, &v.zd })
    {
        const float opd = dot(*ax, op), dd = dot(*ax, d);
        const float a = (0.0f - opd) / dd, b = (1.0f - opd) / dd;
        lo = std::fmax(lo, std::fmin(a, b));
        hi = std::fmin(hi, std::fmax(a, b));
    }
    if(lo > hi)
        return -1.0f;
    if(lo > 0.0f)
        return lo;
    return hi < tmax ? hi : -1.0f;
}

static bool vol_next_boundary(const Scene &s, Vec a, Vec b, Vec &out)
{
    const Vec dir = sub(b, a);
    float best = FMAX, tmax = 1.0f;
    for(const Scene::Vol &v : s.vols)
    {
        const float t = vol_cross(v, a, dir, tmax);
        if(t >= 0.0f)
        {
            best = std::fmin(best, t);
            tmax = best;
        }
    }
    if(best > 10.0f)
        return false;
    out = add(scl(a, 1.0f - best), scl(b, best));
    return true;
}

static std::vector<int> vol_set(const Scene &s, Vec p)
{
    std::vector<int> r;
    for(int i = 0; i < (int)s.vols.size() && r.size() < 4; ++i)
        if(vol_inside(s.vols[i], p))
            r.push_back(i);
    return r;
}

static float vol_density(const Scene::Vol &v, Vec uvw)
{
    return v.nx == 0 ? v.sigma : tex3(v.dens, v.nx, v.ny, v.nz, 1, 0, uvw);
}

static Vec vol_albedo(const Scene::Vol &v, Vec uvw)
{
    if(v.ax == 0)
        return v.albedo;
    return mk(tex3(v.alb, v.ax, v.ay, v.az, 3, 0, uvw), tex3(v.alb, v.ax, v.ay, v.az, 3, 1, uvw),
              tex3(v.alb, v.ax, v.ay, v.az, 3, 2, uvw));
}

// tracking over one piece a -> b with the primitives `set`; sample: true on
// scattering (pos, color set); otherwise multiplies the piece's
// transmittance into *tr when tr is given
template<class S>
static bool vol_piece(const Scene &s, const std::vector<int> &set, Vec a, Vec b, S &smp, Vec *pos, Vec *color,
                      float *tr)
{
    float maj = 0.0f;
    for(int i : set)
        maj = maj + s.vols[i].maxd;
    const float inv = 1.0f / fmx(maj, 0.01f);
    std::vector<Vec> ua, ub;
    for(int i : set)
    {
        ua.push_back(vol_uvw(s.vols[i], a));
        ub.push_back(vol_uvw(s.vols[i], b));
    }
    const float len_ab = len(sub(b, a));
    float t = 0.0f, seg = 1.0f;
    for(;;)
    {
        t = t + -dm::log(1.0f - smp.get1d()) * inv;
        if(t >= len_ab)
            break;
        const float f = t / len_ab;
        std::vector<float> dv;
        float sum = 0.0f;
        for(size_t k = 0; k < set.size(); ++k)
        {
            dv.push_back(vol_density(s.vols[set[k]], add(scl(ua[k], 1.0f - f), scl(ub[k], f))));
            sum = sum + dv.back();
        }
        if(!tr)
        {
            if(smp.get1d() < sum * inv)
            {
                *pos = add(scl(a, 1.0f - f), scl(b, f));
                Vec c = mk(0, 0, 0);
                for(size_t k = 0; k < set.size(); ++k)
                    c = add(c, sclf(dv[k] / sum, vol_albedo(s.vols[set[k]], add(scl(ua[k], 1.0f - f), scl(ub[k], f)))));
                *color = c;
                return true;
            }
        }
        else
            seg = seg * (1.0f - sum * inv);
    }
    if(tr)
        *tr = *tr * seg;
    return false;
}

template<class S>
static bool vol_sample(const Scene &s, Vec a, Vec b, S &smp, Vec &pos, Vec &color)
{
    if(s.vols.empty())
        return false;
    Vec o = a;
    const Vec d = nrm(sub(b, a));
    for(;;)
    {
        if(dot(sub(b, o), d) <= 0.0f || dot(sub(b, o), sub(b, o)) < 0.00001f)
            return false;
        Vec c;
        if(!vol_next_boundary(s, o, b, c))
            c = b;
        const std::vector<int> set = vol_set(s, scl(add(o, c), 0.5f));
        if(!set.empty() && vol_piece(s, set, o, c, smp, &pos, &color, nullptr))
            return true;
        o = add(c, sclf(0.001f, d));
    }
}

template<class S>
static Vec vol_tr(const Scene &s, Vec a, Vec b, S &smp)
{
    float tr = 1.0f;
    if(s.vols.empty())
        return mk(tr, tr, tr);
    Vec o = a;
    const Vec d = nrm(sub(b, a));
    for(;;)
    {
        if(dot(sub(b, o), d) <= 0.0f || dot(sub(b, o), sub(b, o)) < 0.0001f)
            break;
        Vec c;
        if(!vol_next_boundary(s, o, b, c))
            c = b;
        const std::vector<int> set = vol_set(s, scl(add(o, c), 0.5f));
        if(!set.empty())
            vol_piece(s, set, o, c, smp, nullptr, nullptr, &tr);
        o = add(c, sclf(0.001f, d));
    }
    return mk(tr, tr, tr);
}

// ------------------------------------------------------------------ path tracer

struct Params
{
    int32_t w, h, spp, min_depth, max_depth;
    float rr_thr, rr_p;
    Vec eye, lbc, fx, fy;
    float world_diag = 0.0f; // 1.2 |bbox of scene and eye| (wavefront.cpp:138-139)
    bool tile = false, indep = false, gaussian = false;
    float radius = 0.0f, alpha = 0.0f;
    std::vector<float> g_accept;   // Gaussian filter table, 256 x 256 cells
    std::vector<uint32_t> g_alias;
};

// GaussianFilter::commit (film_filter/gaussian.cpp:27-52) with
// AliasTable::initialize (utils/math/alias.cpp:13-58, sum in index order)
static void gaussian_table(Params &p)
{
    const int G = 256;
    const float r = p.radius, al = p.alpha;
    const float ev = std::exp(-al * r * r);
    std::vector<float> v((size_t)G * G, 0.0f);
    for(int y = 0; y < G; ++y)
        for(int x = 0; x < G; ++x)
        {
            const float xf = 2 * r * (x + 0.5f) / G - r, yf = 2 * r * (y + 0.5f) / G - r;
            if(xf * xf + yf * yf >= r * r)
                continue;
            const float gx = std::max(0.0f, std::exp(-al * xf * xf) - ev);
            const float gy = std::max(0.0f, std::exp(-al * yf * yf) - ev);
            v[(size_t)y * G + x] = gx * gy;
        }
    float sum = 0.0f;
    for(float a : v)
        sum += a;
    const uint32_t n = (uint32_t)v.size();
    p.g_accept.assign(n, 0.0f);
    p.g_alias.assign(n, 0);
    const float ratio = (float)n / sum;
    std::vector<uint32_t> ov, un;
    for(uint32_t i = 0; i < n; ++i)
    {
        p.g_accept[i] = v[i] * ratio;
        p.g_alias[i] = i;
        if(p.g_accept[i] > 1) ov.push_back(i);
        else if(p.g_accept[i] < 1) un.push_back(i);
    }
    while(!ov.empty() && !un.empty())
    {
        const uint32_t o = ov.back(), u = un.back();
        ov.pop_back();
        un.pop_back();
        p.g_accept[o] -= 1 - p.g_accept[u];
        p.g_alias[u] = o;
        if(p.g_accept[o] > 1) ov.push_back(o);
        else if(p.g_accept[o] < 1) un.push_back(o);
    }
    for(auto i : ov) p.g_accept[i] = 1;
    for(auto i : un) p.g_accept[i] = 1;
}

// GaussianFilter::sample (gaussian.cpp:54-66) with CAliasTable::sample
// (utils/cmath/calias.cpp:15-21): cell by u0, position in the cell by u1, u2
static void gaussian_sample(const Params &p, float u0, float u1, float u2, float &fx, float &fy)
{
    const uint32_t n = 256u * 256u;
    const float nu = (float)n * u0;
    const uint32_t k = std::min((uint32_t)nu, n - 1);
    const uint32_t cell = (nu - (float)k) <= p.g_accept[k] ? k : p.g_alias[k];
    const float w = 2 * p.radius / 256;
    const float x0 = -p.radius + w * (float)(cell % 256u), x1 = -p.radius + w * (float)(cell % 256u + 1);
    const float y0 = -p.radius + w * (float)(cell / 256u), y1 = -p.radius + w * (float)(cell / 256u + 1);
    fx = x0 * (1.0f - u1) + x1 * u1;
    fy = y0 * (1.0f - u2) + y1 * u2;
}

static inline void splat(std::vector<double> &film, uint32_t pixel, Vec v)
{
    // Film::splat drops non-finite values (core/film.cpp:145-163)
    if(!finite3(v))
        return;
    film[3 * (size_t)pixel + 0] += (double)v.x;
    film[3 * (size_t)pixel + 1] += (double)v.y;
    film[3 * (size_t)pixel + 2] += (double)v.z;
}

// One camera path: generate_kernel (wavefront/generate.cpp:33-87), then per
// bounce the trace programs (trace.cpp:38-78), shade_kernel
// (shade.cpp:29-236) and the shadow programs (shadow.cpp:24-76).
static void trace_path(const Scene &s, const Params &p, uint64_t path_index, std::vector<double> &film,
                       std::vector<double> *aov, uint64_t &ext_rays, uint64_t &shadow_rays)
{
    const uint64_t pixels = (uint64_t)p.w * (uint64_t)p.h;
    // generate.cpp:46-62: uniform or tile path order
    const uint32_t pixel = p.tile ? (uint32_t)(path_index / (uint64_t)p.spp) : (uint32_t)(path_index % pixels);
    const int32_t sample = p.tile ? (int32_t)(path_index % (uint64_t)p.spp) : (int32_t)(path_index / pixels);
    const uint32_t px = pixel % (uint32_t)p.w, py = pixel / (uint32_t)p.w;
    Smp hs;
    hs.indep = p.indep;
    if(p.indep)
        hs.p.init(px, py, sample);
    else
        hs.h.init(p.w, p.h, px, py, sample);
    float fx, fy;
    if(p.gaussian)
    {
        const float u0 = hs.get1d();
        const float u1 = hs.get1d();
        const float u2 = hs.get1d();
        gaussian_sample(p, u0, u1, u2, fx, fy);
    }
    else
    {
        fx = hs.get1d() - 0.5f; // BoxFilter::sample, film_filter/box.cpp:5-9
        fy = hs.get1d() - 0.5f;
    }
    const float film_x = ((float)px + 0.5f + fx) / (float)p.w;
    const float film_y = ((float)py + 0.5f + fy) / (float)p.h;
    (void)hs.get1d();
    // PinholeCamera::generate_ray_inline, camera/pinhole.cpp:84-99
    const Vec dst = add(add(p.lbc, sclf(film_x, p.fx)), sclf(film_y, p.fy));
    Vec o = p.eye, d = nrm(sub(dst, p.eye));
    Vec beta = mk(1, 1, 1), rad = mk(0, 0, 0), beta_le = mk(1, 1, 1);
    float bsdf_pdf = -1.0f;
    uint32_t ray_med = (uint32_t)s.R; // the volume-primitive medium (generate.cpp:84)

    // shadow ray: occlusion over all geometry, then the transmittance of its
    // medium on a miss (shadow.cpp:43-71)
    auto shadow = [&](Vec so, Vec sd, float t1, uint32_t med, Vec bl) {
        ++shadow_rays;
        const uint64_t seed = shadow_seed(hs.pack(pixel));
        if(trace(s, BTRC_TRI_WATERTIGHT, true, true, so, sd, 0.0f, t1).gid != 0xffffffffu)
            return;
        Pcg tr_rng;
        tr_rng.seed(seed, mixbits(seed));
        if(t1 > 1.0f)
        {
            // toward the environment: the volume-primitive medium over the
            // world's diagonal (shadow.cpp:56-60)
            if(!s.vols.empty())
                bl = mul(bl, vol_tr(s, so, add(so, sclf(p.world_diag, nrm(sd))), tr_rng));
        }
        else if(med < (uint32_t)s.R)
            bl = mul(bl, med_tr(s.media[med], so, add(so, sclf(t1, sd)), tr_rng));
        else if(med == (uint32_t)s.R && !s.vols.empty())
            bl = mul(bl, vol_tr(s, so, add(so, sclf(t1, sd)), tr_rng));
        splat(film, pixel, bl);
    };
    // simple_russian_roulette, helper.cpp:67-97; true: the path ends
    auto roulette = [&](int depth) {
        if(depth < p.min_depth)
            return false;
        if(depth >= p.max_depth)
            return true;
        const float u = hs.get1d();
        if(lum(beta) < p.rr_thr)
        {
            if(u > p.rr_p)
                return true;
            beta = dvs(beta, p.rr_p);
        }
        return false;
    };

    for(int depth = 0;; ++depth)
    {
        ++ext_rays;
        const Hit h = trace(s, BTRC_TRI_WATERTIGHT, false, true, o, d, 0.0f, FMAX);
        const bool hit = h.gid != 0xffffffffu;

        // medium stage (wavefront/medium.cpp:28-162): segment, delta tracking,
        // transmittance into beta_le
        bool scattered = false;
        Vec mpos{}, mcolor{};
        float mg = 0.0f;
        uint32_t med = 0;
        if(s.R > 0 || !s.vols.empty())
        {
            Vec end;
            if(hit)
            {
                const Inst &in = s.insts[s.tinst[h.gid]];
                const Vec nor = xnor(in.xf, s.gz[s.geos[in.geo].off + s.tprim[h.gid]]);
                const uint32_t im = dot(d, nor) < 0.0f ? in.outer : in.inner;
                med = std::min(im, ray_med);
                end = offset_dir(add(o, sclf(h.t, d)), nor, neg(d));
            }
            else
            {
                // get_medium_segment, medium.cpp:244-248
                med = (uint32_t)s.R;
                end = add(o, sclf(p.world_diag, nrm(d)));
            }
            ray_med = med;
            if(med < (uint32_t)s.R)
            {
                const Med &m = s.media[med];
                scattered = med_sample(m, o, end, hs, mpos, mcolor);
                mg = m.g;
                beta_le = mul(beta_le, med_tr(m, o, end, hs));
            }
            else if(med == (uint32_t)s.R && !s.vols.empty())
            {
                // the volume-primitive medium: isotropic scattering
                scattered = vol_sample(s, o, end, hs, mpos, mcolor);
                mg = 0.0f;
                beta_le = mul(beta_le, vol_tr(s, o, end, hs));
            }
        }

        if(!hit)
        {
            // eval_miss_le, helper.cpp:144-172
            Vec sp = mk(0, 0, 0);
            if(s.env >= 0)
            {
                const Vec le = env_le(s, s.lights[s.env], d);
                if(bsdf_pdf < 0.0f)
                    sp = dvs(mul(beta_le, le), -bsdf_pdf);
                else
                {
                    const float lp = (1.0f / (float)s.lights.size()) * env_pdf(s.lights[s.env], d);
                    sp = dvs(mul(beta_le, le), bsdf_pdf + lp);
                }
            }
            if(!scattered)
            {
                splat(film, pixel, add(sp, rad));
                return;
            }
            splat(film, pixel, sp);
        }
        const Inst *inp = hit ? &s.insts[s.tinst[h.gid]] : nullptr;
        SP sp{};
        if(hit)
        {
            const Inst &in = *inp;
            sp = hitinfo(s, o, d, in, s.tprim[h.gid], h.t, h.u, h.v);
            // handle_intersected_light, helper.cpp:174-209
            Vec c = mk(0, 0, 0);
            if(in.light >= 0 && s.lights[in.light].type == 0)
            {
                const Light &l = s.lights[in.light];
                const Vec le = dot(sp.frame.z, neg(d)) >= 0.0f ? l.intensity : mk(0, 0, 0);
                if(bsdf_pdf < 0.0f)
                    c = dvs(mul(beta_le, le), -bsdf_pdf);
                else
                {
                    float lpdf = 0.0f;
                    const Vec toref = sub(o, sp.pos);
                    if(dot(sp.frame.z, toref) > 0.0f)
                    {
                        const float d2 = dot(toref, toref);
                        const float d3 = std::sqrt(d2) * d2;
                        lpdf = 1.0f / (l.scale * l.scale) * (1.0f / s.geos[l.geo].total_area) * d3
                             / std::fabs(dot(sp.frame.z, toref));
                    }
                    c = dvs(mul(beta_le, le), bsdf_pdf + (1.0f / (float)s.lights.size()) * lpdf);
                }
            }
            if(scattered)
                splat(film, pixel, c);
            else
                rad = add(rad, c);
        }

        if(scattered)
        {
            // the medium vertex (medium.cpp:86-160)
            if(roulette(depth))
            {
                splat(film, pixel, rad);
                return;
            }
            float l0, l1, l2;
            hs.get3d(l0, l1, l2);
            const float lsel = hs.get1d();
            Shadow sh = sample_light(s, mpos, mk(0, 0, 0), false, l0, l1, l2, lsel);
            Vec bl{};
            if(sh.ok)
            {
                const float pv = hg(mg, -dot(nrm(sh.d), nrm(neg(d))));
                bl = dvs(mul(mul(sh.li, beta), scl(mcolor, pv)), pv + sh.pdf);
            }
            float s0, s1, s2;
            hs.get3d(s0, s1, s2);
            Vec pdir;
            float ppdf;
            const Vec phase = hg_sample(mg, mcolor, neg(d), s0, s1, pdir, ppdf);
            if(sh.ok)
                shadow(sh.o, sh.d, sh.t1, sh.t1 > 1.0f ? (uint32_t)s.R : med, bl);
            if(zero(phase))
            {
                splat(film, pixel, rad);
                return;
            }
            d = nrm(pdir);
            beta_le = mul(beta, phase);
            beta = dvs(beta_le, ppdf);
            bsdf_pdf = ppdf;
            o = mpos;
            ray_med = med;
            continue;
        }

        const Inst &in = *inp;
        if(roulette(depth))
        {
            splat(film, pixel, rad);
            return;
        }
        float l0, l1, l2;
        hs.get3d(l0, l1, l2);
        const float lsel = hs.get1d();
        Shadow sh = sample_light(s, sp.pos, sp.frame.z, true, l0, l1, l2, lsel);

        const btrc_material_desc m = material_at(s, in.mat, sp);
        Shd shd;
        shd.geo = sp.frame;
        shd.sh = normal_mapped(s, m, sp, rot_z(sp.frame, sp.iz));
        shd.st = m.shadow_terminator != 0;
        shd.disney = m.type == BTRC_MATERIAL_DISNEY;
        shd.invisible = m.type == BTRC_MATERIAL_INVISIBLE;
        shd.albedo = shd.invisible ? mk(1, 1, 1) : mk(m.color[0], m.color[1], m.color[2]);
        if(shd.disney)
            shd.d = dis_make(m);
        if(aov && depth == 0)
        {
            // first-hit g-buffer: albedo = material colour, normal = shading z
            // (shade.cpp:133-139,196-204; aggregate.cpp:211-219); an invisible
            // surface reports one and its geometric normal (invisible.cpp:41-55)
            const Vec gb[2] = { shd.albedo, shd.invisible ? shd.geo.z : shd.sh.z };
            for(int k = 0; k < 2; ++k)
                if(finite3(gb[k]))
                {
                    (*aov)[6 * (size_t)pixel + 3 * k + 0] += (double)gb[k].x;
                    (*aov)[6 * (size_t)pixel + 3 * k + 1] += (double)gb[k].y;
                    (*aov)[6 * (size_t)pixel + 3 * k + 2] += (double)gb[k].z;
                }
        }
        float s0, s1, s2;
        hs.get3d(s0, s1, s2);
        const BS bs = shd_sample(shd, neg(d), s0, s1, s2);

        if(sh.ok)
        {
            const Vec f = shd_eval(shd, sh.d, neg(d));
            if(!zero(f))
            {
                const float pb = shd_pdf(shd, sh.d, neg(d));
                const float c = std::fabs(dot(nrm(sh.d), sp.frame.z));
                const Vec bl = dvs(scl(mul(mul(sh.li, beta), f), c), pb + sh.pdf);
                // shade.cpp:166-180: the medium on the shadow ray's side
                const bool sh_out = dot(sp.frame.z, sh.d) > 0.0f, last_out = dot(sp.frame.z, d) < 0.0f;
                const uint32_t smed = sh_out == last_out ? ray_med : (sh_out ? in.outer : in.inner);
                shadow(sh.o, sh.d, sh.t1, smed, bl);
            }
        }
        if(zero(bs.f))
        {
            splat(film, pixel, rad);
            return;
        }
        const Vec dir = nrm(bs.dir);
        const float c = std::fabs(dot(dir, sp.frame.z));
        beta_le = mul(scl(beta, c), bs.f);
        beta = dvs(beta_le, bs.pdf);
        bsdf_pdf = bs.delta ? -bs.pdf : bs.pdf;
        o = offset_dir(sp.pos, sp.frame.z, dir);
        ray_med = dot(sp.frame.z, dir) > 0.0f ? in.outer : in.inner;
        d = dir;
    }
}

// ------------------------------------------------------------------ radiance cache

// The cache network in FP64 (no reference counterpart): 64 inputs, five
// hidden layers of 64 ReLU units, 3 linear outputs, no biases; weights as
// six [rows][64] blocks (the output block has 32 rows, 3 used). With
// `bf16`, the device's rounding points are reproduced: weights, inputs and
// hidden activations, and the scaled backward gradients, rounded to BF16
// (round to nearest even); the sums stay FP64.
namespace nrc {

const int IN = 64, W = 64, L = 6, OUT = 3, OUT_ROWS = 32;
const int PARAMS = 5 * W * IN + OUT_ROWS * IN;

static double bf16(double x)
{
    const float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if((u & 0x7fffffffu) > 0x7f800000u)
        return (double)f;
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return (double)r;
}

static int rows(int l) { return l == L - 1 ? OUT_ROWS : W; }
static int offset(int l) { return l * W * IN; }

// input encoding: input k = 16 s + 8 h + j (s slice, h half, j 0..7) is
// element e = 8 s + j of half h: for e < 24 sin (e even) or cos (e odd) of
// 2^f pi x_a with f = 4 h + e / 6, a = (e % 6) / 2 (each half's chain starts
// at 2^(4 h) pi x_a and doubles by the double-angle identities in FP32);
// then half 0: direction, normal, roughness, diffuse red; half 1: diffuse
// green, blue, specular, three constant ones
static void encode(const float *v, const float *box, double *x)
{
    for(int h = 0; h < 2; ++h)
    {
        float e[32];
        const float base = h ? 16.0f * PI : PI;
        float sn[3], cs[3];
        for(int a = 0; a < 3; ++a)
        {
            const float q = (v[a] - box[a]) * box[3 + a];
            dm::sincos(base * q, &sn[a], &cs[a]);
        }
        int k = 0;
        for(int f = 0; f < 4; ++f)
            for(int a = 0; a < 3; ++a)
            {
                e[k++] = sn[a];
                e[k++] = cs[a];
                const float s2 = 2.0f * sn[a] * cs[a];
                const float c2 = (cs[a] - sn[a]) * (cs[a] + sn[a]);
                sn[a] = s2;
                cs[a] = c2;
            }
        // v = p[3], wo[3], n[3], roughness, diffuse[3], specular[3]
        const float t0[8] = { v[3], v[4], v[5], v[6], v[7], v[8], v[9], v[10] };
        const float t1[8] = { v[11], v[12], v[13], v[14], v[15], 1.0f, 1.0f, 1.0f };
        for(int j = 0; j < 8; ++j)
            e[k++] = h ? t1[j] : t0[j];
        for(int s = 0; s < 4; ++s)
            for(int j = 0; j < 8; ++j)
                x[16 * s + 8 * h + j] = e[8 * s + j];
    }
}

struct Act
{
    double x[L][IN];   // input of each layer
    double y[OUT];
};

static void forward(const std::vector<double> &w, const float *v, const float *box, bool b16, Act &a)
{
    encode(v, box, a.x[0]);
    if(b16)
        for(int k = 0; k < IN; ++k)
            a.x[0][k] = bf16(a.x[0][k]);
    for(int l = 0; l < L; ++l)
    {
        const int n = l == L - 1 ? OUT : W;
        for(int r = 0; r < n; ++r)
        {
            double z = 0.0;
            for(int c = 0; c < IN; ++c)
                z += w[offset(l) + r * IN + c] * a.x[l][c];
            if(l == L - 1)
                a.y[r] = z;
            else
            {
                const double h = z > 0.0 ? z : 0.0;
                a.x[l + 1][r] = b16 ? bf16(h) : h;
            }
        }
    }
}

static std::vector<double> weights(const float *wf, bool b16)
{
    std::vector<double> w(PARAMS);
    for(int k = 0; k < PARAMS; ++k)
        w[k] = b16 ? bf16(wf[k]) : (double)wf[k];
    return w;
}

} // namespace nrc

} // namespace orc

// ================================================================== C API

extern "C" {

void *oracle_scene_create(const btrc_scene_desc *sd) { return orc::make_scene(sd); }
void oracle_scene_destroy(void *s) { delete static_cast<orc::Scene *>(s); }

int64_t oracle_scene_bvh_nodes(void *s) { return (int64_t)static_cast<orc::Scene *>(s)->nodes.size(); }

// closest hits as btrc_hit, or occlusion flags when any != 0
void oracle_trace(void *sp, const btrc_ray *rays, int64_t n, int32_t mode, int32_t any, int32_t use_bvh,
                  btrc_hit *hits, uint8_t *occluded, int32_t threads)
{
    const orc::Scene &s = *static_cast<orc::Scene *>(sp);
    auto work = [&](int64_t b, int64_t e) {
        for(int64_t i = b; i < e; ++i)
        {
            const btrc_ray &r = rays[i];
            const orc::Hit h = orc::trace(s, mode, any != 0, use_bvh != 0, orc::mk(r.o[0], r.o[1], r.o[2]),
                                          orc::mk(r.d[0], r.d[1], r.d[2]), r.tmin, r.tmax);
            if(any)
                occluded[i] = h.gid != 0xffffffffu;
            else
            {
                btrc_hit &o = hits[i];
                const bool hit = h.gid != 0xffffffffu;
                o.t = hit ? h.t : r.tmax;
                o.instance = hit ? s.tinst[h.gid] : 0xffffffffu;
                o.prim = hit ? s.tprim[h.gid] : 0xffffffffu;
                o.u = h.u;
                o.v = h.v;
            }
        }
    };
    threads = std::max(1, threads);
    std::vector<std::thread> pool;
    const int64_t chunk = (n + threads - 1) / threads;
    for(int t = 0; t < threads; ++t)
    {
        const int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
        if(b < e)
            pool.emplace_back(work, b, e);
    }
    for(auto &th : pool)
        th.join();
}

// Renders paths [first, first + count) of the uniform path order and adds
// their radiance to film (3 doubles per pixel, not normalised). Returns the
// number of extension and shadow rays traced through ext/shadow.
// filter may be NULL (box); aov_out (6 doubles per pixel: albedo, normal
// sums of the first hits) may be NULL.
void oracle_render(void *sp, const btrc_camera_desc *cam, const btrc_render_params *rp,
                   const btrc_filter_desc *filter, uint64_t first, uint64_t count, double *film_out, double *aov_out,
                   int32_t threads, uint64_t *ext_out, uint64_t *shadow_out)
{
    const orc::Scene &s = *static_cast<orc::Scene *>(sp);
    orc::Params p;
    p.w = rp->width;
    p.h = rp->height;
    p.spp = rp->spp;
    p.min_depth = rp->min_depth;
    p.max_depth = rp->max_depth;
    p.rr_thr = rp->rr_threshold;
    p.rr_p = rp->rr_cont_prob;
    p.tile = rp->tile != 0;
    p.indep = rp->sampler == BTRC_SAMPLER_INDEPENDENT;
    if(filter && filter->type == BTRC_FILTER_GAUSSIAN)
    {
        p.gaussian = true;
        p.radius = filter->radius;
        p.alpha = filter->alpha;
        orc::gaussian_table(p);
    }
    // PinholeCamera::commit, camera/pinhole.cpp:55-77
    {
        using namespace orc;
        const Vec eye = mk(cam->eye[0], cam->eye[1], cam->eye[2]);
        const Vec fwd = nrm(sub(mk(cam->dst[0], cam->dst[1], cam->dst[2]), eye));
        const Vec ex = nrm(cross(fwd, mk(cam->up[0], cam->up[1], cam->up[2])));
        const Vec ey = nrm(cross(fwd, ex));
        const float yl = 2.0f * std::tan(0.5f * cam->fov_y_deg * PI / 180.0f);
        const float xl = cam->w_over_h * yl;
        p.eye = eye;
        const Vec lo = mk(fmn(s.bb_lo.x, eye.x), fmn(s.bb_lo.y, eye.y), fmn(s.bb_lo.z, eye.z));
        const Vec hi = mk(fmx(s.bb_hi.x, eye.x), fmx(s.bb_hi.y, eye.y), fmx(s.bb_hi.z, eye.z));
        p.world_diag = 1.2f * len(sub(hi, lo));
        p.lbc = sub(add(eye, fwd), sclf(0.5f, add(sclf(xl, ex), sclf(yl, ey))));
        p.fx = sclf(xl, ex);
        p.fy = sclf(yl, ey);
    }
    const size_t pixels = (size_t)p.w * (size_t)p.h;
    threads = std::max(1, threads);
    std::vector<std::vector<double>> films(threads, std::vector<double>(3 * pixels, 0.0));
    std::vector<std::vector<double>> aovs(aov_out ? threads : 0, std::vector<double>(6 * pixels, 0.0));
    std::vector<uint64_t> ext(threads, 0), shd(threads, 0);
    std::vector<std::thread> pool;
    for(int t = 0; t < threads; ++t)
    {
        pool.emplace_back([&, t] {
            for(uint64_t i = (uint64_t)t; i < count; i += (uint64_t)threads)
                orc::trace_path(s, p, first + i, films[t], aov_out ? &aovs[t] : nullptr, ext[t], shd[t]);
        });
    }
    for(auto &th : pool)
        th.join();
    uint64_t e = 0, sh = 0;
    for(int t = 0; t < threads; ++t)
    {
        for(size_t k = 0; k < 3 * pixels; ++k)
            film_out[k] += films[t][k];
        if(aov_out)
            for(size_t k = 0; k < 6 * pixels; ++k)
                aov_out[k] += aovs[t][k];
        e += ext[t];
        sh += shd[t];
    }
    if(ext_out) *ext_out = e;
    if(shadow_out) *shadow_out = sh;
}

void oracle_halton(int32_t w, int32_t h, const uint32_t *px, const uint32_t *py, const int32_t *si, int32_t n,
                   int32_t dims, float *out)
{
    for(int32_t i = 0; i < n; ++i)
    {
        orc::Halton hs;
        hs.init(w, h, px[i], py[i], si[i]);
        for(int32_t k = 0; k < dims; ++k)
            out[(size_t)i * dims + k] = hs.get1d();
    }
}

void oracle_independent(const uint32_t *px, const uint32_t *py, const int32_t *si, int32_t n, int32_t dims,
                        float *out)
{
    for(int32_t i = 0; i < n; ++i)
    {
        orc::Indep smp;
        smp.init(px[i], py[i], si[i]);
        for(int32_t k = 0; k < dims; ++k)
            out[(size_t)i * dims + k] = smp.get1d();
    }
}

void oracle_gaussian_filter(float radius, float alpha, const float *u, int32_t n, float *out)
{
    orc::Params p{};
    p.gaussian = true;
    p.radius = radius;
    p.alpha = alpha;
    orc::gaussian_table(p);
    for(int32_t i = 0; i < n; ++i)
        orc::gaussian_sample(p, u[3 * i], u[3 * i + 1], u[3 * i + 2], out[2 * i], out[2 * i + 1]);
}

void oracle_volume_eval(void *sp, const btrc_ray *seg, const uint64_t *seeds, int32_t n, float *out)
{
    const orc::Scene &s = *static_cast<orc::Scene *>(sp);
    for(int32_t i = 0; i < n; ++i)
    {
        const orc::Vec a = orc::mk(seg[i].o[0], seg[i].o[1], seg[i].o[2]);
        const orc::Vec d = orc::mk(seg[i].d[0], seg[i].d[1], seg[i].d[2]);
        const orc::Vec b = orc::add(a, orc::scl(d, seg[i].tmax));
        orc::Pcg r;
        r.seed(seeds[i], orc::mixbits(seeds[i]));
        orc::Vec pos = orc::mk(0, 0, 0), col = orc::mk(0, 0, 0);
        const bool sc = orc::vol_sample(s, a, b, r, pos, col);
        const orc::Vec tr = orc::vol_tr(s, a, b, r);
        float *o = out + 10 * (size_t)i;
        o[0] = sc ? 1.0f : 0.0f;
        o[1] = pos.x; o[2] = pos.y; o[3] = pos.z;
        o[4] = col.x; o[5] = col.y; o[6] = col.z;
        o[7] = tr.x; o[8] = tr.y; o[9] = tr.z;
    }
}

void oracle_texture_eval(void *sp, int32_t texture, const float *uv, int32_t n, float *out)
{
    const orc::Scene &s = *static_cast<orc::Scene *>(sp);
    for(int32_t i = 0; i < n; ++i)
    {
        const orc::F4 r = orc::tex_sample(s, texture, uv[2 * i], uv[2 * i + 1]);
        out[4 * i] = r.x;
        out[4 * i + 1] = r.y;
        out[4 * i + 2] = r.z;
        out[4 * i + 3] = r.w;
    }
}

void oracle_medium_eval(void *sp, int32_t medium, const btrc_ray *seg, const uint64_t *seeds, int32_t n, float *out)
{
    const orc::Scene &s = *static_cast<orc::Scene *>(sp);
    const orc::Med &m = s.media[medium];
    for(int32_t i = 0; i < n; ++i)
    {
        const orc::Vec a = orc::mk(seg[i].o[0], seg[i].o[1], seg[i].o[2]);
        const orc::Vec d = orc::mk(seg[i].d[0], seg[i].d[1], seg[i].d[2]);
        const orc::Vec b = orc::add(a, orc::scl(d, seg[i].tmax));
        orc::Pcg r;
        r.seed(seeds[i], orc::mixbits(seeds[i]));
        orc::Vec pos = orc::mk(0, 0, 0), col = orc::mk(0, 0, 0);
        const bool sc = orc::med_sample(m, a, b, r, pos, col);
        const orc::Vec tr = orc::med_tr(m, a, b, r);
        float *o = out + 10 * (size_t)i;
        o[0] = sc ? 1.0f : 0.0f;
        o[1] = pos.x; o[2] = pos.y; o[3] = pos.z;
        o[4] = col.x; o[5] = col.y; o[6] = col.z;
        o[7] = tr.x; o[8] = tr.y; o[9] = tr.z;
    }
}

void oracle_hg(float g, const float *u, int32_t n, float *out)
{
    // out: dir xyz and pdf for wo = +z, then the pdf re-evaluated from dir
    for(int32_t i = 0; i < n; ++i)
    {
        orc::Vec dir;
        float pdf;
        orc::hg_sample(g, orc::mk(1, 1, 1), orc::mk(0, 0, 1), u[2 * i], u[2 * i + 1], dir, pdf);
        out[5 * i + 0] = dir.x; out[5 * i + 1] = dir.y; out[5 * i + 2] = dir.z;
        out[5 * i + 3] = pdf;
        out[5 * i + 4] = orc::hg(g, -orc::dot(orc::nrm(dir), orc::mk(0, 0, 1)));
    }
}

float oracle_tex3(const float *grid, int32_t nx, int32_t ny, int32_t nz, float u, float v, float w)
{
    const std::vector<float> g(grid, grid + (size_t)nx * ny * nz);
    return orc::tex3(g, nx, ny, nz, 1, 0, orc::mk(u, v, w));
}

void oracle_nrc_forward(const float *w, const float *box, const float *v, int32_t n, int32_t b16, double *out)
{
    const std::vector<double> wd = orc::nrc::weights(w, b16 != 0);
    orc::nrc::Act a;
    for(int32_t i = 0; i < n; ++i)
    {
        orc::nrc::forward(wd, v + 16 * (size_t)i, box, b16 != 0, a);
        for(int c = 0; c < 3; ++c)
            out[3 * (size_t)i + c] = a.y[c];
    }
}

// gradient of mean_i,c (y - t)^2 / (y^2 + 0.01) (the prediction in the
// denominator held constant) with respect to the weights, and the loss
void oracle_nrc_gradients(const float *w, const float *box, const float *v, const float *t, int32_t n, int32_t b16,
                          float loss_scale, double *grad, double *loss)
{
    using namespace orc::nrc;
    const std::vector<double> wd = weights(w, b16 != 0);
    std::fill(grad, grad + PARAMS, 0.0);
    double lsum = 0.0;
    Act a;
    for(int32_t i = 0; i < n; ++i)
    {
        forward(wd, v + 16 * (size_t)i, box, b16 != 0, a);
        double dz[IN] = {};
        for(int c = 0; c < OUT; ++c)
        {
            const double y = a.y[c], tc = t[3 * (size_t)i + c];
            const float yf = (float)y;
            const double den = (double)(yf * yf + 0.01f);
            lsum += (y - tc) * (y - tc) / den;
            double g = 2.0 * (y - tc) / den;
            if(b16)
                g = bf16(g * loss_scale);
            dz[c] = g;
        }
        for(int l = L - 1; l >= 0; --l)
        {
            const int nr = l == L - 1 ? OUT : W;
            for(int r = 0; r < nr; ++r)
                for(int c = 0; c < IN; ++c)
                    grad[offset(l) + r * IN + c] += dz[r] * a.x[l][c];
            if(l == 0)
                break;
            double dx[IN];
            for(int c = 0; c < IN; ++c)
            {
                double sum = 0.0;
                for(int r = 0; r < nr; ++r)
                    sum += wd[offset(l) + r * IN + c] * dz[r];
                const double g = a.x[l][c] > 0.0 ? sum : 0.0;
                dx[c] = b16 ? bf16(g) : g;
            }
            for(int c = 0; c < IN; ++c)
                dz[c] = dx[c];
        }
    }
    // per-sample gradients were summed (and scaled by loss_scale with bf16)
    const double norm = (b16 ? (double)loss_scale : 1.0) * 3.0 * n;
    for(int k = 0; k < PARAMS; ++k)
        grad[k] /= norm;
    *loss = lsum / (3.0 * n);
}

uint64_t oracle_murmur64a(const uint8_t *key, uint64_t len, uint64_t seed)
{
    return orc::murmur64a(key, (size_t)len, seed);
}

// n outputs of PCG32 seeded (seq, offset) after jumping `skip` numbers ahead
void oracle_pcg32(uint64_t seq, uint64_t offset, uint64_t skip, int32_t n, uint32_t *out)
{
    orc::Pcg r;
    r.seed(seq, offset);
    r.advance(skip);
    for(int32_t i = 0; i < n; ++i)
        out[i] = r.next();
}

// the environment light at n directions / sample triples: 12 floats each
// (le rgb, pdf_li, sampled direction xyz, pdf, radiance rgb, 0)
void oracle_envir_eval(void *sp, const float *dirs, const float *sam, int32_t n, float *out)
{
    const orc::Scene &s = *static_cast<orc::Scene *>(sp);
    const orc::Light &l = s.lights[s.env];
    for(int32_t i = 0; i < n; ++i)
    {
        const orc::Vec d = orc::mk(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
        const orc::Vec le = orc::env_le(s, l, d);
        const float pdf = orc::env_pdf(l, d);
        orc::Vec sd, li;
        float spdf;
        orc::env_sample(s, l, sam[3 * i], sam[3 * i + 1], sam[3 * i + 2], sd, li, spdf);
        const float r[12] = { le.x, le.y, le.z, pdf, sd.x, sd.y, sd.z, spdf, li.x, li.y, li.z, 0.0f };
        for(int k = 0; k < 12; ++k)
            out[12 * i + k] = r[k];
    }
}

// the IBL's tile probabilities; returns the table size
int32_t oracle_envir_table(void *sp, float *probs, int32_t capacity)
{
    const orc::Scene &s = *static_cast<orc::Scene *>(sp);
    if(s.env < 0)
        return 0;
    const std::vector<float> &p = s.lights[s.env].probs;
    for(int32_t i = 0; i < (int32_t)p.size() && i < capacity; ++i)
        probs[i] = p[i];
    return (int32_t)p.size();
}

void oracle_detmath(int32_t fn, const float *x, const float *y, int32_t n, float *out)
{
    namespace dm = btrc_detmath;
    for(int32_t i = 0; i < n; ++i)
    {
        const float a = x[i];
        float r;
        switch(fn)
        {
        case 0: r = dm::sin(a); break;
        case 1: r = dm::cos(a); break;
        case 2: r = dm::log(a); break;
        case 3: r = dm::exp(a); break;
        case 4: r = dm::pow(a, y[i]); break;
        case 5: r = std::sqrt(a); break;
        case 7: r = dm::atan2(a, y[i]); break;
        case 8: r = dm::acos(a); break;
        case 9: r = orc::offset_pos(orc::mk(a, 0.0f, 0.0f), orc::mk(y[i], 0.0f, 0.0f)).x; break;
        default: r = a / y[i]; break;
        }
        out[i] = r;
    }
}

// medium-stage work counters since the last call (segments, lookups); resets them
void oracle_tracking_stats(uint64_t *segments, uint64_t *lookups)
{
    *segments = orc::g_track_segments.exchange(0);
    *lookups = orc::g_track_lookups.exchange(0);
}

// ACESToneMap, postprocess/tonemap.cpp:14-24,35-51: the Narkowicz fit of
// exposure-scaled RGB, saturated; alpha unchanged
void oracle_aces(const float *rgba, int32_t n, float exposure, float *out)
{
    for(int32_t i = 0; i < n; ++i)
    {
        for(int c = 0; c < 3; ++c)
        {
            const float x = rgba[4 * i + c] * exposure;
            const float y = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
            out[4 * i + c] = !(y > 0.0f) ? 0.0f : (y > 1.0f ? 1.0f : y); // saturate: NaN -> 0
        }
        out[4 * i + 3] = rgba[4 * i + 3];
    }
}

// BSDF evaluation hooks for the statistical tests: a material at a surface
// whose geometric and shading frames are the identity frame.
void oracle_bsdf_sample(const btrc_material_desc *m, const float *wo, const float *u, int32_t n,
                        float *f_out, float *dir_out, float *pdf_out)
{
    using namespace orc;
    Shd s;
    s.geo = Frm{ mk(1, 0, 0), mk(0, 1, 0), mk(0, 0, 1) };
    s.sh = s.geo;
    s.st = m->shadow_terminator != 0;
    s.disney = m->type == BTRC_MATERIAL_DISNEY;
    s.albedo = mk(m->color[0], m->color[1], m->color[2]);
    if(s.disney)
        s.d = dis_make(*m);
    for(int32_t i = 0; i < n; ++i)
    {
        const BS b = shd_sample(s, mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), u[3 * i], u[3 * i + 1], u[3 * i + 2]);
        f_out[3 * i] = b.f.x; f_out[3 * i + 1] = b.f.y; f_out[3 * i + 2] = b.f.z;
        dir_out[3 * i] = b.dir.x; dir_out[3 * i + 1] = b.dir.y; dir_out[3 * i + 2] = b.dir.z;
        pdf_out[i] = b.pdf;
    }
}

void oracle_bsdf_eval(const btrc_material_desc *m, const float *wi, const float *wo, int32_t n,
                      float *f_out, float *pdf_out)
{
    using namespace orc;
    Shd s;
    s.geo = Frm{ mk(1, 0, 0), mk(0, 1, 0), mk(0, 0, 1) };
    s.sh = s.geo;
    s.st = m->shadow_terminator != 0;
    s.disney = m->type == BTRC_MATERIAL_DISNEY;
    s.albedo = mk(m->color[0], m->color[1], m->color[2]);
    if(s.disney)
        s.d = dis_make(*m);
    for(int32_t i = 0; i < n; ++i)
    {
        const Vec a = mk(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        const Vec b = mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
        const Vec f = shd_eval(s, a, b);
        f_out[3 * i] = f.x; f_out[3 * i + 1] = f.y; f_out[3 * i + 2] = f.z;
        pdf_out[i] = shd_pdf(s, a, b);
    }
}

} // extern "C"

// The lib/bvh image known-answer test: camera rays of lib/bvh's benchmark
// renderer (lib/bvh/test/benchmark.cpp:120-145, defaults 1080x720) and its
// |normal| shading written as 8-bit RGB rows bottom-up (benchmark.cpp:170-176,
// 253-264).
extern "C" void oracle_kat_rays(int32_t width, int32_t height, const float *eye, const float *dir, const float *up,
                                float fov, btrc_ray *out)
{
    using namespace orc;
    const Vec d = nrm(mk(dir[0], dir[1], dir[2]));
    Vec iu = nrm(cross(d, mk(up[0], up[1], up[2])));
    Vec iv = nrm(cross(iu, d));
    const float iw = std::tan(fov * (float)(3.14159265 * (1.0 / 180.0) * 0.5));
    const float ratio = (float)height / (float)width;
    iu = scl(iu, iw);
    iv = scl(scl(iv, iw), ratio);
    for(int32_t i = 0; i < width; ++i)
        for(int32_t j = 0; j < height; ++j)
        {
            const float u = 2.0f * ((float)i + 0.5f) / (float)width - 1.0f;
            const float v = 2.0f * ((float)j + 0.5f) / (float)height - 1.0f;
            const Vec rd = nrm(add(add(scl(iu, u), scl(iv, v)), d));
            btrc_ray &r = out[(size_t)width * j + i];
            r.o[0] = eye[0]; r.o[1] = eye[1]; r.o[2] = eye[2];
            r.d[0] = rd.x; r.d[1] = rd.y; r.d[2] = rd.z;
            r.tmin = 0.0f;
            r.tmax = FMAX;
        }
}

// tris: 9 floats per triangle; hits in the ray order of oracle_kat_rays;
// out: width * height * 3 bytes, first row = top of the image
extern "C" void oracle_kat_pixels(int32_t width, int32_t height, const float *tris, const btrc_hit *hits,
                                  uint8_t *out)
{
    using namespace orc;
    for(int32_t j = 0; j < height; ++j)
        for(int32_t i = 0; i < width; ++i)
        {
            const btrc_hit &h = hits[(size_t)width * j + i];
            float px[3] = { 0, 0, 0 };
            if(h.instance != 0xffffffffu)
            {
                const float *p = tris + 9 * (size_t)h.prim;
                const Vec p0 = mk(p[0], p[1], p[2]), p1 = mk(p[3], p[4], p[5]), p2 = mk(p[6], p[7], p[8]);
                const Vec n = nrm(cross(sub(p0, p1), sub(p2, p0)));
                px[0] = std::fabs(n.x);
                px[1] = std::fabs(n.y);
                px[2] = std::fabs(n.z);
            }
            uint8_t *o = out + 3 * ((size_t)width * (height - 1 - j) + i);
            for(int c = 0; c < 3; ++c)
                o[c] = (uint8_t)std::max(std::min(px[c] * 255.0f, 255.0f), 0.0f);
        }
}

// Halton index of (pixel, sample) for the sampler tests (halton.cpp:46-66)
extern "C" uint64_t oracle_halton_index(int32_t w, int32_t h, uint32_t px, uint32_t py, int32_t sample)
{
    orc::Halton hs;
    hs.init(w, h, px, py, sample);
    return hs.idx;
}
