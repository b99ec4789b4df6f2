// cih_scene.cpp -- seeded synthetic workloads as command streams.
//
// The reference ships no workload generator; these scenes are the five
// benchmark configurations (plain-512, large-2048, effects-1024, batch and
// serve) written as command streams (cih_stream.h) so that exactly the same
// calls replay into the HIP canvas, the CPU oracle and the reference.
//
// PRNG: SplitMix64 seeds xoshiro128**; a float draw is the top 24 bits of a
// 32-bit output scaled by 2^-24, i.e. uniform on [0,1).
#include "cih_abi.h"
#include "cih_stream.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct rng {
    uint32_t s[4];
    explicit rng(uint64_t seed) {
        for (int k = 0; k < 2; ++k) {
            uint64_t z = (seed += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            s[2 * k] = (uint32_t)z;
            s[2 * k + 1] = (uint32_t)(z >> 32);
        }
        if (!(s[0] | s[1] | s[2] | s[3]))
            s[0] = 1;
    }
    static uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
    uint32_t next() {
        uint32_t result = rotl(s[1] * 5, 7) * 9;
        uint32_t t = s[1] << 9;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 11);
        return result;
    }
    float u() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
    float u(float lo, float hi) { return lo + (hi - lo) * u(); }
    int range(int lo, int hi) { return lo + (int)(next() % (uint32_t)(hi - lo + 1)); }
    float log_uniform(float lo, float hi) {
        return std::exp(std::log(lo) + (std::log(hi) - std::log(lo)) * u());
    }
};

const float kTau = 6.28318531f;
const int kOps[11] = {1, 2, 3, 4, 7, 10, 11, 12, 13, 14, 15};

struct writer {
    std::vector<uint32_t> w;
    writer(int width, int height) {
        w.push_back(CIH_STREAM_MAGIC);
        w.push_back(CIH_STREAM_VERSION);
        w.push_back((uint32_t)width);
        w.push_back((uint32_t)height);
    }
    void op(uint32_t code) { w.push_back(code); }
    void f(float v) { w.push_back(cih_f2w(v)); }
    void i(int v) { w.push_back((uint32_t)v); }
    void op_f(uint32_t code, int n, const float *v) {
        op(code);
        for (int k = 0; k < n; ++k)
            f(v[k]);
    }
    void xy(uint32_t code, float x, float y) {
        float v[2] = {x, y};
        op_f(code, 2, v);
    }
    void color(uint32_t code, int type, float r, float g, float b, float a) {
        op(code);
        i(type);
        f(r), f(g), f(b), f(a);
    }
    void image(int id, int width, int height, const std::vector<unsigned char> &texels) {
        op(CIH_OP_DEFINE_IMAGE);
        i(id), i(width), i(height);
        size_t base = w.size();
        w.resize(base + (size_t)width * (size_t)height);
        std::memcpy(&w[base], &texels[0], texels.size());
    }
    void identity() {
        float v[6] = {1, 0, 0, 1, 0, 0};
        op_f(CIH_OP_SET_TRANSFORM, 6, v);
    }
};

// A random sRGB colour with alpha in [0.3, 1).
void random_color(writer &out, rng &r, uint32_t code, int type) {
    float red = r.u(), green = r.u(), blue = r.u(), alpha = r.u(0.3f, 1.0f);
    out.color(code, type, red, green, blue, alpha);
}

// Plain-512 fill: a closed path of 4-12 cubic Beziers with every point
// uniform over the canvas, solid colour, non-zero fill.
void c1_fill(writer &out, rng &r, float width, float height) {
    out.op(CIH_OP_BEGIN_PATH);
    out.xy(CIH_OP_MOVE_TO, r.u() * width, r.u() * height);
    int k = r.range(4, 12);
    for (int j = 0; j < k; ++j) {
        float v[6];
        for (int m = 0; m < 3; ++m) {
            v[2 * m] = r.u() * width;
            v[2 * m + 1] = r.u() * height;
        }
        out.op_f(CIH_OP_BEZIER_CURVE_TO, 6, v);
    }
    out.op(CIH_OP_CLOSE_PATH);
    random_color(out, r, CIH_OP_SET_COLOR, 0);
    out.op(CIH_OP_FILL);
}

// Plain-512 stroke: an open polyline of 3-20 uniform points, width U(1,8),
// round joins and caps.
void c1_stroke(writer &out, rng &r, float width, float height) {
    out.op(CIH_OP_BEGIN_PATH);
    int n = r.range(3, 20);
    out.xy(CIH_OP_MOVE_TO, r.u() * width, r.u() * height);
    for (int j = 1; j < n; ++j)
        out.xy(CIH_OP_LINE_TO, r.u() * width, r.u() * height);
    float lw = r.u(1.0f, 8.0f);
    out.op_f(CIH_OP_SET_LINE_WIDTH, 1, &lw);
    out.op(CIH_OP_SET_LINE_JOIN), out.i(2);
    out.op(CIH_OP_SET_LINE_CAP), out.i(2);
    random_color(out, r, CIH_OP_SET_COLOR, 1);
    out.op(CIH_OP_STROKE);
}

void c1_mix_path(writer &out, rng &r, float width, float height) {
    if (r.u() < 0.8f)
        c1_fill(out, r, width, height);
    else
        c1_stroke(out, r, width, height);
}

void random_shadow(writer &out, rng &r) {
    float blur = r.u(2.0f, 32.0f);
    out.op_f(CIH_OP_SET_SHADOW_BLUR, 1, &blur);
    out.xy(CIH_OP_SET_SHADOW_OFFSET, r.u(-20.0f, 20.0f), r.u(-20.0f, 20.0f));
    float v[4] = {r.u(), r.u(), r.u(), r.u(0.3f, 1.0f)};
    out.op_f(CIH_OP_SET_SHADOW_COLOR, 4, v);
}

// Large-2048 brush: solid, linear or radial with 2-8 stops.
void c2_brush(writer &out, rng &r, int type, float width, float height) {
    int kind = r.range(0, 2);
    if (kind == 0) {
        random_color(out, r, CIH_OP_SET_COLOR, type);
        return;
    }
    if (kind == 1) {
        out.op(CIH_OP_SET_LINEAR_GRADIENT), out.i(type);
        out.f(r.u() * width), out.f(r.u() * height);
        out.f(r.u() * width), out.f(r.u() * height);
    } else {
        float cx = r.u() * width, cy = r.u() * height;
        out.op(CIH_OP_SET_RADIAL_GRADIENT), out.i(type);
        out.f(cx), out.f(cy), out.f(r.u(0.0f, 64.0f));
        out.f(cx + r.u(-64.0f, 64.0f)), out.f(cy + r.u(-64.0f, 64.0f));
        out.f(r.u(64.0f, 512.0f));
    }
    int stops = r.range(2, 8);
    for (int k = 0; k < stops; ++k) {
        out.op(CIH_OP_ADD_COLOR_STOP), out.i(type);
        out.f(r.u()), out.f(r.u()), out.f(r.u()), out.f(r.u()), out.f(r.u(0.3f, 1.0f));
    }
}

void c2_shape(writer &out, rng &r, float width, float height, bool stroke) {
    out.op(CIH_OP_BEGIN_PATH);
    int kind = r.range(0, 3);
    if (kind <= 1) {
        out.xy(CIH_OP_MOVE_TO, r.u() * width, r.u() * height);
        int k = r.range(4, 12);
        for (int j = 0; j < k; ++j) {
            float v[6];
            int n = kind == 0 ? 6 : 4;
            for (int m = 0; m < n / 2; ++m) {
                v[2 * m] = r.u() * width;
                v[2 * m + 1] = r.u() * height;
            }
            out.op_f(kind == 0 ? CIH_OP_BEZIER_CURVE_TO : CIH_OP_QUADRATIC_CURVE_TO, n, v);
        }
        if (!stroke || r.u() < 0.5f)
            out.op(CIH_OP_CLOSE_PATH);
    } else if (kind == 2) {
        float start = r.u() * kTau;
        float v[5] = {r.u() * width, r.u() * height, r.u(4.0f, 256.0f), start,
                      start + r.u() * kTau};
        out.op_f(CIH_OP_ARC, 5, v);
        out.i(r.u() < 0.5f ? 1 : 0);
    } else {
        float v[4] = {r.u() * width, r.u() * height, r.u(4.0f, 512.0f), r.u(4.0f, 512.0f)};
        out.op_f(CIH_OP_RECTANGLE, 4, v);
    }
}

void c2_draw(writer &out, rng &r, float width, float height, float evenodd) {
    bool stroke = r.u() >= 0.6f;
    c2_shape(out, r, width, height, stroke);
    c2_brush(out, r, stroke ? 1 : 0, width, height);
    float alpha = r.u(0.5f, 1.0f);
    out.op_f(CIH_OP_SET_GLOBAL_ALPHA, 1, &alpha);
    if (stroke) {
        float lw = r.u(1.0f, 8.0f);
        out.op_f(CIH_OP_SET_LINE_WIDTH, 1, &lw);
        out.op(CIH_OP_SET_LINE_JOIN), out.i(r.range(0, 2));
        out.op(CIH_OP_SET_LINE_CAP), out.i(r.range(0, 2));
        float limit = r.u(1.0f, 10.0f);
        out.op_f(CIH_OP_SET_MITER_LIMIT, 1, &limit);
        if (r.u() < 0.5f) {
            int n = r.range(2, 6);
            out.op(CIH_OP_SET_LINE_DASH), out.i(n);
            for (int k = 0; k < n; ++k)
                out.f(r.u(2.0f, 40.0f));
            float offset = r.u(0.0f, 40.0f);
            out.op_f(CIH_OP_SET_LINE_DASH_OFFSET, 1, &offset);
        } else {
            out.op(CIH_OP_SET_LINE_DASH), out.i(-1);
        }
        out.op(CIH_OP_STROKE);
    } else {
        if (evenodd > 0.0f) {
            out.op(CIH_OP_SET_FILL_RULE), out.i(r.u() < evenodd ? 1 : 0);
        }
        out.op(CIH_OP_FILL);
    }
}

void define_images(writer &out, rng &r) {
    std::vector<unsigned char> noise(256 * 256 * 4), checker(256 * 256 * 4);
    for (size_t k = 0; k < noise.size(); ++k)
        noise[k] = (unsigned char)(r.next() >> 24);
    unsigned char colors[2][4];
    for (int c = 0; c < 2; ++c)
        for (int k = 0; k < 4; ++k)
            colors[c][k] = k == 3 ? 255 : (unsigned char)(r.next() >> 24);
    for (int y = 0; y < 256; ++y)
        for (int x = 0; x < 256; ++x)
            std::memcpy(&checker[4 * (size_t)(y * 256 + x)], colors[((x >> 3) ^ (y >> 3)) & 1], 4);
    out.image(0, 256, 256, noise);
    out.image(1, 256, 256, checker);
}

void c3_image(writer &out, rng &r, float width, float height) {
    out.identity();
    out.xy(CIH_OP_TRANSLATE, r.u() * width, r.u() * height);
    float angle = r.u() * kTau;
    out.op_f(CIH_OP_ROTATE, 1, &angle);
    float side = 256.0f * r.u(0.3f, 3.0f);
    out.op(CIH_OP_DRAW_IMAGE), out.i(r.u() < 0.5f ? 0 : 1);
    out.f(-0.5f * side), out.f(-0.5f * side), out.f(side), out.f(side);
    out.identity();
}

int clamp_size(float v, int lo, int hi) {
    int s = (int)(v + 0.5f);
    return s < lo ? lo : s > hi ? hi : s;
}

} // namespace

extern "C" int cih_scene_generate(int config, uint64_t seed, int width, int height,
                                  int shapes, int flags, void **out, size_t *bytes)
{
    if (!out || !bytes)
        return CIH_ERROR_INVALID;
    *out = 0;
    *bytes = 0;
    rng r(seed);
    if (config == 4 || config == 5) {
        float lo = config == 4 ? 64.0f : 128.0f, hi = config == 4 ? 1024.0f : 2048.0f;
        int w = clamp_size(r.log_uniform(lo, hi), (int)lo, (int)hi);
        int h = clamp_size(r.log_uniform(lo, hi), (int)lo, (int)hi);
        if (!width)
            width = w;
        if (!height)
            height = h;
    }
    static const int default_size[6] = {0, 512, 2048, 1024, 0, 0};
    if (config < 1 || config > 5)
        return CIH_ERROR_INVALID;
    if (!width)
        width = default_size[config];
    if (!height)
        height = default_size[config];
    if (width < 1 || height < 1 || width > 32768 || height > 32768 || shapes < 0)
        return CIH_ERROR_INVALID;
    writer w(width, height);
    float fw = (float)width, fh = (float)height;
    if (config == 1) {
        int fills = shapes ? shapes - shapes / 5 : 200;
        int strokes = shapes ? shapes / 5 : 50;
        for (int k = 0; k < fills; ++k)
            c1_fill(w, r, fw, fh);
        for (int k = 0; k < strokes; ++k)
            c1_stroke(w, r, fw, fh);
    } else if (config == 2) {
        int n = shapes ? shapes : 5000;
        float evenodd = (flags & CIH_SCENE_EVENODD) ? 0.25f : 0.0f;
        for (int k = 0; k < n; ++k)
            c2_draw(w, r, fw, fh, evenodd);
    } else if (config == 3) {
        int paths = shapes ? shapes : 500;
        int images = (flags & CIH_SCENE_NO_IMAGES) ? 0 : (shapes ? (shapes + 4) / 5 : 100);
        if (images)
            define_images(w, r);
        std::vector<int> order((size_t)(paths + images), 0);
        for (int k = 0; k < images; ++k)
            order[(size_t)k] = 1;
        for (size_t k = order.size(); k > 1; --k) // Fisher-Yates interleave
            std::swap(order[k - 1], order[(size_t)(r.next() % (uint32_t)k)]);
        for (size_t k = 0; k < order.size(); ++k) {
            w.op(CIH_OP_SET_COMPOSITE_OP), w.i(kOps[r.range(0, 10)]);
            if (order[k])
                c3_image(w, r, fw, fh);
            else {
                random_shadow(w, r);
                c1_mix_path(w, r, fw, fh);
            }
        }
    } else if (config == 4) {
        int n = shapes ? shapes : r.range(10, 400);
        for (int k = 0; k < n; ++k)
            c1_mix_path(w, r, fw, fh);
    } else {
        int n = shapes ? shapes : r.range(50, 2000);
        bool shadows = (flags & CIH_SCENE_SHADOWS) || r.u() < 0.1f;
        int images = 0;
        if (shadows && !(flags & CIH_SCENE_NO_IMAGES)) {
            define_images(w, r);
            images = 4;
        }
        for (int k = 0; k < n; ++k) {
            if (images && r.u() < 0.01f) {
                c3_image(w, r, fw, fh);
                continue;
            }
            if (shadows)
                random_shadow(w, r);
            c1_mix_path(w, r, fw, fh);
        }
    }
    size_t size = w.w.size() * 4;
    void *buffer = std::malloc(size);
    if (!buffer)
        return CIH_ERROR_OUT_OF_MEMORY;
    std::memcpy(buffer, &w.w[0], size);
    *out = buffer;
    *bytes = size;
    return CIH_OK;
}

extern "C" uint64_t cih_scene_seed(uint64_t seed, uint64_t index)
{
    return seed ^ (index * 0x9E3779B97F4A7C15ull);
}

extern "C" void cih_free(void *pointer) { std::free(pointer); }
