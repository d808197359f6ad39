// pt_obj.cpp — Wavefront OBJ triangles as the reference's Mesh::Load receives them.
//
// Mesh::Load (Core/Geometry/Mesh.cpp:18-125) reads a file through Assimp 4.1.0 (the
// reference's pinned dependency, conanfile.txt:2) with no post-processing flags, copies
// every vertex (position, normalised normal) and makes one triangle of the first three
// indices of every face. Assimp is not linked here; what its OBJ importer hands
// over for such a call is restated here:
//  - one output vertex per face corner, faces in file order (objects, groups and
//    material changes split meshes but keep that order);
//  - no triangulation, so a polygon keeps its first three corners;
//  - indices are 1-based, negative ones count back from the elements read so far;
//  - "v" takes 3 components, 4 (divided by w) or 6 (position and colour);
//  - "vt" gives a corner's texture coordinates (u, v; zero without), Vertex::uv;
//  - a corner's normal only when its face gives normal indices, else the Vertex default
//    +Z (Core/Geometry/Vertex.h:19); Assimp fails the whole import (no triangles) when
//    faces of one mesh mix both, which this rejects as an error;
//  - numbers are read with Assimp's fast_atof, which is not correctly rounded: the
//    integer part converted to float, plus the first 15 fractional digits as an integer
//    times 10^-digits in double, rounded to float and added, then an exponent applied
//    as powf(10, e) (a ',' also separates the fraction).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/petrichor_hip.h"

namespace {

struct Cursor {
    const char* p;
    const char* end;
};

bool is_digit(char c) { return c >= '0' && c <= '9'; }

// strtoul10_64 of Assimp's fast_atof.h: at most max_digits digits are accumulated, the
// rest skipped; *digits receives how many were accumulated.
uint64_t read_digits(Cursor& c, unsigned max_digits, unsigned* digits)
{
    uint64_t v = 0;
    unsigned n = 0;
    while (c.p < c.end && is_digit(*c.p)) {
        if (n < max_digits) {
            v = v * 10 + (uint64_t)(*c.p - '0');
            ++n;
        }
        ++c.p;
    }
    if (digits) *digits = n;
    return v;
}

// fast_atoreal_move<float>: false when the text does not start a number.
bool read_float(Cursor& c, float* out)
{
    static const double kNegPow10[16] = {1.0,   0.1,   0.01,  0.001, 1e-4,  1e-5,  1e-6,  1e-7,
                                         1e-8,  1e-9,  1e-10, 1e-11, 1e-12, 1e-13, 1e-14, 1e-15};
    const char* p = c.p;
    const bool neg = p < c.end && *p == '-';
    if (p < c.end && (*p == '-' || *p == '+')) ++p;
    auto starts = [&](const char* s) {
        for (size_t i = 0; s[i]; ++i)
            if (p + i >= c.end || (p[i] | 0x20) != s[i]) return false;
        return true;
    };
    if (starts("nan")) {
        c.p = p + 3;
        *out = neg ? -NAN : NAN;
        return true;
    }
    if (starts("inf")) {
        c.p = p + 3;
        if (starts("infinity")) c.p = p + 8;
        *out = neg ? -INFINITY : INFINITY;
        return true;
    }
    const bool frac_start = p + 1 < c.end && (*p == '.' || *p == ',') && is_digit(p[1]);
    if (!(p < c.end && (is_digit(*p) || frac_start))) return false;
    c.p = p;
    float f = 0.0f;
    if (*c.p != '.' && *c.p != ',') f = (float)read_digits(c, 64, nullptr);
    if (c.p + 1 < c.end && (*c.p == '.' || *c.p == ',') && is_digit(c.p[1])) {
        ++c.p;
        unsigned n = 0;
        double pl = (double)read_digits(c, 15, &n);
        pl *= kNegPow10[n];
        f += (float)pl;
    } else if (c.p < c.end && *c.p == '.') {
        ++c.p;
    }
    if (c.p < c.end && (*c.p == 'e' || *c.p == 'E')) {
        ++c.p;
        const bool eneg = c.p < c.end && *c.p == '-';
        if (c.p < c.end && (*c.p == '-' || *c.p == '+')) ++c.p;
        float e = (float)read_digits(c, 64, nullptr);
        if (eneg) e = -e;
        f *= powf(10.0f, e);
    }
    *out = neg ? -f : f;
    return true;
}

void skip_blank(Cursor& c)
{
    while (c.p < c.end && (*c.p == ' ' || *c.p == '\t' || *c.p == '\r')) ++c.p;
}

// atoi of a face field: optional sign and digits (0 when there are none). Values beyond
// 10^12 saturate there (they are out of range for any index anyway).
long read_int(Cursor& c)
{
    bool neg = false;
    if (c.p < c.end && (*c.p == '-' || *c.p == '+')) neg = *c.p++ == '-';
    long v = 0;
    while (c.p < c.end && is_digit(*c.p)) {
        v = v < 1000000000000L ? v * 10 + (*c.p - '0') : v;
        ++c.p;
    }
    return neg ? -v : v;
}

struct Corner {
    long v, t, n; // position, texture coordinate and normal index; -1: none
};

} // namespace

extern "C" int pt_obj_parse(const char* text, int64_t len, int64_t max_triangles, float* positions, float* normals,
                            float* uvs, int64_t* num_triangles, char* err, int32_t err_len)
{
    auto fail = [&](int code, const std::string& msg) {
        if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", msg.c_str());
        return code;
    };
    if (!text || len < 0 || !num_triangles || max_triangles < 0 ||
        (max_triangles > 0 && (!positions || !normals)))
        return fail(PT_ERR_ARG, "bad OBJ arguments");
    std::vector<float> pos, nrm, tex; // 3, 3 and 2 per element
    int64_t count = 0;
    int with_normals = -1; // faces so far: -1 none, 0 without normals, 1 with
    Cursor c{text, text + len};
    long line = 0;
    while (c.p < c.end) {
        ++line;
        const char* eol = (const char*)std::memchr(c.p, '\n', (size_t)(c.end - c.p));
        if (!eol) eol = c.end;
        Cursor l{c.p, eol};
        c.p = eol < c.end ? eol + 1 : eol;
        skip_blank(l);
        if (l.p == l.end || *l.p == '#') continue;
        const char* tag = l.p;
        while (l.p < l.end && *l.p != ' ' && *l.p != '\t') ++l.p;
        const std::string t(tag, l.p);
        if (t == "vt") {
            float v[3] = {0.0f, 0.0f, 0.0f};
            int k = 0;
            for (;;) {
                skip_blank(l);
                if (l.p == l.end || *l.p == '#') break;
                if (k == 3 || !read_float(l, &v[k])) return fail(PT_ERR_ARG, "OBJ line " + std::to_string(line) + ": bad number");
                ++k;
            }
            if (k < 1) return fail(PT_ERR_ARG, "OBJ line " + std::to_string(line) + ": a texture coordinate needs a number");
            tex.insert(tex.end(), v, v + 2);
        } else if (t == "v" || t == "vn") {
            float v[6];
            int k = 0;
            for (;;) {
                skip_blank(l);
                if (l.p == l.end || *l.p == '#') break;
                if (k == 6 || !read_float(l, &v[k])) return fail(PT_ERR_ARG, "OBJ line " + std::to_string(line) + ": bad number");
                ++k;
            }
            if (t == "vn") {
                if (k < 3) return fail(PT_ERR_ARG, "OBJ line " + std::to_string(line) + ": a normal needs 3 numbers");
                nrm.insert(nrm.end(), v, v + 3);
            } else if (k == 3 || k == 6) {
                pos.insert(pos.end(), v, v + 3);
            } else if (k == 4) {
                if (v[3] == 0.0f) return fail(PT_ERR_ARG, "OBJ line " + std::to_string(line) + ": w = 0");
                const float p3[3] = {v[0] / v[3], v[1] / v[3], v[2] / v[3]};
                pos.insert(pos.end(), p3, p3 + 3);
            } else {
                return fail(PT_ERR_ARG, "OBJ line " + std::to_string(line) + ": a vertex needs 3, 4 or 6 numbers");
            }
        } else if (t == "f") {
            Corner corners[3];
            int k = 0;
            bool has_n = false, no_n = false;
            const long nv = (long)(pos.size() / 3), nn = (long)(nrm.size() / 3), nt = (long)(tex.size() / 2);
            for (;;) {
                skip_blank(l);
                if (l.p == l.end || *l.p == '#') break;
                long fields[3] = {0, 0, 0};
                for (int s = 0; s < 3; ++s) {
                    fields[s] = read_int(l);
                    if (l.p < l.end && *l.p == '/') {
                        ++l.p;
                    } else {
                        break;
                    }
                }
                while (l.p < l.end && *l.p != ' ' && *l.p != '\t' && *l.p != '\r') ++l.p;
                const long vi = fields[0] > 0 ? fields[0] - 1 : nv + fields[0];
                const long ti = fields[1] > 0 ? fields[1] - 1 : (fields[1] < 0 ? nt + fields[1] : -1);
                const long ni = fields[2] > 0 ? fields[2] - 1 : (fields[2] < 0 ? nn + fields[2] : -1);
                if (fields[0] == 0 || vi < 0 || vi >= nv || (ni >= 0 && ni >= nn) || (fields[2] < 0 && ni < 0) ||
                    (ti >= 0 && ti >= nt) || (fields[1] < 0 && ti < 0))
                    return fail(PT_ERR_ARG, "OBJ line " + std::to_string(line) + ": index out of range");
                (ni >= 0 ? has_n : no_n) = true;
                if (k < 3) corners[k] = Corner{vi, ti, ni};
                ++k;
            }
            if (k < 3) return fail(PT_ERR_ARG, "OBJ line " + std::to_string(line) + ": a face needs three corners");
            if (has_n && no_n) return fail(PT_ERR_ARG, "OBJ line " + std::to_string(line) + ": a face mixes corners with and without normals");
            if (with_normals >= 0 && with_normals != (int)has_n)
                return fail(PT_ERR_ARG, "OBJ line " + std::to_string(line) + ": faces with and without normals in one file");
            with_normals = has_n;
            if (count < max_triangles) {
                float* P = positions + 9 * count;
                float* N = normals + 9 * count;
                for (int j = 0; j < 3; ++j) {
                    for (int a = 0; a < 3; ++a) P[3 * j + a] = pos[3 * corners[j].v + a];
                    if (uvs) {
                        uvs[6 * count + 2 * j] = corners[j].t >= 0 ? tex[2 * corners[j].t] : 0.0f;
                        uvs[6 * count + 2 * j + 1] = corners[j].t >= 0 ? tex[2 * corners[j].t + 1] : 0.0f;
                    }
                    if (corners[j].n >= 0) {
                        // Vector3f::Normalized: v / sqrt(dot(v, v)) (Math/Vector3f.h:170-189)
                        const float x = nrm[3 * corners[j].n], y = nrm[3 * corners[j].n + 1], z = nrm[3 * corners[j].n + 2];
                        const float d = x * x + y * y + z * z;
                        const float l2 = (float)std::sqrt((double)d);
                        N[3 * j] = x / l2;
                        N[3 * j + 1] = y / l2;
                        N[3 * j + 2] = z / l2;
                    } else {
                        N[3 * j] = 0.0f;
                        N[3 * j + 1] = 0.0f;
                        N[3 * j + 2] = 1.0f;
                    }
                }
            }
            ++count;
        } else if (t == "l" || t == "p") {
            return fail(PT_ERR_UNSUPPORTED, "OBJ line " + std::to_string(line) + ": point and line elements are not triangles");
        }
        // o, g, s, usemtl, mtllib and others do not change the triangles
    }
    *num_triangles = count;
    return PT_OK;
}
