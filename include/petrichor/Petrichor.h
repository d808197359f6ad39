// Petrichor.h — C++ host API of the MI355X render path with the names and behaviour of
// LibPetrichor Core: Math::Vector3f, Scene, Camera, the materials, Texture2D,
// Environment, RenderSetting and Petrichor::Render. Host code written against the
// reference builds against this header; Render flattens the scene and calls the C-ABI
// of petrichor_hip.h (libpetrichor_hip.so).
//
// Only what the GPU path renders is here: triangles (flat or smooth), spheres, Lambert, GGX,
// Glass, Emission and MixMaterial with their image textures (Lambert albedo, GGX
// reflectance, roughness and normal maps), an equirect environment given as a Texture2D, the MIS
// and simple integrators, the denoising albedo / normal and UV AOVs, Texture2D::Save (PNG
// bytes converted on the GPU, Radiance HDR), the reference's file loading (JSON scene and
// render-setting files, OBJ meshes, PNG textures and Radiance .hdr environment maps), and
// the 2D samplers (RandomSampler2D, MicroJitteredSampler) with XorShift128 and Halton.
//
// Floating point: the camera set-up below must be compiled without FMA contraction
// (-ffp-contract=off) to give the reference's camera state bit for bit.
#pragma once

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../petrichor_hip.h"

namespace Petrichor
{
namespace Math
{

// Math/Vector3f.h: float components, operators evaluated left to right.
struct Vector3f
{
    float x = 0.0f, y = 0.0f, z = 0.0f;

    constexpr Vector3f() = default;
    constexpr Vector3f(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}

    constexpr Vector3f operator+(const Vector3f& v) const { return {x + v.x, y + v.y, z + v.z}; }
    constexpr Vector3f operator-(const Vector3f& v) const { return {x - v.x, y - v.y, z - v.z}; }
    constexpr Vector3f operator-() const { return {-x, -y, -z}; }
    constexpr Vector3f operator*(float c) const { return {x * c, y * c, z * c}; }
    constexpr Vector3f operator*(const Vector3f& v) const { return {x * v.x, y * v.y, z * v.z}; }
    constexpr Vector3f operator/(float c) const { return {x / c, y / c, z / c}; }
    friend constexpr Vector3f operator*(float c, const Vector3f& v) { return v * c; }

    float SquaredLength() const { return x * x + y * y + z * z; }
    // sqrt of a float: the correctly rounded root, as the reference's sqrt(Dot(...))
    float Length() const { return std::sqrt(SquaredLength()); }
    Vector3f Normalized() const { return *this / Length(); }

    static constexpr Vector3f Zero() { return {0.0f, 0.0f, 0.0f}; }
    static constexpr Vector3f One() { return {1.0f, 1.0f, 1.0f}; }
    static constexpr Vector3f UnitX() { return {1.0f, 0.0f, 0.0f}; }
    static constexpr Vector3f UnitY() { return {0.0f, 1.0f, 0.0f}; }
    static constexpr Vector3f UnitZ() { return {0.0f, 0.0f, 1.0f}; }
};

inline float Dot(const Vector3f& a, const Vector3f& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

inline Vector3f Cross(const Vector3f& a, const Vector3f& b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Random/XorShift.h: Marsaglia's xorshift128 with the reference's default x, y, z words and
// the seed as w; next() is random() / (float)max().
class XorShift128
{
public:
    explicit XorShift128(uint32_t seed = 88675123u) : w(seed) {}
    uint32_t random()
    {
        const uint32_t t = x ^ (x << 11);
        x = y;
        y = z;
        z = w;
        return w = (w ^ (w >> 19)) ^ (t ^ (t >> 8));
    }
    static constexpr uint32_t max() { return 0xFFFFFFFFu; }
    float next() { return (float)random() / static_cast<float>(max()); }

private:
    uint32_t x = 123456789u, y = 362436069u, z = 521288629u, w;
};

// Math/Halton.h: the radical-inverse sequence in one base, tabulated by Initialize with the
// reference's float steps (Math/Halton.cpp:18-37).
class Halton
{
public:
    void Initialize(size_t length, unsigned base)
    {
        m_values.assign(length, 0.0f);
        for (size_t i = 0; i < length; ++i) {
            unsigned rest = (unsigned)i;
            float f = 1.0f, r = 0.0f;
            for (; rest > 0; rest /= base) {
                f /= (float)base;
                r += f * (float)(rest % base);
            }
            m_values[i] = r;
        }
        m_isInitialized = true;
    }
    float GetValue(unsigned index) const { return m_values.at(index); }
    size_t GetLength() const { return m_values.size(); }
    bool IsInitialized() const { return m_isInitialized; }

private:
    std::vector<float> m_values;
    bool m_isInitialized = false;
};

} // namespace Math

using Color3f = Math::Vector3f;

struct RenderingResult
{
    float totalSec;
};

namespace Core
{

// Core/Sampler/ISampler2D.h and its two implementations. The renders here draw their 2D
// samples on the device; these host classes give the same sequences to host code.
class ISampler2D
{
public:
    virtual ~ISampler2D() = default;
    virtual std::tuple<float, float> Next() = 0;
};

// Core/Sampler/RandomSampler2D.cpp:8-21: one stream per coordinate.
class RandomSampler2D : public ISampler2D
{
public:
    RandomSampler2D(uint32_t seedX, uint32_t seedY) : m_x(seedX), m_y(seedY) {}
    std::tuple<float, float> Next() override
    {
        const float a = m_x.next();
        return {a, m_y.next()};
    }

private:
    Math::XorShift128 m_x, m_y;
};

// Core/Sampler/MicroJitteredSampler.cpp:10-48: Halton points in bases 2 and 3 plus offsets
// that are redrawn (mu * (next() - 0.5), mu = 2.5 / sqrt(length)) after every pass over
// the sequence, each coordinate taken modulo 1. Petrichor::SetMicroJitteredSampler renders
// with it on the device.
class MicroJitteredSampler : public ISampler2D
{
public:
    explicit MicroJitteredSampler(unsigned seed) : m_xorShift(seed) {}
    void Initialize(int lengthSeq)
    {
        m_lengthSeq = (size_t)lengthSeq;
        // the reference's unqualified sqrt is the double C function
        m_mu = (float)((double)kStarDiscrepancy / std::sqrt((double)static_cast<float>(lengthSeq)));
        m_haltonBase2.Initialize(m_lengthSeq, 2);
        m_haltonBase3.Initialize(m_lengthSeq, 3);
    }
    std::tuple<float, float> Next() override
    {
        float x = m_haltonBase2.GetValue(m_index) + m_offsetX;
        float y = m_haltonBase3.GetValue(m_index) + m_offsetY;
        x = x - 1.0f * std::floor(x / 1.0f); // Math::Mod(x, 1.0f)
        y = y - 1.0f * std::floor(y / 1.0f);
        if (++m_index >= m_lengthSeq) {
            m_offsetX = m_mu * (m_xorShift.next() - 0.5f);
            m_offsetY = m_mu * (m_xorShift.next() - 0.5f);
            m_index = 0;
        }
        return {x, y};
    }
    unsigned GetIndex() const { return m_index; }

private:
    Math::Halton m_haltonBase2, m_haltonBase3;
    Math::XorShift128 m_xorShift;
    unsigned m_index = 0;
    size_t m_lengthSeq = 0;
    float m_mu = 0.0f;
    float m_offsetX = 0.0f, m_offsetY = 0.0f;
    static constexpr float kStarDiscrepancy = 2.5f;
};

// Core/Material/MaterialBase.h:15-22 (the order of the C-ABI's PT_MAT_* codes).
enum class MaterialTypes
{
    Lambert,
    Glass,
    Emission,
    Glossy,
    Mix
};

class MaterialBase
{
public:
    virtual ~MaterialBase() = default;
    virtual MaterialTypes GetMaterialType() const = 0;
    void EnableImportanceSampling(bool isEnabled) { m_isImportanceSamplingEnabled = isEnabled; }
    bool IsImportanceSamplingEnabled() const { return m_isImportanceSamplingEnabled; }

private:
    bool m_isImportanceSamplingEnabled = true;
};

class Texture2D;

// Core/Material/Lambert.h:13-49. With an albedo texture, GetAlbedo is kd x the texture at
// the hit's uv (so the BxDF is kd x kd x texture / pi); without, 1.
class Lambert : public MaterialBase
{
public:
    explicit Lambert(const Color3f& kd) : m_kd(kd) {}
    Lambert(const Color3f& kd, const Texture2D& albedoTex) : m_kd(kd), m_texAlbedo(&albedoTex) {}
    MaterialTypes GetMaterialType() const override { return MaterialTypes::Lambert; }
    const Color3f& GetKd() const { return m_kd; }
    void SetTexAlbedo(const Texture2D* texAlbedo) { m_texAlbedo = texAlbedo; }
    const Texture2D* GetTexAlbedo() const { return m_texAlbedo; }

private:
    Color3f m_kd;
    const Texture2D* m_texAlbedo = nullptr;
};

// Core/Material/GGX.h:13-164 with metalness 0. A reflectance (F0) texture multiplies the
// reflectance; a roughness map replaces the roughness by strength x the luminance of the
// map; a normal map replaces the normal the GGX functions use by GetNormal's, in the frame
// of the hit's tangent (GGX.h:84-119).
class GGX : public MaterialBase
{
public:
    GGX(const Color3f& f0, float roughness = 1.0f) : m_f0(f0), m_roughness(roughness) {}
    MaterialTypes GetMaterialType() const override { return MaterialTypes::Glossy; }
    const Color3f& GetReflectance() const { return m_f0; }
    float GetRoughness() const { return m_roughness; }
    void SetF0Texture(const Texture2D* texture) { m_reflectanceMap = texture; }
    void SetRoughnessMap(const Texture2D* roughnessMap) { m_roughnessMap = roughnessMap; }
    void SetNormalMap(const Texture2D* normalMap) { m_normalMap = normalMap; }
    void SetNormalMapStrength(float strength) { m_normalMapStrength = std::clamp(strength, 0.0f, 1.0f); }
    void SetRoughnessMapStrength(float strength) { m_roughnessMapStrength = strength; }
    const Texture2D* GetF0Texture() const { return m_reflectanceMap; }
    const Texture2D* GetRoughnessMap() const { return m_roughnessMap; }
    const Texture2D* GetNormalMap() const { return m_normalMap; }
    float GetNormalMapStrength() const { return m_normalMapStrength; }
    float GetRoughnessMapStrength() const { return m_roughnessMapStrength; }

private:
    Color3f m_f0;
    float m_roughness;
    float m_normalMapStrength = 1.0f;
    float m_roughnessMapStrength = 1.0f;
    const Texture2D* m_reflectanceMap = nullptr;
    const Texture2D* m_roughnessMap = nullptr;
    const Texture2D* m_normalMap = nullptr;
};

class Glass : public MaterialBase
{
public:
    Glass() = default;
    Glass(const Color3f& color, float ior) noexcept : m_color(color), m_ior(ior) {}
    MaterialTypes GetMaterialType() const override { return MaterialTypes::Glass; }
    const Color3f& GetColor() const { return m_color; }
    float GetIor() const { return m_ior; }

private:
    Color3f m_color = Color3f::One();
    float m_ior = 1.333f; // IOR::Water
};

class Emission : public MaterialBase
{
public:
    Emission(const Color3f& color) : m_color(color) {}
    MaterialTypes GetMaterialType() const override { return MaterialTypes::Emission; }
    const Color3f& GetLightColor() const { return m_color; }

private:
    Color3f m_color;
};

class MixMaterial : public MaterialBase
{
public:
    MixMaterial(const MaterialBase* mat0, const MaterialBase* mat1, float mix) : m_mat0(mat0), m_mat1(mat1), m_mix(mix)
    {
    }
    MaterialTypes GetMaterialType() const override { return MaterialTypes::Mix; }
    // MixMaterial.h:45-49
    const MaterialBase* GetSingleMaterial(float rand) const { return m_mix <= rand ? m_mat0 : m_mat1; }
    const MaterialBase* GetMaterial0() const { return m_mat0; }
    const MaterialBase* GetMaterial1() const { return m_mat1; }
    float GetMix() const { return m_mix; }

private:
    const MaterialBase* m_mat0;
    const MaterialBase* m_mat1;
    float m_mix;
};

// Core/Geometry/Vertex.h
struct Vertex
{
    Vertex() = default;
    Vertex(float x, float y, float z) : pos(x, y, z) {}
    Vertex(const Math::Vector3f& p, const Math::Vector3f& n, const Math::Vector3f& texcoord)
      : pos(p), normal(n), uv(texcoord)
    {
    }
    Math::Vector3f pos;
    Math::Vector3f normal = Math::Vector3f::UnitZ();
    Math::Vector3f uv;
};

enum class ShadingTypes
{
    Flat,
    Smooth
};

class GeometryBase
{
public:
    virtual ~GeometryBase() = default;
    void SetMaterial(const MaterialBase* material) { m_material = material; }
    const MaterialBase* GetMaterial() const { return m_material; }

private:
    const MaterialBase* m_material = nullptr;
};

// Core/Geometry/Triangle.h: three shared vertices and a shading type.
class Triangle : public GeometryBase
{
public:
    explicit Triangle(ShadingTypes shadingType = ShadingTypes::Flat) : m_shadingType(shadingType) {}
    Triangle(const Vertex* v0, const Vertex* v1, const Vertex* v2, ShadingTypes shadingType = ShadingTypes::Flat)
      : m_vertices{v0, v1, v2}, m_shadingType(shadingType)
    {
    }
    void SetVertices(const Vertex* v0, const Vertex* v1, const Vertex* v2) { m_vertices = {v0, v1, v2}; }
    const Vertex* GetVertex(int i) const { return m_vertices[i]; }
    ShadingTypes GetShadingType() const { return m_shadingType; }

private:
    std::array<const Vertex*, 3> m_vertices{};
    ShadingTypes m_shadingType;
};

// Core/Geometry/Sphere.h: centre and radius. The GPU path renders spheres as primitives
// after all the scene's triangles (Render numbers them so).
class Sphere : public GeometryBase
{
public:
    Sphere() = default;
    Sphere(const Math::Vector3f& o, float r) : m_origin(o), m_radius(r) {}
    const Math::Vector3f& GetOrigin() const { return m_origin; }
    float GetRadius() const { return m_radius; }

private:
    Math::Vector3f m_origin;
    float m_radius = 1.0f;
};

namespace detail
{

inline bool readFile(const std::filesystem::path& path, std::string& out)
{
    FILE* f = std::fopen(path.string().c_str(), "rb");
    if (!f) return false;
    out.clear();
    char buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, k);
    std::fclose(f);
    return true;
}

// A JSON value, as much of it as the reference's scene and setting files use (numbers
// are read as doubles, as nlohmann::json reads them).
struct Json
{
    enum Type
    {
        Null,
        Bool,
        Number,
        String,
        Array,
        Object
    } type = Null;
    bool boolean = false;
    double number = 0.0;
    std::string str;
    std::vector<Json> items;
    std::vector<std::pair<std::string, Json>> members;

    const Json* find(std::string_view key) const
    {
        if (type != Object) return nullptr;
        for (const auto& m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
};

inline bool parseJson(const char*& p, const char* e, Json& out, int depth = 0)
{
    auto ws = [&] {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    };
    auto str = [&](std::string& s) {
        if (p >= e || *p != '"') return false;
        ++p;
        s.clear();
        while (p < e && *p != '"') {
            char c = *p++;
            if (c == '\\') {
                if (p >= e) return false;
                c = *p++;
                switch (c) {
                case 'n': c = '\n'; break;
                case 't': c = '\t'; break;
                case 'r': c = '\r'; break;
                case 'b': c = '\b'; break;
                case 'f': c = '\f'; break;
                case 'u': {
                    if (e - p < 4) return false;
                    unsigned cp = 0;
                    for (int k = 0; k < 4; ++k) {
                        const char h = p[k];
                        const int dgt = h >= '0' && h <= '9'   ? h - '0'
                                        : h >= 'a' && h <= 'f' ? h - 'a' + 10
                                        : h >= 'A' && h <= 'F' ? h - 'A' + 10
                                                               : -1;
                        if (dgt < 0) return false;
                        cp = cp * 16 + (unsigned)dgt;
                    }
                    p += 4;
                    if (cp < 0x80) {
                        s += (char)cp;
                    } else if (cp < 0x800) {
                        s += (char)(0xc0 | (cp >> 6));
                        s += (char)(0x80 | (cp & 0x3f));
                    } else {
                        s += (char)(0xe0 | (cp >> 12));
                        s += (char)(0x80 | ((cp >> 6) & 0x3f));
                        s += (char)(0x80 | (cp & 0x3f));
                    }
                    continue;
                }
                default: break; // \" \\ \/
                }
            }
            s += c;
        }
        if (p >= e) return false;
        ++p;
        return true;
    };
    if (depth > 64) return false;
    ws();
    if (p >= e) return false;
    if (*p == '{') {
        out.type = Json::Object;
        ++p;
        ws();
        if (p < e && *p == '}') return ++p, true;
        for (;;) {
            std::string key;
            ws();
            if (!str(key)) return false;
            ws();
            if (p >= e || *p != ':') return false;
            ++p;
            Json v;
            if (!parseJson(p, e, v, depth + 1)) return false;
            out.members.emplace_back(std::move(key), std::move(v));
            ws();
            if (p < e && *p == ',') {
                ++p;
                continue;
            }
            if (p < e && *p == '}') return ++p, true;
            return false;
        }
    }
    if (*p == '[') {
        out.type = Json::Array;
        ++p;
        ws();
        if (p < e && *p == ']') return ++p, true;
        for (;;) {
            Json v;
            if (!parseJson(p, e, v, depth + 1)) return false;
            out.items.push_back(std::move(v));
            ws();
            if (p < e && *p == ',') {
                ++p;
                continue;
            }
            if (p < e && *p == ']') return ++p, true;
            return false;
        }
    }
    if (*p == '"') {
        out.type = Json::String;
        return str(out.str);
    }
    auto word = [&](const char* w) {
        const size_t n = std::strlen(w);
        if ((size_t)(e - p) >= n && std::strncmp(p, w, n) == 0) return p += n, true;
        return false;
    };
    if (word("true")) return out.type = Json::Bool, out.boolean = true, true;
    if (word("false")) return out.type = Json::Bool, true;
    if (word("null")) return true;
    const char* q = p;
    while (q < e && (std::strchr("+-.eE", *q) || (*q >= '0' && *q <= '9'))) ++q;
    if (q == p) return false;
    out.type = Json::Number;
    out.number = std::strtod(std::string(p, q).c_str(), nullptr);
    p = q;
    return true;
}

inline bool loadJson(const std::filesystem::path& path, Json& out)
{
    std::string text;
    if (!readFile(path, text)) return false;
    const char* p = text.data();
    return parseJson(p, p + text.size(), out);
}

// stbi_loadf of a Radiance file (Core/Texture2D.cpp:165-190): flat or run-length RGBE
// scanlines, byte * 2^(exponent - 136).
inline bool loadHdr(const std::filesystem::path& path, int& w, int& h, std::vector<float>& rgb)
{
    std::string d;
    if (!readFile(path, d)) return false;
    if (d.rfind("#?RADIANCE\n", 0) != 0 && d.rfind("#?RGBE\n", 0) != 0) return false;
    size_t pos = d.find('\n') + 1;
    bool fmt = false;
    for (;;) {
        const size_t end = d.find('\n', pos);
        if (end == std::string::npos) return false;
        const std::string line = d.substr(pos, end - pos);
        pos = end + 1;
        if (line.empty()) break;
        if (line == "FORMAT=32-bit_rle_rgbe") fmt = true;
    }
    const size_t end = d.find('\n', pos);
    if (!fmt || end == std::string::npos) return false;
    if (std::sscanf(d.substr(pos, end - pos).c_str(), "-Y %d +X %d", &h, &w) != 2 || w <= 0 || h <= 0) return false;
    pos = end + 1;
    const uint8_t* b = (const uint8_t*)d.data() + pos;
    const size_t avail = d.size() - pos;
    std::vector<uint8_t> e4((size_t)w * h * 4);
    const bool rle = w >= 8 && w < 32768 && avail >= 4 && b[0] == 2 && b[1] == 2 && !(b[2] & 0x80);
    if (!rle) {
        if (avail < e4.size()) return false;
        std::memcpy(e4.data(), b, e4.size());
    } else {
        size_t i = 0;
        for (int y = 0; y < h; ++y) {
            if (i + 4 > avail || b[i] != 2 || b[i + 1] != 2 || ((b[i + 2] << 8) | b[i + 3]) != w) return false;
            i += 4;
            for (int ch = 0; ch < 4; ++ch) {
                for (int x = 0; x < w;) {
                    if (i >= avail) return false;
                    int count = b[i++];
                    if (count > 128) {
                        count -= 128;
                        if (x + count > w || i >= avail) return false;
                        for (int k = 0; k < count; ++k) e4[4 * ((size_t)y * w + x + k) + ch] = b[i];
                        ++i;
                    } else {
                        if (count == 0 || x + count > w || i + count > avail) return false;
                        for (int k = 0; k < count; ++k) e4[4 * ((size_t)y * w + x + k) + ch] = b[i + k];
                        i += count;
                    }
                    x += count;
                }
            }
        }
    }
    rgb.resize((size_t)w * h * 3);
    for (size_t k = 0; k < (size_t)w * h; ++k) {
        const uint8_t* q = &e4[4 * k];
        const float f = q[3] ? (float)std::ldexp(1.0f, q[3] - 136) : 0.0f;
        for (int c = 0; c < 3; ++c) rgb[3 * k + c] = q[3] ? q[c] * f : 0.0f;
    }
    return true;
}

} // namespace detail

// Core/Geometry/Mesh.h, built in memory: vertices and triangles over them.
class Mesh
{
public:
    Mesh() = default;
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;
    // Reserves room so that vertex pointers stay valid while vertices are added.
    explicit Mesh(size_t maxVertices) { m_vertices.reserve(maxVertices); }
    const Vertex* AddVertex(const Vertex& v)
    {
        if (m_vertices.size() == m_vertices.capacity()) throw std::length_error("Mesh: vertex capacity reached");
        m_vertices.push_back(v);
        return &m_vertices.back();
    }
    void AddTriangle(const Triangle& t) { m_triangles.push_back(t); }
    const std::vector<Triangle>& GetTriangles() const { return m_triangles; }

    // Core/Geometry/Mesh.cpp:18-125: the triangles of an OBJ file as Assimp 4.1 hands them
    // over (pt_obj_parse), one vertex per face corner, all with the given material and
    // shading. The mesh must be empty. Returns false (error text in *error) when the file
    // cannot be read or parsed; the reference logs the error and leaves the mesh empty.
    bool Load(const std::filesystem::path& path, const MaterialBase* material,
              ShadingTypes shadingType = ShadingTypes::Flat, std::string* error = nullptr)
    {
        auto fail = [&](const std::string& msg) {
            if (error) *error = msg;
            return false;
        };
        if (!m_vertices.empty() || !m_triangles.empty()) return fail("Mesh::Load needs an empty mesh");
        std::string text;
        if (!detail::readFile(path, text)) return fail("cannot read " + path.string());
        int64_t n = 0;
        char err[256] = {0};
        if (pt_obj_parse(text.data(), (int64_t)text.size(), 0, nullptr, nullptr, nullptr, &n, err, sizeof err) != PT_OK)
            return fail(err);
        std::vector<float> p(9 * (size_t)n), nr(9 * (size_t)n), uv(6 * (size_t)n);
        if (n > 0 && pt_obj_parse(text.data(), (int64_t)text.size(), n, p.data(), nr.data(), uv.data(), &n, err,
                                  sizeof err) != PT_OK)
            return fail(err);
        m_vertices.reserve(3 * (size_t)n);
        m_triangles.reserve((size_t)n);
        for (size_t t = 0; t < (size_t)n; ++t) {
            const Vertex* v[3];
            for (int k = 0; k < 3; ++k) {
                const float* q = &p[9 * t + 3 * k];
                const float* m = &nr[9 * t + 3 * k];
                m_vertices.emplace_back(Math::Vector3f(q[0], q[1], q[2]), Math::Vector3f(m[0], m[1], m[2]),
                                        Math::Vector3f(uv[6 * t + 2 * k], uv[6 * t + 2 * k + 1], 0.0f));
                v[k] = &m_vertices.back();
            }
            Triangle tri(v[0], v[1], v[2], shadingType);
            tri.SetMaterial(material);
            m_triangles.push_back(tri);
        }
        return true;
    }

private:
    std::vector<Vertex> m_vertices;
    std::vector<Triangle> m_triangles;
};

// Core/Texture2D.h: row-major RGB float pixels.
class Texture2D
{
public:
    // Core/Texture2D.h:33-37: colour textures get gamma 2.2 when loaded, non-colour ones not
    enum class TextureColorType
    {
        Color,
        NonColor
    };

    Texture2D() = default;
    Texture2D(int width, int height) : m_width(width), m_height(height), m_pixels((size_t)width * height) {}
    void Clear(const Color3f& color = Color3f::Zero()) { std::fill(m_pixels.begin(), m_pixels.end(), color); }
    Color3f GetPixel(int i, int j) const { return m_pixels[(size_t)j * m_width + i]; }
    void SetPixel(int i, int j, const Color3f& color) { m_pixels[(size_t)j * m_width + i] = color; }
    int GetWidth() const { return m_width; }
    int GetHeight() const { return m_height; }
    bool IsValid() const { return m_width > 0 && m_height > 0; }
    const float* GetRawDataPtr() const { return &m_pixels[0].x; }
    float* GetRawDataPtr() { return &m_pixels[0].x; }

    // Core/Texture2D.cpp:107-195: a ".png" path is read as stbi_load(..., 3) reads it, the
    // format told by the file's header (pt_image_decode: PNG, BMP, TGA, binary PNM), bytes /
    // 255, then gamma 2.2 for a colour texture; a ".hdr" path as Radiance RGBE floats, no
    // gamma; any other extension, a missing file or an unreadable one gives false and
    // leaves the texture as it was.
    bool Load(const std::filesystem::path& path, TextureColorType textureColorType)
    {
        int32_t w = 0, h = 0;
        std::vector<float> rgb;
        if (path.extension() == ".png") {
            std::string data;
            if (!detail::readFile(path, data)) return false;
            const auto* bytes = (const uint8_t*)data.data();
            const int32_t color = textureColorType == TextureColorType::Color ? 1 : 0;
            if (pt_image_decode(bytes, (int64_t)data.size(), color, &w, &h, nullptr, nullptr, 0) != PT_OK) return false;
            rgb.resize(3 * (size_t)w * h);
            if (pt_image_decode(bytes, (int64_t)data.size(), color, &w, &h, rgb.data(), nullptr, 0) != PT_OK)
                return false;
        } else if (path.extension() == ".hdr") {
            int hw = 0, hh = 0;
            if (!detail::loadHdr(path, hw, hh, rgb)) return false;
            w = hw;
            h = hh;
        } else {
            return false;
        }
        m_width = w;
        m_height = h;
        m_pixels.assign((size_t)w * h, Color3f::Zero());
        std::memcpy(GetRawDataPtr(), rgb.data(), rgb.size() * sizeof(float));
        return true;
    }

    // Core/Texture2D.cpp:197-246: ".png" writes the reference's 8-bit bytes (ACES filmic,
    // RTCamp7 curve, degamma, truncation), converted on the GPU by pt_tonemap_rgb8;
    // ".hdr" writes the floats as Radiance RGBE. Returns false for another extension or a
    // failure (the reference logs the error and returns).
    bool Save(const std::filesystem::path& path, int device = 0) const
    {
        if (!IsValid()) return false;
        const size_t n = (size_t)m_width * m_height;
        if (path.extension() == ".png") {
            std::vector<uint8_t> px(3 * n);
            pt_context* ctx = nullptr;
            if (pt_ctx_create(device, &ctx) != PT_OK) return false;
            const int rc = pt_tonemap_rgb8(ctx, (int64_t)n, GetRawDataPtr(), px.data(), 0, nullptr);
            pt_ctx_destroy(ctx);
            return rc == PT_OK && writePng(path, px);
        }
        if (path.extension() == ".hdr") return writeHdr(path);
        return false;
    }

private:
    static uint32_t crc32(uint32_t crc, const uint8_t* p, size_t n)
    {
        // CRC-32 of the PNG chunks (ISO 3309, reflected polynomial 0xedb88320)
        static const std::array<uint32_t, 256> table = [] {
            std::array<uint32_t, 256> t{};
            for (uint32_t i = 0; i < 256; ++i) {
                uint32_t c = i;
                for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xedb88320u ^ (c >> 1) : c >> 1;
                t[i] = c;
            }
            return t;
        }();
        crc = ~crc;
        for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xff] ^ (crc >> 8);
        return ~crc;
    }

    static void put32(std::vector<uint8_t>& b, uint32_t v)
    {
        for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }

    // An 8-bit RGB PNG: unfiltered rows in one zlib stream of stored (uncompressed)
    // deflate blocks.
    bool writePng(const std::filesystem::path& path, const std::vector<uint8_t>& px) const
    {
        std::vector<uint8_t> raw;
        raw.reserve((size_t)m_height * (1 + 3 * (size_t)m_width));
        for (int y = 0; y < m_height; ++y) {
            raw.push_back(0);
            raw.insert(raw.end(), px.begin() + (size_t)y * 3 * m_width, px.begin() + (size_t)(y + 1) * 3 * m_width);
        }
        std::vector<uint8_t> z = {0x78, 0x01};
        uint32_t a = 1, b = 0; // Adler-32 of the uncompressed stream
        for (uint8_t c : raw) {
            a = (a + c) % 65521u;
            b = (b + a) % 65521u;
        }
        for (size_t pos = 0; pos < raw.size(); pos += 65535) {
            const size_t len = std::min<size_t>(65535, raw.size() - pos);
            z.push_back(pos + len >= raw.size() ? 1 : 0);
            z.push_back((uint8_t)(len & 0xff));
            z.push_back((uint8_t)(len >> 8));
            z.push_back((uint8_t)(~len & 0xff));
            z.push_back((uint8_t)((~len >> 8) & 0xff));
            z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        }
        put32(z, (b << 16) | a);
        std::vector<uint8_t> file = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
        auto chunk = [&](const char* kind, const std::vector<uint8_t>& data) {
            put32(file, (uint32_t)data.size());
            const size_t at = file.size();
            file.insert(file.end(), kind, kind + 4);
            file.insert(file.end(), data.begin(), data.end());
            put32(file, crc32(0, file.data() + at, 4 + data.size()));
        };
        std::vector<uint8_t> ihdr;
        put32(ihdr, (uint32_t)m_width);
        put32(ihdr, (uint32_t)m_height);
        ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0}); // 8-bit RGB, deflate, no filter, no interlace
        chunk("IHDR", ihdr);
        chunk("IDAT", z);
        chunk("IEND", {});
        FILE* f = std::fopen(path.string().c_str(), "wb");
        if (!f) return false;
        const bool ok = std::fwrite(file.data(), 1, file.size(), f) == file.size();
        return std::fclose(f) == 0 && ok;
    }

    // Radiance RGBE (Ward): each pixel's shared exponent from frexp of its largest component.
    bool writeHdr(const std::filesystem::path& path) const
    {
        FILE* f = std::fopen(path.string().c_str(), "wb");
        if (!f) return false;
        std::fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", m_height, m_width);
        std::vector<uint8_t> row(4 * (size_t)m_width);
        bool ok = true;
        for (int y = 0; y < m_height && ok; ++y) {
            for (int x = 0; x < m_width; ++x) {
                const Color3f& c = m_pixels[(size_t)y * m_width + x];
                const float m = std::max(c.x, std::max(c.y, c.z));
                uint8_t* e = &row[4 * (size_t)x];
                if (!(m >= 1e-32f)) {
                    e[0] = e[1] = e[2] = e[3] = 0;
                    continue;
                }
                int ex = 0;
                const float scale = (float)std::frexp(m, &ex) * 256.0f / m;
                e[0] = (uint8_t)(c.x * scale);
                e[1] = (uint8_t)(c.y * scale);
                e[2] = (uint8_t)(c.z * scale);
                e[3] = (uint8_t)(ex + 128);
            }
            ok = std::fwrite(row.data(), 1, row.size(), f) == row.size();
        }
        return std::fclose(f) == 0 && ok;
    }

    int m_width = 0, m_height = 0;
    std::vector<Color3f> m_pixels;
};

static_assert(sizeof(Color3f) == 3 * sizeof(float), "Texture2D pixels must be packed RGB floats");

// Core/Environment.h. The map is given as a texture (the reference loads it from a file).
class Environment
{
public:
    void SetBaseColor(const Color3f& baseColor) { m_baseColor = baseColor; }
    void SetZAxisRotation(float angle) { m_ZAxisRotation = angle; }
    void SetTexture(const Texture2D* texture) { m_texEnv = texture; }
    void EnableImportanceSampling(bool isEnabled) { m_useImportanceSampling = isEnabled; }
    const Texture2D* GetTexture() const { return m_texEnv; }
    const Color3f& GetBaseColor() const { return m_baseColor; }
    float GetZAxisRotation() const { return m_ZAxisRotation; }
    bool UseEnvImportanceSampling() const { return m_texEnv && m_texEnv->IsValid() && m_useImportanceSampling; }
    // Environment::Load (Core/Environment.cpp:15-27): Texture2D::Load of the map as a colour
    // texture (a Radiance .hdr, or a .png path), which the environment then owns (shared by
    // its copies). False for another format or a read error.
    bool Load(const std::filesystem::path& path)
    {
        auto tex = std::make_shared<Texture2D>();
        if (!tex->Load(path, Texture2D::TextureColorType::Color)) return false;
        m_ownedTex = tex;
        m_texEnv = tex.get();
        return true;
    }

private:
    std::shared_ptr<Texture2D> m_ownedTex;
    const Texture2D* m_texEnv = nullptr;
    Color3f m_baseColor;
    float m_ZAxisRotation = 0.0f;
    bool m_useImportanceSampling = true;
};

// Core/Camera.h and Camera.cpp:14-50: the same set-up operations.
class Camera
{
public:
    Camera() = default;
    Camera(const Math::Vector3f& pos, const Math::Vector3f& dir) : m_pos(pos) { SetViewDir(dir); }

    void SetPosition(const Math::Vector3f& pos) { m_pos = pos; }
    void SetFocalLength(float focalLength) { m_hPerf = m_sensorHeight / focalLength; }
    void FocusTo(const Math::Vector3f& target) { m_focusDist = (target - m_pos).Length(); }
    // 2 tan(fov / 2) with the double-precision tangent, rounded once
    void SetFOV(float fov) { m_hPerf = (float)(2.0f * std::tan((double)(0.5f * fov))); }
    void SetViewDir(const Math::Vector3f& dir)
    {
        m_forward = dir.Normalized();
        m_right = Math::Cross(m_forward, Math::Vector3f::UnitZ());
        m_up = Math::Cross(m_right, m_forward);
    }
    void LookAt(const Math::Vector3f& target) { SetViewDir(target - m_pos); }
    float GetFocalLength() const { return m_sensorHeight / m_hPerf; }
    float GetApeture() const { return m_apeture; }
    void SetApeture(float apeture) { m_apeture = apeture; }
    void SetFNumber(float fNumber) { m_apeture = GetFocalLength() / fNumber; }
    Math::Vector3f Right() const { return m_right; }
    Math::Vector3f Up() const { return m_up; }
    Math::Vector3f Forward() const { return m_forward; }

    // The 16 floats of pt_scene_desc.camera.
    void CopyState(float* out) const
    {
        const Math::Vector3f v[4] = {m_pos, m_right, m_up, m_forward};
        for (int i = 0; i < 4; ++i) {
            out[3 * i] = v[i].x;
            out[3 * i + 1] = v[i].y;
            out[3 * i + 2] = v[i].z;
        }
        out[12] = m_focusDist;
        out[13] = m_hPerf;
        out[14] = m_apeture;
        out[15] = 0.0f;
    }

private:
    Math::Vector3f m_pos, m_right, m_up, m_forward;
    float m_focusDist = 1.0f;
    float m_sensorHeight = 24e-3f;
    float m_hPerf = 24e-3f / 55e-3f;
    float m_apeture = 25e-3f;
};

// Core/RenderSetting.h
struct RenderSetting
{
    int outputWidth = 1280;
    int outputHeight = 720;
    int numSamplesPerPixel = 128;
    int numSppForDenoising = 128;
    int numMaxBounces = 16;
    int tileWidth = 16;
    int tileHeight = 16;
    int numThreads = 0;
};

// Core/Scene.h: non-owning geometry and light lists, owned materials, camera and
// environment, non-owning target textures.
class Scene
{
public:
    struct AOVType
    {
        enum Value
        {
            Rendered,
            UV,
            DenoisingAlbedo,
            DenoisingNormal,
            NumAOVTypes
        };
    };

    void AppendGeometry(const GeometryBase* geometry) { m_geometries.emplace_back(geometry); }
    void AppendMesh(const Mesh& mesh)
    {
        for (const auto& t : mesh.GetTriangles()) AppendGeometry(&t);
    }
    void AppendLightMesh(const Mesh& mesh)
    {
        for (const auto& t : mesh.GetTriangles()) {
            AppendGeometry(&t);
            AppendLight(&t);
        }
    }
    void AppendLight(const GeometryBase* geometry) { m_lights.emplace_back(geometry); }
    const std::vector<const GeometryBase*>& GetGeometries() const { return m_geometries; }
    const std::vector<const GeometryBase*>& GetLights() const { return m_lights; }
    const Camera* GetMainCamera() const { return m_mainCamera.get(); }
    void RegisterMaterial(std::string_view materialName, std::unique_ptr<const MaterialBase> material)
    {
        m_materials[std::string(materialName)] = std::move(material);
    }
    const MaterialBase* GetMaterial(std::string_view materialName)
    {
        return m_materials[std::string(materialName)].get();
    }
    void SetMainCamera(std::unique_ptr<Camera> camera) { m_mainCamera = std::move(camera); }
    const Environment& GetEnvironment() const { return m_environment; }
    void SetEnvironment(const Environment& environment) { m_environment = environment; }
    void SetTargetTexture(AOVType::Value aovType, Texture2D* targetTex) { m_targetTextures[aovType] = targetTex; }
    Texture2D* GetTargetTexture(AOVType::Value aovType) const { return m_targetTextures[aovType]; }
    const RenderSetting& GetRenderSetting() const { return m_renderSetting; }
    void SetRenderSetting(const RenderSetting& renderSetting) { m_renderSetting = renderSetting; }

    // Scene::LoadRenderSetting (Core/Scene.cpp:12-26, RenderSettingLoader.cpp:15-64): the
    // keys present in the JSON file override the current values (numbers truncated to int).
    bool LoadRenderSetting(const std::filesystem::path& path)
    {
        detail::Json j;
        if (!detail::loadJson(path, j)) return false;
        auto read = [&](int* v, const char* key) {
            if (const detail::Json* x = j.find(key); x && x->type == detail::Json::Number) *v = (int)x->number;
        };
        read(&m_renderSetting.outputWidth, "outputWidth");
        read(&m_renderSetting.outputHeight, "outputHeight");
        read(&m_renderSetting.numSamplesPerPixel, "spp");
        read(&m_renderSetting.numSppForDenoising, "denoising_spp");
        read(&m_renderSetting.numMaxBounces, "maxBounces");
        read(&m_renderSetting.tileWidth, "tileWidth");
        read(&m_renderSetting.tileHeight, "tileHeight");
        read(&m_renderSetting.numThreads, "numThreads");
        return true;
    }

    // Scene::LoadModel (Core/Scene.cpp:28-51): an OBJ file's triangles, smooth-shaded, with
    // the named material ("default" without a name). The scene owns the mesh. False (with
    // *error) for an unregistered material or a file that cannot be read.
    bool LoadModel(const std::filesystem::path& path, std::string_view materialName = "default",
                   std::string* error = nullptr)
    {
        const MaterialBase* material = GetMaterial(materialName);
        if (!material) {
            if (error) *error = "material '" + std::string(materialName) + "' is not registered";
            return false;
        }
        auto mesh = std::make_unique<Mesh>();
        if (!mesh->Load(path, material, ShadingTypes::Smooth, error)) return false;
        AppendMesh(*mesh);
        m_loadedMeshes.push_back(std::move(mesh));
        return true;
    }

    // Scene::LoadAssets = SceneLoaderJson::Load (Core/SceneLoader.cpp:17-296): materials,
    // camera, environment and OBJ assets of a JSON scene file, paths relative to it. As in
    // the reference: "default" is a white Lambert, "emission" materials register nothing,
    // and the camera takes SetFNumber before SetFocalLength. False (with *error) where the
    // reference would use a missing material, for material textures other than .png and for
    // environment maps other than Radiance .hdr and .png. Material textures are read as the
    // reference reads them (Core/SceneLoader.cpp:106-189): a Lambert's and a GGX's
    // "color_tex" from the scene's folder as colour textures, a GGX's "roughness_tex" from
    // the path as written (relative to the working directory) as a non-colour texture; a
    // file that does not exist leaves the material untextured; a texture registered under
    // a name already taken reuses the first one (Core/Scene.h:131-143).
    bool LoadAssets(const std::filesystem::path& path, std::string* error = nullptr)
    {
        auto fail = [&](const std::string& msg) {
            if (error) *error = msg;
            return false;
        };
        using detail::Json;
        Json j;
        if (!detail::loadJson(path, j)) return fail("cannot read or parse " + path.string());
        const std::filesystem::path base = path.parent_path();
        auto vec3 = [](const Json* o, const char* key) {
            float v[3] = {0.0f, 0.0f, 0.0f};
            if (const Json* a = o ? o->find(key) : nullptr; a && a->type == Json::Array)
                for (size_t i = 0; i < a->items.size() && i < 3; ++i) v[i] = (float)a->items[i].number;
            return Math::Vector3f(v[0], v[1], v[2]);
        };
        auto num = [](const Json* o, const char* key, float def) {
            const Json* x = o ? o->find(key) : nullptr;
            return x && x->type == Json::Number ? (float)x->number : def;
        };
        auto text = [](const Json* o, const char* key) {
            const Json* x = o ? o->find(key) : nullptr;
            return x && x->type == Json::String ? x->str : std::string();
        };
        RegisterMaterial("default", std::make_unique<Lambert>(Color3f::One()));
        if (const Json* mats = j.find("materials"); mats && mats->type == Json::Array) {
            for (const Json& m : mats->items) {
                const std::string type = text(&m, "type"), name = text(&m, "name");
                if (type.empty() || name.empty()) continue;
                // Texture2D::Load then RegisterTexture; nullptr where the file does not exist
                std::string texError;
                auto texture = [&](const std::filesystem::path& file, const std::string& key,
                                   bool color) -> const Texture2D* {
                    if (!std::filesystem::exists(file)) return nullptr;
                    Texture2D t;
                    if (!t.Load(file, color ? Texture2D::TextureColorType::Color
                                            : Texture2D::TextureColorType::NonColor)) {
                        texError = "material '" + name + "': cannot read texture " + file.string() +
                                   " (Texture2D::Load reads .png and .hdr paths)";
                        return nullptr;
                    }
                    return &m_textures.emplace(key, std::move(t)).first->second;
                };
                const std::string colorTex = text(&m, "color_tex"), roughTex = text(&m, "roughness_tex");
                if (type == "lambert") {
                    auto lm = std::make_unique<Lambert>(vec3(&m, "base_color"));
                    if (!colorTex.empty()) lm->SetTexAlbedo(texture(base / colorTex, (base / colorTex).string(), true));
                    if (!texError.empty()) return fail(texError);
                    RegisterMaterial(name, std::move(lm));
                } else if (type == "glass") {
                    RegisterMaterial(name, std::make_unique<Glass>(vec3(&m, "color"), num(&m, "ior", 1.0f)));
                } else if (type == "glossy") {
                    auto gm = std::make_unique<GGX>(vec3(&m, "color"), num(&m, "roughness", 0.0f));
                    if (!colorTex.empty()) gm->SetF0Texture(texture(base / colorTex, colorTex, true));
                    if (!roughTex.empty()) gm->SetRoughnessMap(texture(roughTex, roughTex, false));
                    if (!texError.empty()) return fail(texError);
                    RegisterMaterial(name, std::move(gm));
                } else if (type == "mix") {
                    const MaterialBase* m0 = GetMaterial(text(&m, "mat0_name"));
                    const MaterialBase* m1 = GetMaterial(text(&m, "mat1_name"));
                    if (!m0 || !m1) return fail("material '" + name + "' mixes unregistered materials");
                    RegisterMaterial(name, std::make_unique<MixMaterial>(m0, m1, num(&m, "mix", 0.5f)));
                }
            }
        }
        const Json* cd = j.find("camera");
        auto camera = std::make_unique<Camera>();
        camera->SetPosition(vec3(cd, "position"));
        camera->LookAt(vec3(cd, "look_at"));
        camera->FocusTo(vec3(cd, "focus_pos"));
        camera->SetFNumber(num(cd, "f_number", 2.8f));
        camera->SetFocalLength(num(cd, "focal_length", 55e-3f));
        SetMainCamera(std::move(camera));
        const Json* ed = j.find("env");
        Environment env;
        env.SetBaseColor(vec3(ed, "base_color"));
        if (const std::string tex = text(ed, "texture"); !tex.empty() && !env.Load(base / tex))
            return fail("environment map " + (base / tex).string() + ": not a readable .hdr or .png file");
        SetEnvironment(env);
        if (const Json* assets = j.find("assets"); assets && assets->type == Json::Array) {
            for (const Json& a : assets->items) {
                const std::string mat = a.find("mat") ? text(&a, "mat") : std::string("default");
                if (!LoadModel(base / text(&a, "path"), mat, error)) return false;
            }
        }
        return true;
    }

private:
    std::vector<std::unique_ptr<Mesh>> m_loadedMeshes;
    std::vector<const GeometryBase*> m_geometries;
    std::vector<const GeometryBase*> m_lights;
    std::unordered_map<std::string, std::unique_ptr<const MaterialBase>> m_materials;
    std::map<std::string, Texture2D> m_textures; // loaded material textures by name (nodes stay put)
    Environment m_environment;

    std::unique_ptr<Camera> m_mainCamera;
    std::array<Texture2D*, AOVType::NumAOVTypes> m_targetTextures = {};
    RenderSetting m_renderSetting;
};

// Core/Ray.h:9-49: the ray of the integrators (origin, unit direction and the path's
// throughput, bounce count, survival probability and index of refraction).
enum class RayTypes : uint8_t
{
    Camera,
    Shadow,
    Diffuse,
    Translucent,
    Glossy,
    Refract,
};
struct Ray
{
    Ray() = default;
    Ray(const Math::Vector3f& o, const Math::Vector3f& dir, RayTypes rayType = RayTypes::Camera,
        Color3f throughput = Color3f::One(), int bounce = 0, float prob = 1.0f, float ior = 1.0f)
      : o(o), dir(dir), throughput(throughput), bounce(bounce), prob(prob), ior(ior), rayType(rayType)
    {
    }
    Math::Vector3f o;
    Math::Vector3f dir;
    Color3f throughput = Color3f::One();
    int bounce = 0;
    float prob = 1.0f;
    float ior = 1.0f;
    RayTypes rayType = RayTypes::Camera;
};

// Core/HitInfo.h:15-19: the nearest hit's distance and geometry.
struct HitInfo
{
    float distance = std::numeric_limits<float>::infinity();
    const GeometryBase* hitObj = nullptr;
};

// Core/Accel/AccelBase.h:17-56: Build over a scene's geometries, then closest-hit queries
// within [distMin, distMax] (defaults [0, FLT_MAX], or [distMin, inf]).
enum class AccelType
{
    BruteForce,
    BVH
};
class AccelBase
{
public:
    virtual ~AccelBase() = default;
    virtual void Build(const Scene& scene) = 0;
    virtual std::optional<HitInfo> Intersect(const Ray& ray, const Scene& scene, float distMin,
                                             float distMax) const = 0;
    std::optional<HitInfo> Intersect(const Ray& ray, const Scene& scene) const
    {
        return Intersect(ray, scene, 0.0f, std::numeric_limits<float>::max());
    }
    std::optional<HitInfo> Intersect(const Ray& ray, const Scene& scene, float distMin) const
    {
        return Intersect(ray, scene, distMin, std::numeric_limits<float>::infinity());
    }
};

// Core/Accel/BinnedSAHBVH.h: the binned-SAH tree (16 bins, leaves of up to 4) over the
// scene's triangles and spheres, built and traversed on the GPU (pt_bvh_build,
// pt_trace_closest_host). Build keeps the primitives in the geometries' order (triangles
// first, then spheres, as Render numbers them); a hit reports its geometry. Intersect
// traces one ray per call; IntersectBatch traces many in one launch. Build returns with
// an empty tree (every query a miss) and GetLastError set when the device fails.
class BinnedSAHBVH : public AccelBase
{
public:
    explicit BinnedSAHBVH(int device = 0) : m_device(device) {}
    ~BinnedSAHBVH() override
    {
        release();
        if (m_ctx) pt_ctx_destroy(m_ctx);
    }
    BinnedSAHBVH(const BinnedSAHBVH&) = delete;
    BinnedSAHBVH& operator=(const BinnedSAHBVH&) = delete;

    // Extension: true builds on the host with the reference's sort at every node, so the
    // primitives of a leaf are also in the reference's order among equal centroid keys;
    // false (default) builds on the device (same nodes, ties by primitive index).
    void SetReferenceLeafOrder(bool on) { m_referenceOrder = on; }

    void Build(const Scene& scene) override
    {
        release();
        m_error.clear();
        std::vector<float> pos, sph;
        std::vector<int32_t> tmat, smat;
        std::vector<const GeometryBase*> spheres;
        for (const GeometryBase* g : scene.GetGeometries()) {
            if (dynamic_cast<const Sphere*>(g)) {
                spheres.push_back(g);
            } else if (const auto* t = dynamic_cast<const Triangle*>(g)) {
                for (int k = 0; k < 3; ++k) {
                    const Math::Vector3f& p = t->GetVertex(k)->pos;
                    pos.insert(pos.end(), {p.x, p.y, p.z});
                }
                m_prims.push_back(g);
                tmat.push_back(0);
            }
        }
        for (const GeometryBase* g : spheres) {
            const auto* sp = static_cast<const Sphere*>(g);
            sph.insert(sph.end(), {sp->GetOrigin().x, sp->GetOrigin().y, sp->GetOrigin().z, sp->GetRadius()});
            m_prims.push_back(g);
            smat.push_back(0);
        }
        // geometry only: one placeholder material, no lights, no camera
        const int32_t matType = PT_MAT_LAMBERT;
        const float matParam[8] = {1.0f, 1.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f};
        pt_scene_desc d{};
        d.num_triangles = (int32_t)tmat.size();
        d.positions = pos.empty() ? nullptr : pos.data();
        d.material_ids = tmat.empty() ? nullptr : tmat.data();
        d.num_materials = 1;
        d.material_types = &matType;
        d.material_params = matParam;
        d.num_spheres = (int32_t)smat.size();
        d.spheres = sph.empty() ? nullptr : sph.data();
        d.sphere_material_ids = smat.empty() ? nullptr : smat.data();
        if (!m_ctx && pt_ctx_create(m_device, &m_ctx) != PT_OK) {
            m_error = "device context creation failed";
            m_prims.clear();
            return;
        }
        // BinnedSAHBVH::kNumBins, leaves of up to 4
        pt_bvh_desc bd{16, 4, 2, m_referenceOrder ? PT_BUILD_HOST_REFERENCE_ORDER : PT_BUILD_DEVICE};
        if (pt_scene_create(m_ctx, &d, &m_scene) != PT_OK || pt_bvh_build(m_scene, &bd, &m_bvh) != PT_OK) {
            m_error = pt_last_error(m_ctx);
            release();
        }
    }

    std::optional<HitInfo> Intersect(const Ray& ray, const Scene& scene, float distMin,
                                     float distMax) const override
    {
        std::vector<std::optional<HitInfo>> h = IntersectBatch(std::vector<Ray>{ray}, distMin, distMax);
        (void)scene;
        return h[0];
    }
    using AccelBase::Intersect;

    // Closest hits of many rays in one launch (the same query per ray).
    std::vector<std::optional<HitInfo>> IntersectBatch(const std::vector<Ray>& rays, float distMin,
                                                       float distMax) const
    {
        std::vector<std::optional<HitInfo>> out(rays.size());
        if (!m_bvh || rays.empty()) return out;
        std::vector<float> r(8 * rays.size()), t(rays.size());
        std::vector<int32_t> prim(rays.size());
        for (size_t i = 0; i < rays.size(); ++i) {
            const Ray& q = rays[i];
            const float v[8] = {q.o.x, q.o.y, q.o.z, distMin, q.dir.x, q.dir.y, q.dir.z, distMax};
            std::memcpy(&r[8 * i], v, sizeof v);
        }
        if (pt_trace_closest_host(m_bvh, (int64_t)rays.size(), r.data(), t.data(), prim.data()) != PT_OK) {
            m_error = pt_last_error(m_ctx);
            return out;
        }
        for (size_t i = 0; i < rays.size(); ++i)
            if (prim[i] >= 0) out[i] = HitInfo{t[i], m_prims[(size_t)prim[i]]};
        return out;
    }

    const std::string& GetLastError() const { return m_error; }

private:
    void release()
    {
        if (m_bvh) pt_bvh_destroy(m_bvh);
        if (m_scene) pt_scene_destroy(m_scene);
        m_bvh = nullptr;
        m_scene = nullptr;
        m_prims.clear();
    }

    int m_device;
    bool m_referenceOrder = false;
    pt_context* m_ctx = nullptr;
    pt_scene* m_scene = nullptr;
    pt_bvh* m_bvh = nullptr;
    std::vector<const GeometryBase*> m_prims; // primitive index -> geometry
    mutable std::string m_error;
};

enum class Integrator
{
    PathTracing,      // MIS (Core/Integrator/PathTracing.cpp)
    SimplePathTracing // the integrator Petrichor::Render uses (Petrichor.cpp:58)
};

// Core/Petrichor.h: the render entry. Render blocks until the image is in the target
// texture; the callback then receives the total time.
class Petrichor
{
public:
    explicit Petrichor(int device = 0) : m_device(device) {}
    ~Petrichor()
    {
        if (m_ctx) pt_ctx_destroy(m_ctx);
    }
    Petrichor(const Petrichor&) = delete;
    Petrichor& operator=(const Petrichor&) = delete;

    void SetRenderCallback(const std::function<void(const RenderingResult&)>& onRenderingFinished)
    {
        m_onRenderingFinished = onRenderingFinished;
    }
    uint32_t GetNumRenderedTiles() const { return m_numRenderedTiles; }
    uint32_t GetNumTiles() const { return m_numTiles; }

    // Extensions of the GPU path: integrator choice, the sampler seed, and the 2D sampler
    // (0: RandomSampler2D as Core::Petrichor::Render uses; > 0: MicroJitteredSampler
    // initialised with that sequence length, sample s of a pixel starting at Halton point
    // s; the denoising AOVs keep the random sampler).
    void SetIntegrator(Integrator integrator) { m_integrator = integrator; }
    void SetSeed(uint32_t seed) { m_seed = seed; }
    void SetMicroJitteredSampler(int lengthSeq) { m_samplerLength = lengthSeq; }
    // true: the BVH is built on the host with the reference's sort at every node, so the
    // leaf order is the reference's also among equal centroid keys (see BinnedSAHBVH).
    void SetReferenceLeafOrder(bool on) { m_referenceOrder = on; }
    // Error text of the last failed Render (empty after a success).
    const std::string& GetLastError() const { return m_error; }

    // Returns false (with GetLastError) when the scene has no camera or target texture,
    // holds a material or primitive the GPU path does not render, or the device fails.
    bool Render(const Scene& scene)
    {
        const auto t0 = std::chrono::steady_clock::now();
        m_error.clear();
        Texture2D* target = scene.GetTargetTexture(Scene::AOVType::Rendered);
        if (!target || !target->IsValid()) return fail("Target texture has not been set.");
        if (!scene.GetMainCamera()) return fail("The scene has no camera.");
        const RenderSetting& rs = scene.GetRenderSetting();
        // TileManager.cpp:8-43 counts (W / tw + 1) * (H / th + 1) tiles
        m_numTiles = (uint32_t)((target->GetWidth() / rs.tileWidth + 1) * (target->GetHeight() / rs.tileHeight + 1));
        m_numRenderedTiles = 0;

        Flat f;
        if (!flatten(scene, f)) return false;
        if (!m_ctx && pt_ctx_create(m_device, &m_ctx) != PT_OK) return fail("device context creation failed");
        pt_scene* s = nullptr;
        pt_bvh* bvh = nullptr;
        if (pt_scene_create(m_ctx, &f.desc, &s) != PT_OK) return fail(pt_last_error(m_ctx));
        // BinnedSAHBVH: 16 bins, leaves of up to 4
        pt_bvh_desc bd{16, 4, 2, m_referenceOrder ? PT_BUILD_HOST_REFERENCE_ORDER : PT_BUILD_DEVICE};
        if (pt_bvh_build(s, &bd, &bvh) != PT_OK) {
            pt_scene_destroy(s);
            return fail(pt_last_error(m_ctx));
        }
        pt_render_desc rd{target->GetWidth(), target->GetHeight(), rs.numSamplesPerPixel, rs.numMaxBounces,
                          m_integrator == Integrator::PathTracing ? PT_INTEGRATOR_MIS : PT_INTEGRATOR_SIMPLE, m_seed,
                          0, m_samplerLength};
        int rc = pt_render(s, bvh, &rd, target->GetRawDataPtr(), 0, nullptr);
        // the AOV passes (Core/Petrichor.cpp:104-261) into their target textures: UV at
        // RenderSetting::numSamplesPerPixel samples, the denoising ones at numSppForDenoising
        const std::pair<Scene::AOVType::Value, int32_t> aovs[3] = {
            {Scene::AOVType::UV, PT_AOV_UV},
            {Scene::AOVType::DenoisingAlbedo, PT_AOV_DENOISING_ALBEDO},
            {Scene::AOVType::DenoisingNormal, PT_AOV_DENOISING_NORMAL}};
        for (const auto& [type, aov] : aovs) {
            Texture2D* tex = scene.GetTargetTexture(type);
            if (rc != PT_OK || !tex || !tex->IsValid()) continue;
            pt_render_desc ad{tex->GetWidth(), tex->GetHeight(),
                              aov == PT_AOV_UV ? rs.numSamplesPerPixel : rs.numSppForDenoising, rs.numMaxBounces,
                              PT_INTEGRATOR_MIS, m_seed, 0, 0};
            rc = pt_render_aov(s, bvh, aov, &ad, tex->GetRawDataPtr(), 0, nullptr);
        }
        if (rc != PT_OK) fail(pt_last_error(m_ctx));
        pt_bvh_destroy(bvh);
        pt_scene_destroy(s);
        if (rc != PT_OK) return false;
        m_numRenderedTiles = m_numTiles;
        if (m_onRenderingFinished) {
            const float sec = std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count();
            m_onRenderingFinished(RenderingResult{sec});
        }
        return true;
    }

private:
    struct Flat {
        std::vector<float> pos, nrm, uv, matParam, env, spheres;
        std::vector<int32_t> sphereMat;
        std::vector<int32_t> matId, smooth, matType, lights, matMaps, matNormalMaps, texSizes;
        std::vector<float> matStrengths;
        bool anyNormalMap = false, anyStrength = false;
        std::vector<const float*> texPixels;
        std::unordered_map<const Texture2D*, int32_t> texIndex;
        bool anyMap = false;
        pt_scene_desc desc{};
    };

    bool fail(const std::string& msg)
    {
        m_error = msg;
        return false;
    }

    static int32_t type_code(MaterialTypes t) { return (int32_t)t; }

    // Material table entry (8 floats: colour, p3, p4, p5, importance sampling, 0).
    int32_t material_index(const MaterialBase* m, std::unordered_map<const MaterialBase*, int32_t>& index, Flat& f)
    {
        auto it = index.find(m);
        if (it != index.end()) return it->second;
        float p[8] = {0, 0, 0, 0, 0, 0, m->IsImportanceSamplingEnabled() ? 1.0f : 0.0f, 0};
        int32_t maps[2] = {-1, -1}, normalMap = -1;
        float strengths[2] = {1.0f, 1.0f};
        auto rgb = [&](const Color3f& c) {
            p[0] = c.x;
            p[1] = c.y;
            p[2] = c.z;
        };
        switch (m->GetMaterialType()) {
        case MaterialTypes::Lambert:
            rgb(static_cast<const Lambert*>(m)->GetKd());
            maps[0] = texture_index(static_cast<const Lambert*>(m)->GetTexAlbedo(), f);
            if (maps[0] == -3) return -3;
            break;
        case MaterialTypes::Glossy: {
            const auto* g = static_cast<const GGX*>(m);
            rgb(g->GetReflectance());
            p[3] = g->GetRoughness();
            maps[0] = texture_index(g->GetF0Texture(), f);
            maps[1] = texture_index(g->GetRoughnessMap(), f);
            normalMap = texture_index(g->GetNormalMap(), f);
            if (maps[0] == -3 || maps[1] == -3 || normalMap == -3) return -3;
            strengths[0] = g->GetNormalMapStrength();
            strengths[1] = g->GetRoughnessMapStrength();
            f.anyNormalMap |= normalMap >= 0;
            f.anyStrength |= strengths[0] != 1.0f || strengths[1] != 1.0f;
            break;
        }
        case MaterialTypes::Glass:
            rgb(static_cast<const Glass*>(m)->GetColor());
            p[3] = static_cast<const Glass*>(m)->GetIor();
            break;
        case MaterialTypes::Emission: rgb(static_cast<const Emission*>(m)->GetLightColor()); break;
        case MaterialTypes::Mix: {
            const auto* mix = static_cast<const MixMaterial*>(m);
            // sub-materials first, so that their indices are known
            const int32_t a = material_index(mix->GetMaterial0(), index, f);
            const int32_t b = material_index(mix->GetMaterial1(), index, f);
            if (a < 0 || b < 0) return a < 0 ? a : b;
            p[3] = mix->GetMix();
            p[4] = (float)a;
            p[5] = (float)b;
            break;
        }
        }
        const int32_t id = (int32_t)f.matType.size();
        f.matType.push_back(type_code(m->GetMaterialType()));
        f.matParam.insert(f.matParam.end(), p, p + 8);
        f.matMaps.insert(f.matMaps.end(), maps, maps + 2);
        f.matNormalMaps.push_back(normalMap);
        f.matStrengths.insert(f.matStrengths.end(), strengths, strengths + 2);
        index[m] = id;
        return id;
    }

    // Index of a material texture in the descriptor's list (-1: none; -3: empty texture).
    static int32_t texture_index(const Texture2D* t, Flat& f)
    {
        if (!t) return -1;
        if (!t->IsValid()) return -3;
        auto it = f.texIndex.find(t);
        if (it != f.texIndex.end()) return it->second;
        const int32_t id = (int32_t)f.texPixels.size();
        f.texPixels.push_back(t->GetRawDataPtr());
        f.texSizes.push_back(t->GetWidth());
        f.texSizes.push_back(t->GetHeight());
        f.texIndex[t] = id;
        f.anyMap = true;
        return id;
    }

    bool flatten(const Scene& scene, Flat& f)
    {
        std::unordered_map<const MaterialBase*, int32_t> matIndex;
        std::unordered_map<const GeometryBase*, int32_t> triIndex;
        bool anySmooth = false, anyUV = false;
        for (const GeometryBase* g : scene.GetGeometries())
            if (const auto* t = dynamic_cast<const Triangle*>(g)) {
                anySmooth |= t->GetShadingType() == ShadingTypes::Smooth;
                for (int k = 0; k < 3; ++k)
                    if (t->GetVertex(k) && (t->GetVertex(k)->uv.x != 0.0f || t->GetVertex(k)->uv.y != 0.0f)) anyUV = true;
            }
        std::vector<const Sphere*> spheres;
        for (const GeometryBase* g : scene.GetGeometries()) {
            if (const auto* sp = dynamic_cast<const Sphere*>(g)) { // numbered after the triangles
                if (!sp->GetMaterial()) return fail("a sphere has no material");
                spheres.push_back(sp);
                continue;
            }
            const auto* t = dynamic_cast<const Triangle*>(g);
            if (!t) return fail("the GPU path renders triangles and spheres only");
            if (!t->GetMaterial()) return fail("a triangle has no material");
            triIndex[g] = (int32_t)f.matId.size();
            for (int k = 0; k < 3; ++k) {
                const Vertex* v = t->GetVertex(k);
                f.pos.insert(f.pos.end(), {v->pos.x, v->pos.y, v->pos.z});
                if (anySmooth) f.nrm.insert(f.nrm.end(), {v->normal.x, v->normal.y, v->normal.z});
                if (anyUV) f.uv.insert(f.uv.end(), {v->uv.x, v->uv.y});
            }
            if (anySmooth) f.smooth.push_back(t->GetShadingType() == ShadingTypes::Smooth ? 1 : 0);
            const int32_t mi = material_index(t->GetMaterial(), matIndex, f);
            if (mi == -3) return fail("a material texture is empty");
            f.matId.push_back(mi);
        }
        const int32_t numTris = (int32_t)f.matId.size();
        for (const Sphere* sp : spheres) {
            triIndex[sp] = numTris + (int32_t)f.sphereMat.size();
            const int32_t mi = material_index(sp->GetMaterial(), matIndex, f);
            if (mi == -3) return fail("a material texture is empty");
            f.spheres.insert(f.spheres.end(), {sp->GetOrigin().x, sp->GetOrigin().y, sp->GetOrigin().z, sp->GetRadius()});
            f.sphereMat.push_back(mi);
        }
        for (const GeometryBase* l : scene.GetLights()) {
            auto it = triIndex.find(l);
            if (it == triIndex.end()) return fail("a light is not one of the scene's triangles or spheres");
            f.lights.push_back(it->second);
        }
        if (f.matType.empty()) { // pt_scene_create wants one material even for an empty scene
            f.matType.push_back(type_code(MaterialTypes::Lambert));
            f.matParam.assign(8, 0.0f);
            f.matMaps.assign(2, -1);
            f.matNormalMaps.assign(1, -1);
            f.matStrengths.assign(2, 1.0f);
        }
        pt_scene_desc& d = f.desc;
        d.num_triangles = (int32_t)f.matId.size();
        d.num_spheres = (int32_t)f.sphereMat.size();
        d.spheres = f.spheres.empty() ? nullptr : f.spheres.data();
        d.sphere_material_ids = f.sphereMat.empty() ? nullptr : f.sphereMat.data();
        d.positions = f.pos.data();
        d.normals = anySmooth ? f.nrm.data() : nullptr;
        d.material_ids = f.matId.data();
        d.smooth = anySmooth ? f.smooth.data() : nullptr;
        d.num_materials = (int32_t)f.matType.size();
        d.material_types = f.matType.data();
        d.material_params = f.matParam.data();
        d.num_lights = (int32_t)f.lights.size();
        d.light_triangles = f.lights.data();
        scene.GetMainCamera()->CopyState(d.camera);
        const Environment& env = scene.GetEnvironment();
        const Texture2D* tex = env.GetTexture();
        if (tex && tex->IsValid()) {
            d.env_width = tex->GetWidth();
            d.env_height = tex->GetHeight();
            d.env_rgb = tex->GetRawDataPtr();
        }
        d.env_base_color[0] = env.GetBaseColor().x;
        d.env_base_color[1] = env.GetBaseColor().y;
        d.env_base_color[2] = env.GetBaseColor().z;
        d.env_z_rotation = env.GetZAxisRotation();
        d.env_importance_sampling = env.UseEnvImportanceSampling() ? 1 : 0;
        d.uvs = anyUV ? f.uv.data() : nullptr;
        if (f.anyMap) {
            d.num_textures = (int32_t)f.texPixels.size();
            d.texture_sizes = f.texSizes.data();
            d.textures = f.texPixels.data();
            d.material_maps = f.matMaps.data();
            if (f.anyNormalMap) d.material_normal_maps = f.matNormalMaps.data();
        }
        if (f.anyStrength) d.material_map_strengths = f.matStrengths.data();
        return true;
    }

    int m_device;
    pt_context* m_ctx = nullptr;
    Integrator m_integrator = Integrator::SimplePathTracing;
    uint32_t m_seed = 1;
    int m_samplerLength = 0;
    bool m_referenceOrder = false;
    std::string m_error;
    std::atomic<uint32_t> m_numRenderedTiles = 0;
    uint32_t m_numTiles = 0;
    std::function<void(const RenderingResult&)> m_onRenderingFinished;
};

} // namespace Core
} // namespace Petrichor
