// gltf.cpp -- glTF 2.0 scene ingest, the part of SceneParser::parse_scene_file
// (src/Scene/SceneParser.cpp:19-226) that the reference hands to ASSIMP with
// aiProcess_PreTransformVertices | aiProcess_Triangulate, restated for glTF
// files: .gltf with external or data: buffers, and .glb.
//
// What comes out is the reference's struct Scene for that file:
//   - geometry pre-transformed to world space (node matrices composed from the
//     root; normals by the inverse transpose, renormalised), grouped the way
//     PreTransformVertices joins meshes: material by material, and within one
//     material the primitives in depth-first node order;
//   - one material per glTF material plus the default glTF material the
//     importer appends for primitives without one, read as
//     SceneParser::read_material_properties (SceneParser.cpp:392-436) reads the
//     importer's keys: base colour, emission x KHR_materials_emissive_strength,
//     metallic, roughness, and the KHR_materials_{anisotropy, sheen, specular,
//     clearcoat, ior, transmission, volume} factors;
//   - the emissive triangles (ThreadFunctions.cpp:136-190: every triangle of a
//     material with a non-zero emission colour);
//   - the first camera, placed by the node that holds it.
// Textures, texture coordinates, skins, morph targets, animations, point and
// line primitives and sparse or quantised accessors are not read.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mipt.h"

namespace mipt {

mipt_camera look_at_camera(const double eye[3], const double target[3], const double up[3], double vfov, int32_t w,
                           int32_t h);
void set_no_ctx_error(const std::string& msg);

namespace {

struct GltfError {
    std::string msg;
};
[[noreturn]] void bad(const std::string& m) { throw GltfError{m}; }

// ---------------------------------------------------------------------------
// JSON (RFC 8259)
// ---------------------------------------------------------------------------
struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double n = 0.0;
    std::string s;
    std::vector<Json> a;
    std::vector<std::pair<std::string, Json>> o;

    const Json* get(const char* key) const
    {
        if (kind != OBJ)
            return nullptr;
        for (const auto& kv : o)
            if (kv.first == key)
                return &kv.second;
        return nullptr;
    }
    double num(const char* key, double def) const
    {
        const Json* v = get(key);
        if (!v)
            return def;
        if (v->kind != NUM)
            bad(std::string("\"") + key + "\" is not a number");
        return v->n;
    }
    int64_t index(const char* key, int64_t def) const
    {
        const Json* v = get(key);
        if (!v)
            return def;
        if (v->kind != NUM || v->n < 0.0 || v->n != std::floor(v->n) || v->n > 2147483647.0)
            bad(std::string("\"") + key + "\" is not a non-negative integer");
        return (int64_t)v->n;
    }
    std::string str(const char* key, const char* def) const
    {
        const Json* v = get(key);
        if (!v)
            return def;
        if (v->kind != STR)
            bad(std::string("\"") + key + "\" is not a string");
        return v->s;
    }
    // numbers of an array member of `count` entries, or `def` without it
    std::vector<double> nums(const char* key, size_t count, const std::vector<double>& def) const
    {
        const Json* v = get(key);
        if (!v)
            return def;
        if (v->kind != ARR || v->a.size() != count)
            bad(std::string("\"") + key + "\" is not an array of " + std::to_string(count) + " numbers");
        std::vector<double> out(count);
        for (size_t i = 0; i < count; i++)
        {
            if (v->a[i].kind != NUM)
                bad(std::string("\"") + key + "\" holds a non-number");
            out[i] = v->a[i].n;
        }
        return out;
    }
    const std::vector<Json>& arr(const char* key) const
    {
        static const std::vector<Json> empty;
        const Json* v = get(key);
        if (!v)
            return empty;
        if (v->kind != ARR)
            bad(std::string("\"") + key + "\" is not an array");
        return v->a;
    }
};

class JsonParser {
public:
    JsonParser(const char* begin, const char* end) : p_(begin), end_(end) {}
    Json document()
    {
        Json v = value(0);
        ws();
        if (p_ != end_)
            bad("trailing characters after the JSON document");
        return v;
    }

private:
    const char* p_;
    const char* end_;

    void ws()
    {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\n' || *p_ == '\r' || *p_ == '\t'))
            p_++;
    }
    void expect(char c)
    {
        ws();
        if (p_ >= end_ || *p_ != c)
            bad(std::string("JSON: expected '") + c + "'");
        p_++;
    }
    void literal(const char* word)
    {
        const size_t n = std::strlen(word);
        if ((size_t)(end_ - p_) < n || std::strncmp(p_, word, n) != 0)
            bad("JSON: bad literal");
        p_ += n;
    }
    static void utf8(std::string& out, uint32_t cp)
    {
        if (cp < 0x80)
            out += (char)cp;
        else if (cp < 0x800)
        {
            out += (char)(0xc0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3f));
        }
        else if (cp < 0x10000)
        {
            out += (char)(0xe0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3f));
            out += (char)(0x80 | (cp & 0x3f));
        }
        else
        {
            out += (char)(0xf0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3f));
            out += (char)(0x80 | ((cp >> 6) & 0x3f));
            out += (char)(0x80 | (cp & 0x3f));
        }
    }
    uint32_t hex4()
    {
        if (end_ - p_ < 4)
            bad("JSON: short \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; i++)
        {
            const char c = *p_++;
            v <<= 4;
            if (c >= '0' && c <= '9')
                v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f')
                v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F')
                v |= (uint32_t)(c - 'A' + 10);
            else
                bad("JSON: bad \\u escape");
        }
        return v;
    }
    std::string string()
    {
        expect('"');
        std::string out;
        while (true)
        {
            if (p_ >= end_)
                bad("JSON: unterminated string");
            const char c = *p_++;
            if (c == '"')
                return out;
            if ((unsigned char)c < 0x20)
                bad("JSON: control character in a string");
            if (c != '\\')
            {
                out += c;
                continue;
            }
            if (p_ >= end_)
                bad("JSON: unterminated escape");
            const char e = *p_++;
            switch (e)
            {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u':
            {
                uint32_t cp = hex4();
                if (cp >= 0xd800 && cp < 0xdc00 && end_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u')
                {
                    p_ += 2;
                    const uint32_t lo = hex4();
                    if (lo < 0xdc00 || lo >= 0xe000)
                        bad("JSON: bad surrogate pair");
                    cp = 0x10000 + ((cp - 0xd800) << 10) + (lo - 0xdc00);
                }
                utf8(out, cp);
                break;
            }
            default: bad("JSON: bad escape");
            }
        }
    }
    double number()
    {
        const char* s = p_;
        while (p_ < end_ && (std::strchr("+-0123456789.eE", *p_) != nullptr))
            p_++;
        const std::string tok(s, p_);
        if (tok.empty())
            bad("JSON: unexpected character");
        char* stop = nullptr;
        const double v = std::strtod(tok.c_str(), &stop);
        if (stop != tok.c_str() + tok.size())
            bad("JSON: bad number '" + tok + "'");
        return v;
    }
    Json value(int depth)
    {
        if (depth > 200)
            bad("JSON: nested too deeply");
        ws();
        if (p_ >= end_)
            bad("JSON: unexpected end");
        Json v;
        switch (*p_)
        {
        case '{':
            v.kind = Json::OBJ;
            p_++;
            ws();
            if (p_ < end_ && *p_ == '}')
            {
                p_++;
                return v;
            }
            while (true)
            {
                std::string key = string();
                expect(':');
                Json item = value(depth + 1);
                v.o.emplace_back(std::move(key), std::move(item));
                ws();
                if (p_ < end_ && *p_ == ',')
                {
                    p_++;
                    continue;
                }
                expect('}');
                return v;
            }
        case '[':
            v.kind = Json::ARR;
            p_++;
            ws();
            if (p_ < end_ && *p_ == ']')
            {
                p_++;
                return v;
            }
            while (true)
            {
                v.a.push_back(value(depth + 1));
                ws();
                if (p_ < end_ && *p_ == ',')
                {
                    p_++;
                    continue;
                }
                expect(']');
                return v;
            }
        case '"':
            v.kind = Json::STR;
            v.s = string();
            return v;
        case 't':
            literal("true");
            v.kind = Json::BOOL;
            v.b = true;
            return v;
        case 'f':
            literal("false");
            v.kind = Json::BOOL;
            return v;
        case 'n':
            literal("null");
            return v;
        default:
            v.kind = Json::NUM;
            v.n = number();
            return v;
        }
    }
};

// ---------------------------------------------------------------------------
// Files and buffers
// ---------------------------------------------------------------------------
std::vector<uint8_t> read_file(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        bad("cannot open " + path);
    f.seekg(0, std::ios::end);
    const std::streamoff size = f.tellg();
    if (size < 0)
        bad("cannot read " + path);
    f.seekg(0, std::ios::beg);
    std::vector<uint8_t> data((size_t)size);
    if (size > 0 && !f.read(reinterpret_cast<char*>(data.data()), size))
        bad("cannot read " + path);
    return data;
}

std::vector<uint8_t> base64(const std::string& s, size_t from)
{
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < s.size(); i++)
    {
        if (s[i] == '=')
            break;
        const int v = val(s[i]);
        if (v < 0)
            bad("bad base64 in a data URI");
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8)
        {
            bits -= 8;
            out.push_back((uint8_t)(acc >> bits));
        }
    }
    return out;
}

// %XX escapes of a relative URI
std::string uri_decode(const std::string& u)
{
    std::string out;
    for (size_t i = 0; i < u.size(); i++)
    {
        if (u[i] == '%' && i + 2 < u.size())
        {
            const std::string hx = u.substr(i + 1, 2);
            char* stop = nullptr;
            const long v = std::strtol(hx.c_str(), &stop, 16);
            if (stop == hx.c_str() + 2)
            {
                out += (char)v;
                i += 2;
                continue;
            }
        }
        out += u[i];
    }
    return out;
}

uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }

struct Document {
    Json root;
    std::vector<std::vector<uint8_t>> buffers;
};

Document load_document(const std::string& path)
{
    Document d;
    std::vector<uint8_t> file = read_file(path);
    std::vector<uint8_t> glb_bin;
    bool have_bin = false;
    const char* js_begin;
    const char* js_end;
    if (file.size() >= 12 && std::memcmp(file.data(), "glTF", 4) == 0)
    {
        // binary glTF: 12-byte header, a JSON chunk, an optional BIN chunk
        if (rd32(file.data() + 4) != 2)
            bad("only glTF 2.0 binaries are read");
        const size_t total = std::min<size_t>(rd32(file.data() + 8), file.size());
        size_t pos = 12;
        js_begin = js_end = nullptr;
        while (pos + 8 <= total)
        {
            const uint32_t len = rd32(file.data() + pos), type = rd32(file.data() + pos + 4);
            pos += 8;
            if ((size_t)len > total - pos)
                bad("truncated GLB chunk");
            if (type == 0x4e4f534au && !js_begin)
            {
                js_begin = reinterpret_cast<const char*>(file.data() + pos);
                js_end = js_begin + len;
            }
            else if (type == 0x004e4942u && !have_bin)
            {
                glb_bin.assign(file.data() + pos, file.data() + pos + len);
                have_bin = true;
            }
            pos += len;
        }
        if (!js_begin)
            bad("GLB without a JSON chunk");
    }
    else
    {
        js_begin = reinterpret_cast<const char*>(file.data());
        js_end = js_begin + file.size();
    }
    d.root = JsonParser(js_begin, js_end).document();
    if (d.root.kind != Json::OBJ)
        bad("the glTF document is not a JSON object");
    const Json* asset = d.root.get("asset");
    if (!asset || asset->str("version", "").rfind("2", 0) != 0)
        bad("not a glTF 2.x asset");
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
    const std::vector<Json>& bufs = d.root.arr("buffers");
    for (size_t i = 0; i < bufs.size(); i++)
    {
        const Json& b = bufs[i];
        const size_t length = (size_t)b.index("byteLength", 0);
        std::vector<uint8_t> data;
        const Json* uri = b.get("uri");
        if (!uri)
        {
            if (i != 0 || !have_bin)
                bad("buffer " + std::to_string(i) + " has no uri and no GLB binary chunk");
            data = glb_bin;
        }
        else
        {
            const std::string u = b.str("uri", "");
            if (u.rfind("data:", 0) == 0)
            {
                const size_t comma = u.find(',');
                if (comma == std::string::npos || u.substr(0, comma).find(";base64") == std::string::npos)
                    bad("only base64 data URIs are read");
                data = base64(u, comma + 1);
            }
            else
                data = read_file(dir + uri_decode(u));
        }
        if (data.size() < length)
            bad("buffer " + std::to_string(i) + " is shorter than its byteLength");
        d.buffers.push_back(std::move(data));
    }
    return d;
}

// ---------------------------------------------------------------------------
// Accessors
// ---------------------------------------------------------------------------
int components(const std::string& type)
{
    if (type == "SCALAR") return 1;
    if (type == "VEC2") return 2;
    if (type == "VEC3") return 3;
    if (type == "VEC4") return 4;
    bad("unsupported accessor type " + type);
}

size_t component_size(int64_t ct)
{
    switch (ct)
    {
    case 5120: case 5121: return 1;
    case 5122: case 5123: return 2;
    case 5125: case 5126: return 4;
    default: bad("unsupported componentType " + std::to_string(ct));
    }
}

// count x comps values of an accessor as doubles (normalised integers mapped
// as the specification says), checked against its buffer view
std::vector<double> read_accessor(const Document& d, int64_t index, int want_comps, bool allow_int, size_t& count)
{
    const std::vector<Json>& accs = d.root.arr("accessors");
    if (index < 0 || (size_t)index >= accs.size())
        bad("accessor index out of range");
    const Json& a = accs[(size_t)index];
    if (a.get("sparse"))
        bad("sparse accessors are not read");
    const int comps = components(a.str("type", ""));
    if (want_comps && comps != want_comps)
        bad("accessor " + std::to_string(index) + " has " + std::to_string(comps) + " components");
    const int64_t ct = a.index("componentType", 0);
    const size_t csize = component_size(ct);
    if (ct != 5126 && !allow_int && !(a.get("normalized") && a.get("normalized")->b))
        bad("accessor " + std::to_string(index) + ": quantised attributes are not read");
    count = (size_t)a.index("count", 0);
    std::vector<double> out(count * (size_t)comps, 0.0);
    if (!a.get("bufferView"))
        return out; // all zeros
    const std::vector<Json>& views = d.root.arr("bufferViews");
    const int64_t vi = a.index("bufferView", 0);
    if ((size_t)vi >= views.size())
        bad("bufferView index out of range");
    const Json& v = views[(size_t)vi];
    const int64_t bi = v.index("buffer", 0);
    if ((size_t)bi >= d.buffers.size())
        bad("buffer index out of range");
    const std::vector<uint8_t>& buf = d.buffers[(size_t)bi];
    const size_t view_off = (size_t)v.index("byteOffset", 0), view_len = (size_t)v.index("byteLength", 0);
    if (view_off > buf.size() || view_len > buf.size() - view_off)
        bad("bufferView outside its buffer");
    const size_t elem = csize * (size_t)comps;
    const size_t stride = v.get("byteStride") ? (size_t)v.index("byteStride", 0) : elem;
    const size_t off = (size_t)a.index("byteOffset", 0);
    if (stride < elem || (count > 0 && (off > view_len || (count - 1) * stride + elem > view_len - off)))
        bad("accessor " + std::to_string(index) + " reads outside its bufferView");
    const bool normalized = a.get("normalized") && a.get("normalized")->b;
    const uint8_t* base = buf.data() + view_off + off;
    for (size_t i = 0; i < count; i++)
        for (int c = 0; c < comps; c++)
        {
            const uint8_t* p = base + i * stride + (size_t)c * csize;
            double x;
            switch (ct)
            {
            case 5126: { float f; std::memcpy(&f, p, 4); x = f; break; }
            case 5125: x = (double)rd32(p); break;
            case 5123: x = (double)(uint16_t)(p[0] | (p[1] << 8)); if (normalized) x /= 65535.0; break;
            case 5122: x = (double)(int16_t)(uint16_t)(p[0] | (p[1] << 8)); if (normalized) x = std::fmax(x / 32767.0, -1.0); break;
            case 5121: x = (double)p[0]; if (normalized) x /= 255.0; break;
            default: x = (double)(int8_t)p[0]; if (normalized) x = std::fmax(x / 127.0, -1.0); break;
            }
            out[i * (size_t)comps + (size_t)c] = x;
        }
    return out;
}

// ---------------------------------------------------------------------------
// Node transforms (row-major 4x4, column vectors)
// ---------------------------------------------------------------------------
struct Mat4 {
    double m[4][4];
};

Mat4 identity()
{
    Mat4 r{};
    for (int i = 0; i < 4; i++)
        r.m[i][i] = 1.0;
    return r;
}

Mat4 mul(const Mat4& a, const Mat4& b)
{
    Mat4 r{};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
        {
            double acc = 0.0;
            for (int k = 0; k < 4; k++)
                acc = acc + a.m[i][k] * b.m[k][j];
            r.m[i][j] = acc;
        }
    return r;
}

// a node's local matrix: "matrix" (column-major) or T * R * S
Mat4 local_matrix(const Json& node)
{
    Mat4 r = identity();
    if (node.get("matrix"))
    {
        const std::vector<double> m = node.nums("matrix", 16, {});
        for (int c = 0; c < 4; c++)
            for (int row = 0; row < 4; row++)
                r.m[row][c] = m[(size_t)(4 * c + row)];
        return r;
    }
    const std::vector<double> t = node.nums("translation", 3, {0.0, 0.0, 0.0});
    const std::vector<double> q = node.nums("rotation", 4, {0.0, 0.0, 0.0, 1.0});
    const std::vector<double> s = node.nums("scale", 3, {1.0, 1.0, 1.0});
    const double len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (!(len > 0.0))
        bad("zero rotation quaternion");
    const double x = q[0] / len, y = q[1] / len, z = q[2] / len, w = q[3] / len;
    const double R[3][3] = {{1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - z * w), 2.0 * (x * z + y * w)},
                            {2.0 * (x * y + z * w), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - x * w)},
                            {2.0 * (x * z - y * w), 2.0 * (y * z + x * w), 1.0 - 2.0 * (x * x + y * y)}};
    for (int i = 0; i < 3; i++)
    {
        for (int j = 0; j < 3; j++)
            r.m[i][j] = R[i][j] * s[(size_t)j];
        r.m[i][3] = t[(size_t)i];
    }
    return r;
}

// inverse transpose of the upper 3x3 (cofactors / determinant)
void normal_matrix(const Mat4& M, double N[3][3])
{
    const double(*m)[4] = M.m;
    double cof[3][3];
    cof[0][0] = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    cof[0][1] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    cof[0][2] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    cof[1][0] = m[0][2] * m[2][1] - m[0][1] * m[2][2];
    cof[1][1] = m[0][0] * m[2][2] - m[0][2] * m[2][0];
    cof[1][2] = m[0][1] * m[2][0] - m[0][0] * m[2][1];
    cof[2][0] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
    cof[2][1] = m[0][2] * m[1][0] - m[0][0] * m[1][2];
    cof[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
    const double det = m[0][0] * cof[0][0] + m[0][1] * cof[0][1] + m[0][2] * cof[0][2];
    if (det == 0.0)
        bad("singular node transform");
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            N[i][j] = cof[i][j] / det;
}

// ---------------------------------------------------------------------------
// Materials (SceneParser::read_material_properties, SceneParser.cpp:392-436,
// through the importer's keys)
// ---------------------------------------------------------------------------
void read_material(const Json& mat, mipt_material& m)
{
    mipt_material_default(&m);
    static const Json none;
    const Json* pbr = mat.get("pbrMetallicRoughness");
    if (!pbr)
        pbr = &none;
    const std::vector<double> base = pbr->nums("baseColorFactor", 4, {1.0, 1.0, 1.0, 1.0});
    const std::vector<double> emi = mat.nums("emissiveFactor", 3, {0.0, 0.0, 0.0});
    for (int c = 0; c < 3; c++)
    {
        m.base_color[c] = (float)base[(size_t)c];
        m.emission[c] = (float)emi[(size_t)c];
    }
    m.metallic = (float)pbr->num("metallicFactor", 1.0);
    m.roughness = (float)pbr->num("roughnessFactor", 1.0);
    const Json* ext = mat.get("extensions");
    if (!ext)
        return;
    if (const Json* e = ext->get("KHR_materials_emissive_strength"))
        m.emission_strength = (float)e->num("emissiveStrength", 1.0);
    if (const Json* e = ext->get("KHR_materials_anisotropy"))
        m.anisotropy = (float)e->num("anisotropyStrength", 0.0);
    if (const Json* e = ext->get("KHR_materials_sheen"))
    {
        const std::vector<double> sc = e->nums("sheenColorFactor", 3, {0.0, 0.0, 0.0});
        for (int c = 0; c < 3; c++)
            m.sheen_color[c] = (float)sc[(size_t)c];
        m.sheen_roughness = (float)e->num("sheenRoughnessFactor", 0.0);
        m.sheen = 1.0f;
    }
    if (const Json* e = ext->get("KHR_materials_specular"))
    {
        m.specular = (float)e->num("specularFactor", 1.0);
        m.specular_tint = 1.0f;
        for (int c = 0; c < 3; c++)
            m.specular_color[c] = 1.0f;
    }
    if (const Json* e = ext->get("KHR_materials_clearcoat"))
    {
        m.coat = (float)e->num("clearcoatFactor", 0.0);
        m.coat_roughness = (float)e->num("clearcoatRoughnessFactor", 0.0);
    }
    if (const Json* e = ext->get("KHR_materials_ior"))
        m.ior = (float)e->num("ior", 1.5);
    if (const Json* e = ext->get("KHR_materials_transmission"))
        m.specular_transmission = (float)e->num("transmissionFactor", 0.0);
    if (const Json* e = ext->get("KHR_materials_volume"))
    {
        const std::vector<double> ac = e->nums("attenuationColor", 3, {1.0, 1.0, 1.0});
        for (int c = 0; c < 3; c++)
            m.absorption_color[c] = (float)ac[(size_t)c];
        if (e->get("attenuationDistance"))
            m.absorption_at_distance = (float)e->num("attenuationDistance", 1.0);
    }
}

// ---------------------------------------------------------------------------
// Scene assembly
// ---------------------------------------------------------------------------
struct Instance {
    int64_t material;  // glTF material, or the number of materials for the default one
    Mat4 world;
    const Json* primitive;
};

struct Loaded {
    std::vector<float> verts, normals;
    std::vector<uint8_t> has_normals;
    std::vector<int32_t> idx, mat_idx, emissive;
    std::vector<mipt_material> mats;
    mipt_camera camera{};
};

void visit(const Document& d, int64_t node_index, const Mat4& parent, int depth, std::vector<Instance>& inst,
           int64_t& camera_node, Mat4& camera_world)
{
    const std::vector<Json>& nodes = d.root.arr("nodes");
    if (node_index < 0 || (size_t)node_index >= nodes.size())
        bad("node index out of range");
    if (depth > 64)
        bad("node hierarchy deeper than 64 (or cyclic)");
    const Json& node = nodes[(size_t)node_index];
    const Mat4 world = mul(parent, local_matrix(node));
    if (node.get("camera") && camera_node < 0 && node.index("camera", 0) == 0)
    {
        camera_node = node_index;
        camera_world = world;
    }
    if (node.get("mesh"))
    {
        const std::vector<Json>& meshes = d.root.arr("meshes");
        const int64_t mi = node.index("mesh", 0);
        if ((size_t)mi >= meshes.size())
            bad("mesh index out of range");
        const int64_t nmat = (int64_t)d.root.arr("materials").size();
        for (const Json& prim : meshes[(size_t)mi].arr("primitives"))
        {
            const int64_t mode = prim.index("mode", 4);
            if (mode < 4 || mode > 6)
                continue; // points and lines
            int64_t mat = prim.index("material", nmat);
            if (mat > nmat)
                bad("material index out of range");
            inst.push_back({mat, world, &prim});
        }
    }
    for (const Json& c : node.arr("children"))
    {
        if (c.kind != Json::NUM)
            bad("bad child index");
        visit(d, (int64_t)c.n, world, depth + 1, inst, camera_node, camera_world);
    }
}

void append_primitive(const Document& d, const Instance& in, Loaded& L)
{
    const Json& prim = *in.primitive;
    const Json* attrs = prim.get("attributes");
    if (!attrs || !attrs->get("POSITION"))
        bad("a primitive without POSITION");
    size_t nv = 0;
    const std::vector<double> pos = read_accessor(d, attrs->index("POSITION", 0), 3, false, nv);
    std::vector<double> nrm;
    if (attrs->get("NORMAL"))
    {
        size_t nn = 0;
        nrm = read_accessor(d, attrs->index("NORMAL", 0), 3, false, nn);
        if (nn != nv)
            bad("NORMAL and POSITION counts differ");
    }
    std::vector<uint32_t> ind;
    if (prim.get("indices"))
    {
        size_t ni = 0;
        const std::vector<double> iv = read_accessor(d, prim.index("indices", 0), 1, true, ni);
        ind.resize(ni);
        for (size_t i = 0; i < ni; i++)
        {
            if (iv[i] >= (double)nv)
                bad("vertex index out of range");
            ind[i] = (uint32_t)iv[i];
        }
    }
    else
    {
        ind.resize(nv);
        for (size_t i = 0; i < nv; i++)
            ind[i] = (uint32_t)i;
    }
    const size_t base = L.verts.size() / 3;
    if (base + nv > (size_t)INT32_MAX)
        bad("too many vertices");
    const double(*M)[4] = in.world.m;
    double N[3][3];
    if (!nrm.empty())
        normal_matrix(in.world, N);
    for (size_t v = 0; v < nv; v++)
    {
        const double x = pos[3 * v], y = pos[3 * v + 1], z = pos[3 * v + 2];
        for (int r = 0; r < 3; r++)
            L.verts.push_back((float)(M[r][0] * x + M[r][1] * y + M[r][2] * z + M[r][3]));
        if (nrm.empty())
        {
            L.normals.insert(L.normals.end(), {0.0f, 0.0f, 0.0f});
            L.has_normals.push_back(0);
            continue;
        }
        const double a = nrm[3 * v], b = nrm[3 * v + 1], c = nrm[3 * v + 2];
        double t[3];
        for (int r = 0; r < 3; r++)
            t[r] = N[r][0] * a + N[r][1] * b + N[r][2] * c;
        const double len = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        for (int r = 0; r < 3; r++)
            L.normals.push_back(len > 0.0 ? (float)(t[r] / len) : 0.0f);
        L.has_normals.push_back(1);
    }
    // triangles, strips and fans (glTF 2.0 section 3.7.2.1)
    const int64_t mode = prim.index("mode", 4);
    const mipt_material& m = L.mats[(size_t)in.material];
    const bool emissive = m.emission[0] != 0.0f || m.emission[1] != 0.0f || m.emission[2] != 0.0f;
    auto tri = [&](uint32_t i0, uint32_t i1, uint32_t i2) {
        L.idx.push_back((int32_t)(base + i0));
        L.idx.push_back((int32_t)(base + i1));
        L.idx.push_back((int32_t)(base + i2));
        L.mat_idx.push_back((int32_t)in.material);
        if (emissive)
            L.emissive.push_back((int32_t)L.mat_idx.size() - 1);
    };
    if (mode == 4)
    {
        for (size_t i = 0; i + 2 < ind.size(); i += 3)
            tri(ind[i], ind[i + 1], ind[i + 2]);
    }
    else if (mode == 5)
    {
        for (size_t i = 0; i + 2 < ind.size(); i++)
            if (i % 2 == 0)
                tri(ind[i], ind[i + 1], ind[i + 2]);
            else
                tri(ind[i + 1], ind[i], ind[i + 2]);
    }
    else
    {
        for (size_t i = 1; i + 1 < ind.size(); i++)
            tri(ind[i], ind[i + 1], ind[0]);
    }
}

Loaded load(const std::string& path, int32_t width, int32_t height)
{
    const Document d = load_document(path);
    Loaded L;
    const std::vector<Json>& mats = d.root.arr("materials");
    L.mats.resize(mats.size() + 1);
    for (size_t i = 0; i < mats.size(); i++)
        read_material(mats[i], L.mats[i]);
    read_material(Json{Json::OBJ}, L.mats.back()); // the importer's default material

    // the scene's nodes, depth first
    const std::vector<Json>& scenes = d.root.arr("scenes");
    std::vector<Instance> inst;
    int64_t camera_node = -1;
    Mat4 camera_world = identity();
    if (!scenes.empty())
    {
        const int64_t si = d.root.index("scene", 0);
        if ((size_t)si >= scenes.size())
            bad("scene index out of range");
        for (const Json& n : scenes[(size_t)si].arr("nodes"))
        {
            if (n.kind != Json::NUM)
                bad("bad root node index");
            visit(d, (int64_t)n.n, identity(), 0, inst, camera_node, camera_world);
        }
    }
    // PreTransformVertices: one output mesh per material, in material order
    for (int64_t m = 0; m <= (int64_t)mats.size(); m++)
        for (const Instance& in : inst)
            if (in.material == m)
                append_primitive(d, in, L);
    if (L.idx.empty())
        bad("the scene holds no triangles");

    // the first camera (SceneParser::parse_camera, SceneParser.cpp:227-287):
    // at the node's origin, looking down its -Z with +Y up
    const std::vector<Json>& cams = d.root.arr("cameras");
    double eye[3] = {0.0, 0.0, 0.0}, target[3] = {0.0, 0.0, -1.0}, up[3] = {0.0, 1.0, 0.0};
    const double aspect = (double)width / (double)height;
    double vfov = 2.0 * std::atan(std::tan(40.0 / 180.0 * M_PI * 0.5) / aspect); // 40 degrees horizontal
    const Json* persp = cams.empty() ? nullptr : cams[0].get("perspective");
    if (persp)
    {
        vfov = persp->num("yfov", vfov);
        const double(*W)[4] = camera_world.m;
        double fwd[3], u[3];
        for (int r = 0; r < 3; r++)
        {
            eye[r] = W[r][3];
            fwd[r] = -W[r][2];
            u[r] = W[r][1];
        }
        const double fl = std::sqrt(fwd[0] * fwd[0] + fwd[1] * fwd[1] + fwd[2] * fwd[2]);
        if (!(fl > 0.0))
            bad("degenerate camera transform");
        for (int r = 0; r < 3; r++)
        {
            target[r] = eye[r] + fwd[r] / fl;
            up[r] = u[r];
        }
    }
    if (!(vfov > 0.0 && vfov < M_PI))
        bad("camera yfov out of (0, pi)");
    L.camera = look_at_camera(eye, target, up, vfov, width, height);
    return L;
}

template <class T> T* dup(const std::vector<T>& v)
{
    T* p = new T[v.empty() ? 1 : v.size()];
    if (!v.empty())
        std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

} // namespace
} // namespace mipt

extern "C" int mipt_scene_load_gltf(const char* path, int32_t width, int32_t height, mipt_generated_scene** out)
{
    if (!out || !path || width <= 0 || height <= 0)
        return MIPT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    try
    {
        mipt::Loaded L = mipt::load(path, width, height);
        auto* g = new mipt_generated_scene;
        std::memset(g, 0, sizeof(*g));
        g->desc.triangles_vertex_indices = mipt::dup(L.idx);
        g->desc.triangle_count = (int32_t)(L.idx.size() / 3);
        g->desc.vertices_positions = mipt::dup(L.verts);
        g->desc.vertex_count = (int32_t)(L.verts.size() / 3);
        g->desc.has_vertex_normals = mipt::dup(L.has_normals);
        g->desc.vertex_normals = mipt::dup(L.normals);
        g->desc.material_indices = mipt::dup(L.mat_idx);
        g->desc.materials = mipt::dup(L.mats);
        g->desc.material_count = (int32_t)L.mats.size();
        g->desc.emissive_triangles_primitive_indices = mipt::dup(L.emissive);
        g->desc.emissive_triangle_count = (int32_t)L.emissive.size();
        g->camera = L.camera;
        g->width = width;
        g->height = height;
        *out = g;
        return MIPT_OK;
    }
    catch (const mipt::GltfError& e)
    {
        mipt::set_no_ctx_error(std::string("glTF ") + path + ": " + e.msg);
        return MIPT_ERR_INVALID_ARGUMENT;
    }
    catch (const std::bad_alloc&)
    {
        mipt::set_no_ctx_error("out of memory loading a glTF scene");
        return MIPT_ERR_OUT_OF_MEMORY;
    }
}
