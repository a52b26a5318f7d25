// rd_json.cpp — native loader of the reference's scene configs.
//
// Reads a scene config the way Radiantium.Cli does (Radiantium.Cli/Program.cs:25-140:
// renderer, integrator, camera (+ its medium as the global medium), accel,
// assets, instanced, scene entities breadth first) and builds what
// RendererBuilder.Build builds (Config/RendererBuilder.cs:867-1010), as an
// rd_scene_desc:
//   - JSON values are read as JsonConfigProvider reads them
//     (Config/JsonConfigProvider.cs): numbers as float / int32 from their
//     text, "to_world"-style 16-number matrices transposed, the last of
//     duplicate keys wins;
//   - entity transforms ReadTransform (RendererBuilder.cs:277-305) composed as
//     Scale * Rotation * Translation under the parent's ModelToWorld;
//   - every entity creates its own material, media and area lights; an entity
//     with neither material nor medium but a light is an infinite light;
//   - OBJ models with WavefrontObjReader's rules (Radiantium.Core/WavefrontObjReader.cs:
//     triangles only, v / v/vt / v//vn / v/vt/vn, 1-based, vertices deduplicated
//     by (v, vn, vt) in face order, optional normal generation);
//   - images as TextureUtility.LoadImageFromPath delivers them
//     (Radiantium.Core/TextureUtility.cs:118-200): PNG, OpenEXR, Radiance and
//     PFM decoded by rd_image.cpp, rows flipped unless is_flip_y is false,
//     sRGB -> linear unless is_cast_to_linear is false (default: false for
//     .exr/.hdr, true otherwise), sampled through ColorBuffer.GetRGB.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/radiantium_hip.h"
#include "rd_camera.h"
#include "rd_image.h"
#include "rd_metals.h"

namespace rd {
void set_last_error(const std::string& msg);
}

namespace {

struct LoadError : std::runtime_error {
    int code;
    LoadError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void bad(const std::string& m) { throw LoadError(RD_ERR_INVALID, m); }
[[noreturn]] void unsupported(const std::string& m) { throw LoadError(RD_ERR_UNSUPPORTED, m); }

// ---------------------------------------------------------------- JSON
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    std::string s;   // string value, or the literal text of a number
    std::vector<JVal> a;
    std::vector<std::pair<std::string, JVal>> o;
    // JsonElement.GetProperty: the last property of that name
    const JVal* get(const std::string& k) const {
        if (kind != Obj) return nullptr;
        for (size_t i = o.size(); i-- > 0;)
            if (o[i].first == k) return &o[i].second;
        return nullptr;
    }
};

struct JParser {
    const char* p;
    const char* end;
    int depth = 0;
    void ws() { while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++; }
    [[noreturn]] void fail(const char* what) { bad(std::string("config JSON: ") + what); }
    static void put_utf8(std::string& out, uint32_t c) {
        if (c < 0x80) out += (char)c;
        else if (c < 0x800) { out += (char)(0xC0 | (c >> 6)); out += (char)(0x80 | (c & 0x3F)); }
        else if (c < 0x10000) {
            out += (char)(0xE0 | (c >> 12)); out += (char)(0x80 | ((c >> 6) & 0x3F)); out += (char)(0x80 | (c & 0x3F));
        } else {
            out += (char)(0xF0 | (c >> 18)); out += (char)(0x80 | ((c >> 12) & 0x3F));
            out += (char)(0x80 | ((c >> 6) & 0x3F)); out += (char)(0x80 | (c & 0x3F));
        }
    }
    uint32_t hex4() {
        if (end - p < 4) fail("bad \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; i++) {
            char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= c - '0';
            else if (c >= 'a' && c <= 'f') v |= c - 'a' + 10;
            else if (c >= 'A' && c <= 'F') v |= c - 'A' + 10;
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string str() {
        if (p >= end || *p != '"') fail("expected a string");
        p++;
        std::string out;
        while (true) {
            if (p >= end) fail("unterminated string");
            char c = *p++;
            if (c == '"') break;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') { out += c; continue; }
            if (p >= end) fail("unterminated escape");
            char e = *p++;
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t c1 = hex4();
                    if (c1 >= 0xD800 && c1 < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        uint32_t c2 = hex4();
                        if (c2 < 0xDC00 || c2 >= 0xE000) fail("bad surrogate pair");
                        c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
                    }
                    put_utf8(out, c1);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }
    JVal value() {
        if (++depth > 128) fail("nesting too deep");
        ws();
        if (p >= end) fail("unexpected end");
        JVal v;
        char c = *p;
        if (c == '{') {
            v.kind = JVal::Obj;
            p++; ws();
            if (p < end && *p == '}') { p++; depth--; return v; }
            while (true) {
                ws();
                std::string k = str();
                ws();
                if (p >= end || *p != ':') fail("expected ':'");
                p++;
                v.o.emplace_back(std::move(k), value());
                ws();
                if (p < end && *p == ',') { p++; continue; }
                if (p < end && *p == '}') { p++; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = JVal::Arr;
            p++; ws();
            if (p < end && *p == ']') { p++; depth--; return v; }
            while (true) {
                v.a.push_back(value());
                ws();
                if (p < end && *p == ',') { p++; continue; }
                if (p < end && *p == ']') { p++; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = JVal::Str;
            v.s = str();
        } else if (c == 't' && end - p >= 4 && !strncmp(p, "true", 4)) {
            v.kind = JVal::Bool; v.b = true; p += 4;
        } else if (c == 'f' && end - p >= 5 && !strncmp(p, "false", 5)) {
            v.kind = JVal::Bool; v.b = false; p += 5;
        } else if (c == 'n' && end - p >= 4 && !strncmp(p, "null", 4)) {
            v.kind = JVal::Null; p += 4;
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            const char* s0 = p;
            if (*p == '-') p++;
            if (p >= end || !(*p >= '0' && *p <= '9')) fail("bad number");
            if (*p == '0') p++;
            else while (p < end && *p >= '0' && *p <= '9') p++;
            if (p < end && *p == '.') {
                p++;
                if (p >= end || !(*p >= '0' && *p <= '9')) fail("bad number");
                while (p < end && *p >= '0' && *p <= '9') p++;
            }
            if (p < end && (*p == 'e' || *p == 'E')) {
                p++;
                if (p < end && (*p == '+' || *p == '-')) p++;
                if (p >= end || !(*p >= '0' && *p <= '9')) fail("bad number");
                while (p < end && *p >= '0' && *p <= '9') p++;
            }
            v.kind = JVal::Num;
            v.s.assign(s0, p - s0);
        } else {
            fail("unexpected character");
        }
        depth--;
        return v;
    }
};

JVal parse_json(const std::string& text) {
    JParser jp{text.data(), text.data() + text.size()};
    // a UTF-8 byte order mark is allowed, as JsonDocument allows it
    if (text.size() >= 3 && (unsigned char)text[0] == 0xEF && (unsigned char)text[1] == 0xBB && (unsigned char)text[2] == 0xBF)
        jp.p += 3;
    JVal v = jp.value();
    jp.ws();
    if (jp.p != jp.end) bad("config JSON: trailing characters");
    return v;
}

// ---- JsonConfigProvider reads
enum ParamType { PObject, PArray, PVec2, PVec3, PVec4, PMat4, PString, PNumber, PBool, PNull };
ParamType param_type(const JVal& v) {   // GetParamType (JsonConfigProvider.cs:24-64)
    switch (v.kind) {
        case JVal::Obj: return PObject;
        case JVal::Arr: {
            for (const JVal& e : v.a) if (e.kind != JVal::Num) return PArray;
            switch (v.a.size()) {
                case 2: return PVec2;
                case 3: return PVec3;
                case 4: return PVec4;
                case 16: return PMat4;
                default: return PArray;
            }
        }
        case JVal::Str: return PString;
        case JVal::Num: return PNumber;
        case JVal::Bool: return PBool;
        default: return PNull;
    }
}
float num_f(const JVal& v, const std::string& key) {   // JsonElement.GetSingle
    if (v.kind != JVal::Num) bad("\"" + key + "\" is not a number");
    return strtof(v.s.c_str(), nullptr);
}
int num_i(const JVal& v, const std::string& key) {     // JsonElement.GetInt32
    if (v.kind != JVal::Num) bad("\"" + key + "\" is not a number");
    char* e = nullptr;
    long long x = strtoll(v.s.c_str(), &e, 10);
    if (*e != 0 || x < INT32_MIN || x > INT32_MAX) bad("\"" + key + "\" is not an int32");
    return (int)x;
}
bool has(const JVal* o, const char* k) { return o && o->get(k); }
float read_f(const JVal* o, const char* k, float d) { const JVal* v = o ? o->get(k) : nullptr; return v ? num_f(*v, k) : d; }
int read_i(const JVal* o, const char* k, int d) { const JVal* v = o ? o->get(k) : nullptr; return v ? num_i(*v, k) : d; }
bool read_b(const JVal* o, const char* k, bool d) {
    const JVal* v = o ? o->get(k) : nullptr;
    if (!v) return d;
    if (v->kind != JVal::Bool) bad(std::string("\"") + k + "\" is not a bool");
    return v->b;
}
std::string read_s(const JVal* o, const char* k, const std::string& d) {
    const JVal* v = o ? o->get(k) : nullptr;
    if (!v) return d;
    if (v->kind != JVal::Str) bad(std::string("\"") + k + "\" is not a string");
    return v->s;
}
const JVal& need_arr(const JVal* v, const char* k, size_t n) {
    if (v->kind != JVal::Arr || v->a.size() < n) bad(std::string("\"") + k + "\" needs " + std::to_string(n) + " numbers");
    return *v;
}
struct V3f { float x, y, z; };
V3f read_v3(const JVal* o, const char* k, V3f d) {
    const JVal* v = o ? o->get(k) : nullptr;
    if (!v) return d;
    const JVal& a = need_arr(v, k, 3);
    return {num_f(a.a[0], k), num_f(a.a[1], k), num_f(a.a[2], k)};
}
// ReadMat4: 16 numbers into M11..M44, then transposed (JsonConfigProvider.cs:106-118)
void read_m4(const JVal* v, const char* k, float* M) {
    const JVal& a = need_arr(v, k, 16);
    float r[16];
    for (int i = 0; i < 16; i++) r[i] = num_f(a.a[i], k);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) M[4 * i + j] = r[4 * j + i];
}
const JVal* sub(const JVal* o, const char* k) { return o ? o->get(k) : nullptr; }

// ---------------------------------------------------------------- matrices (row vectors)
struct M4 {
    float m[16];
    static M4 identity() { M4 r; std::memset(r.m, 0, sizeof r.m); r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1; return r; }
    bool is_identity() const {
        M4 I = identity();
        for (int i = 0; i < 16; i++) if (m[i] != I.m[i]) return false;
        return true;
    }
};
M4 mul(const M4& a, const M4& b) { M4 r; rd::mul4(a.m, b.m, r.m); return r; }
M4 translation(V3f p) { M4 r = M4::identity(); r.m[12] = p.x; r.m[13] = p.y; r.m[14] = p.z; return r; }
M4 scaling(V3f s) { M4 r = M4::identity(); r.m[0] = s.x; r.m[5] = s.y; r.m[10] = s.z; return r; }
// Rotation about the (unit) axis by `angle` radians in the row-vector
// convention: R = c I - s [a]x + (1 - c) a a^T  (Matrix4x4.CreateFromAxisAngle).
M4 axis_angle(V3f a, float angle) {
    float s = std::sin(angle), c = std::cos(angle);
    float xx = a.x * a.x, yy = a.y * a.y, zz = a.z * a.z, xy = a.x * a.y, xz = a.x * a.z, yz = a.y * a.z;
    M4 r = M4::identity();
    r.m[0] = xx + c * (1.0f - xx); r.m[1] = xy - c * xy + s * a.z; r.m[2] = xz - c * xz - s * a.y;
    r.m[4] = xy - c * xy - s * a.z; r.m[5] = yy + c * (1.0f - yy); r.m[6] = yz - c * yz + s * a.x;
    r.m[8] = xz - c * xz + s * a.y; r.m[9] = yz - c * yz - s * a.x; r.m[10] = zz + c * (1.0f - zz);
    return r;
}
// Rotation of the unit quaternion (x, y, z, w), row-vector convention
// (Matrix4x4.CreateFromQuaternion).
M4 quaternion(float x, float y, float z, float w) {
    float xx = x * x, yy = y * y, zz = z * z, xy = x * y, wz = z * w, xz = z * x, wy = y * w, yz = y * z, wx = x * w;
    M4 r = M4::identity();
    r.m[0] = 1.0f - 2.0f * (yy + zz); r.m[1] = 2.0f * (xy + wz); r.m[2] = 2.0f * (xz - wy);
    r.m[4] = 2.0f * (xy - wz); r.m[5] = 1.0f - 2.0f * (zz + xx); r.m[6] = 2.0f * (yz + wx);
    r.m[8] = 2.0f * (xz + wy); r.m[9] = 2.0f * (yz - wx); r.m[10] = 1.0f - 2.0f * (yy + xx);
    return r;
}
V3f xform_point(V3f v, const M4& M) {   // Vector3.Transform
    return {v.x * M.m[0] + v.y * M.m[4] + v.z * M.m[8] + M.m[12], v.x * M.m[1] + v.y * M.m[5] + v.z * M.m[9] + M.m[13],
            v.x * M.m[2] + v.y * M.m[6] + v.z * M.m[10] + M.m[14]};
}
V3f xform_normal(V3f v, const M4& M) {  // Vector3.TransformNormal
    return {v.x * M.m[0] + v.y * M.m[4] + v.z * M.m[8], v.x * M.m[1] + v.y * M.m[5] + v.z * M.m[9],
            v.x * M.m[2] + v.y * M.m[6] + v.z * M.m[10]};
}

// ReadTransform (RendererBuilder.cs:277-305) -> Scale * Rotation * Translation
M4 read_transform(const JVal* t) {
    if (!t) return M4::identity();
    V3f pos = read_v3(t, "position", {0, 0, 0});
    M4 rot = M4::identity();
    if (const JVal* r = t->get("rotation")) {
        switch (param_type(*r)) {
            case PMat4: read_m4(r, "rotation", rot.m); break;
            case PObject: {
                V3f axis = read_v3(r, "axis", {0, 1, 0});
                float angle = read_f(r, "angle", 0.0f);
                rot = axis_angle(axis, angle * (3.14159265358979323846f / 180.0f));   // MathExt.Radian
                break;
            }
            case PVec4:
                rot = quaternion(num_f(r->a[0], "rotation"), num_f(r->a[1], "rotation"), num_f(r->a[2], "rotation"),
                                 num_f(r->a[3], "rotation"));
                break;
            default: bad("invalid rotation");
        }
    }
    V3f scl = read_v3(t, "scale", {1, 1, 1});
    return mul(mul(scaling(scl), rot), translation(pos));
}

// ---------------------------------------------------------------- files
std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) bad("cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}
bool file_exists(const std::string& p) { std::ifstream f(p, std::ios::binary); return (bool)f; }
// RendererBuilder.GetPath (:604-610): the location as given, else under the search path
std::string resolve(const std::string& loc, const std::string& dir) {
    if (file_exists(loc)) return loc;
    std::string p = dir.empty() ? loc : dir + "/" + loc;
    if (!file_exists(p)) bad("invalid location " + loc);
    return p;
}

// ---------------------------------------------------------------- OBJ (WavefrontObjReader)
struct TriModel {
    std::vector<V3f> p, n;
    std::vector<float> uv;   // 2 per vertex
    std::vector<int> idx;
};

// float.TryParse / int.TryParse (leading / trailing white space, sign,
// thousands separators, decimal point, exponent)
bool try_float(const char* b, const char* e, float* out) {
    while (b < e && (*b == ' ' || *b == '\t' || *b == '\r')) b++;
    while (e > b && (e[-1] == ' ' || e[-1] == '\t' || e[-1] == '\r')) e--;
    if (b >= e) return false;
    std::string s;
    const char* q = b;
    if (*q == '+' || *q == '-') s += *q++;
    bool digits = false, dot = false;
    while (q < e && ((*q >= '0' && *q <= '9') || (*q == ',' && !dot) || (*q == '.' && !dot))) {
        if (*q == '.') dot = true;
        if (*q >= '0' && *q <= '9') digits = true;
        if (*q != ',') s += *q;
        q++;
    }
    if (!digits) return false;
    if (q < e && (*q == 'e' || *q == 'E')) {
        s += *q++;
        if (q < e && (*q == '+' || *q == '-')) s += *q++;
        if (q >= e || !(*q >= '0' && *q <= '9')) return false;
        while (q < e && *q >= '0' && *q <= '9') s += *q++;
    }
    if (q != e) return false;
    *out = strtof(s.c_str(), nullptr);
    return true;
}
bool try_int(const char* b, const char* e, int* out) {
    while (b < e && (*b == ' ' || *b == '\t' || *b == '\r')) b++;
    while (e > b && (e[-1] == ' ' || e[-1] == '\t' || e[-1] == '\r')) e--;
    if (b >= e) return false;
    const char* q = b;
    if (*q == '+' || *q == '-') q++;
    if (q >= e) return false;
    for (const char* r = q; r < e; r++) if (!(*r >= '0' && *r <= '9')) return false;
    long long v = strtoll(std::string(b, e).c_str(), nullptr, 10);
    if (v < INT32_MIN || v > INT32_MAX) return false;
    *out = (int)v;
    return true;
}
// TrySplit3: positions of the first two separators; false when fewer than two
bool split3(const char* b, const char* e, char sep, int* b0, int* b1) {
    int cnt = 0;
    for (const char* q = b; q < e; q++)
        if (*q == sep) { if (cnt == 0) *b0 = (int)(q - b); else if (cnt == 1) *b1 = (int)(q - b); cnt++; }
    return cnt >= 2;
}
bool read_vec3(const char* b, const char* e, V3f* v) {
    int b0 = 0, b1 = 0;
    if (!split3(b, e, ' ', &b0, &b1)) return false;
    float x, y, z;
    bool ok = try_float(b, b + b0, &x);
    ok = try_float(b + b0 + 1, b + b1, &y) && ok;
    ok = try_float(b + b1 + 1, e, &z) && ok;
    if (ok) *v = {x, y, z};
    return ok;
}
bool read_vec2(const char* b, const char* e, float* u, float* v) {
    const char* sp = nullptr;
    for (const char* q = b; q < e; q++) if (*q == ' ') { sp = q; break; }
    if (!sp || sp == b) return false;
    float x, y;
    bool ok = try_float(b, sp, &x);
    ok = try_float(sp + 1, e, &y) && ok;
    if (ok) { *u = x; *v = y; }
    return ok;
}
struct Idx3 { int v, t, n; };
bool read_face_vertex(const char* b, const char* e, Idx3* out) {   // TryReadFaceVertex
    int b0 = 0, b1 = 0;
    if (split3(b, e, '/', &b0, &b1)) {
        int rv = 0, rvt = 0, rvn = 0;
        bool ok = try_int(b, b + b0, &rv);
        if (b1 - b0 == 1) rvt = 0;
        else ok = try_int(b + b0 + 1, b + b1, &rvt) && ok;
        ok = try_int(b + b1 + 1, e, &rvn) && ok;
        if (ok) *out = {rv, rvt, rvn};
        return ok;
    }
    const char* slash = nullptr;
    for (const char* q = b; q < e; q++) if (*q == '/') { slash = q; break; }
    if (slash && slash > b) {
        int rv = 0, rvt = 0;
        bool ok = try_int(b, slash, &rv);
        ok = try_int(slash + 1, e, &rvt) && ok;
        if (ok) *out = {rv, rvt, -1};
        return ok;
    }
    int rv = 0;
    if (!try_int(b, e, &rv)) return false;
    *out = {rv, -1, -1};
    return true;
}

struct ObjFile {
    std::vector<V3f> pos, nrm;
    std::vector<float> tex;            // 2 per texcoord
    struct Face { Idx3 c[3]; };        // 0-based (parsed - 1)
    std::vector<Face> faces;
    std::vector<std::pair<std::string, std::vector<int>>> objs;

    void parse(const std::string& text) {
        size_t i = 0;
        while (i < text.size()) {
            size_t j = text.find('\n', i);
            if (j == std::string::npos) j = text.size();
            std::string line = text.substr(i, j - i);
            if (!line.empty() && line.back() == '\r') line.pop_back();
            i = j + 1;
            line_(line);
        }
    }
    static int first_non_space(const std::string& s, size_t start) {
        for (size_t k = start; k < s.size(); k++) if (s[k] != ' ') return (int)k;
        return -1;
    }
    static bool starts(const std::string& s, const char* pre) { return s.compare(0, strlen(pre), pre) == 0; }
    void line_(const std::string& line) {
        bool blank = true;
        for (char c : line) if (!isspace((unsigned char)c)) { blank = false; break; }
        if (blank || line[0] == '#') return;
        if (starts(line, "o ") || starts(line, "g ")) {
            int st = first_non_space(line, 1);
            std::string name = st < 0 ? "" : line.substr(st);
            for (auto& o : objs) if (o.first == name) return;   // duplicate name: ignored
            objs.emplace_back(name, std::vector<int>());
        } else if (starts(line, "v ")) {
            int st = first_non_space(line, 1);
            V3f v{0, 0, 0};
            if (st < 0 || !read_vec3(line.data() + st, line.data() + line.size(), &v)) v = {0, 0, 0};
            pos.push_back(v);
        } else if (starts(line, "f ")) {
            int st = first_non_space(line, 1);
            if (st < 0) return;
            const char* b = line.data() + st;
            const char* e = line.data() + line.size();
            int b0 = 0, b1 = 0;
            if (!split3(b, e, ' ', &b0, &b1)) return;
            Idx3 x, y, z;
            bool ok = read_face_vertex(b, b + b0, &x);
            ok = read_face_vertex(b + b0 + 1, b + b1, &y) && ok;
            ok = read_face_vertex(b + b1 + 1, e, &z) && ok;
            if (!ok) return;                                   // can't parse face: ignored
            Face f;
            f.c[0] = {x.v - 1, x.t - 1, x.n - 1};
            f.c[1] = {y.v - 1, y.t - 1, y.n - 1};
            f.c[2] = {z.v - 1, z.t - 1, z.n - 1};
            faces.push_back(f);
            if (!objs.empty()) objs.back().second.push_back((int)faces.size() - 1);
        } else if (starts(line, "vt ")) {
            int st = first_non_space(line, 2);
            float u = 0, v = 0;
            if (st < 0 || !read_vec2(line.data() + st, line.data() + line.size(), &u, &v)) { u = 0; v = 0; }
            tex.push_back(u);
            tex.push_back(v);
        } else if (starts(line, "vn ")) {
            int st = first_non_space(line, 2);
            V3f v{0, 1, 0};
            if (st < 0 || !read_vec3(line.data() + st, line.data() + line.size(), &v)) v = {0, 1, 0};
            nrm.push_back(v);
        }
        // mtllib, usemtl, s and unknown commands: ignored
    }

    // ToModel (WavefrontObjReader.cs:444-500)
    TriModel to_model(const std::vector<int>& face_ids, bool gen_normal) const {
        struct Key { int p, n, t; bool operator<(const Key& o) const {
            return p != o.p ? p < o.p : (n != o.n ? n < o.n : t < o.t); } };
        std::map<Key, int> unique;
        TriModel m;
        int count = 0;
        for (int fi : face_ids) {
            const Face& f = faces[fi];
            for (int k = 0; k < 3; k++) {
                Key key{f.c[k].v, f.c[k].n, f.c[k].t};
                auto it = unique.find(key);
                int index;
                if (it == unique.end()) {
                    unique.emplace(key, count);
                    index = count;
                    if (key.p < 0 || key.p >= (int)pos.size()) bad("OBJ face references a missing vertex");
                    m.p.push_back(pos[key.p]);
                    if (key.n >= 0) {
                        if (key.n >= (int)nrm.size()) bad("OBJ face references a missing normal");
                        m.n.push_back(nrm[key.n]);
                    } else if (gen_normal) {
                        m.n.push_back({0, 0, 0});
                    }
                    if (key.t >= 0) {
                        if (2 * key.t + 1 >= (int)tex.size()) bad("OBJ face references a missing texcoord");
                        m.uv.push_back(tex[2 * key.t]);
                        m.uv.push_back(tex[2 * key.t + 1]);
                    }
                    count++;
                } else {
                    index = it->second;
                }
                m.idx.push_back(index);
            }
        }
        if (gen_normal && nrm.empty()) {
            // as written: na, nb and nc all read the first corner's normal
            int tris = (int)m.idx.size() / 3;
            auto normalize = [](V3f v) {
                float l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
                return V3f{v.x / l, v.y / l, v.z / l};
            };
            auto zero = [](V3f v) { return v.x == 0 && v.y == 0 && v.z == 0; };
            for (int i = 0; i < tris; i++) {
                V3f na = m.n[m.idx[3 * i]], nb = na, nc = na;
                V3f a = m.p[m.idx[3 * i]], b = m.p[m.idx[3 * i + 1]], c = m.p[m.idx[3 * i + 2]];
                V3f e1{b.x - a.x, b.y - a.y, b.z - a.z}, e2{c.x - a.x, c.y - a.y, c.z - a.z};
                V3f nor = normalize({e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x});
                m.n[m.idx[3 * i]] = zero(na) ? nor : normalize({nor.x + na.x, nor.y + na.y, nor.z + na.z});
                m.n[m.idx[3 * i + 1]] = zero(nb) ? nor : normalize({nor.x + nb.x, nor.y + nb.y, nor.z + nb.z});
                m.n[m.idx[3 * i + 2]] = zero(nc) ? nor : normalize({nor.x + nc.x, nor.y + nc.y, nor.z + nc.z});
            }
        }
        // a file mixing vertices with and without normals / texcoords gives
        // arrays the reference indexes out of step; refuse it
        if (!m.n.empty() && m.n.size() != m.p.size()) bad("OBJ mixes vertices with and without normals");
        if (!m.uv.empty() && m.uv.size() != 2 * m.p.size()) bad("OBJ mixes vertices with and without texcoords");
        return m;
    }
};

// ---------------------------------------------------------------- images
struct Image {
    int w = 0, h = 0, channels = 0;
    std::vector<float> data;   // ColorBuffer: (x * h + y) * channels
};

bool ext_is_hdr(const std::string& loc) {   // TextureUtility.IsHdr: the extension contains "exr" or "hdr"
    size_t dot = loc.find_last_of('.');
    size_t slash = loc.find_last_of("/\\");
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return false;
    std::string ext = loc.substr(dot);
    return ext.find("exr") != std::string::npos || ext.find("hdr") != std::string::npos;
}

Image load_image(const std::string& path, const std::string& name, int need_channel, bool flip_y, bool to_linear) {
    rd::RawImage raw;
    try {
        raw = rd::decode_image(read_file(path), name);
    } catch (const rd::ImageError& e) {
        throw LoadError(e.code, e.what());
    }
    const int w = raw.w, h = raw.h, ch = raw.channels;
    const std::vector<float>& px = raw.px;
    Image im;
    im.w = w; im.h = h;
    im.channels = need_channel > 0 ? need_channel : ch;
    im.data.assign((size_t)w * h * im.channels, 0.0f);
    for (int i = 0; i < w; i++)
        for (int j = 0; j < h; j++) {
            int k = flip_y ? h - j - 1 : j;
            const float* s = px.data() + ((size_t)k * w + i) * ch;
            float* o = im.data.data() + ((size_t)i * h + j) * im.channels;
            for (int a = 0; a < im.channels; a++) o[a] = a >= ch ? 0.0f : s[a];
        }
    if (to_linear)
        for (float& e : im.data)   // Color3F.ToLinear (Radiantium.Core/Color3F.cs:86-89)
            e = e <= 0.04045f ? e * (1.0f / 12.92f) : std::pow((e + 0.055f) * (1.0f / 1.055f), 2.4f);
    return im;
}

// ---------------------------------------------------------------- the scene
struct Built {
    std::vector<float> pos, nrm, uv;
    std::vector<int32_t> idx;
    std::vector<uint8_t> tflags;
    bool any_n = false, any_uv = false;
    std::vector<float> spheres, sphere_w2m;
    bool any_sphere_rot = false;
    std::vector<rd_prim_desc> prims;
    std::vector<rd_material_desc> mats;
    std::vector<rd_texture_desc> texs;
    std::deque<std::vector<float>> texels;   // stable storage for image data
    std::vector<rd_light_desc> lights;
    std::vector<rd_medium_desc> media;
    std::vector<rd_instance_set_desc> sets;
    std::vector<rd_instance_desc> insts;
    rd_scene_desc desc{};
    std::vector<std::string> notes;
};

const MetalEtaK* find_metal(const std::string& name) {   // Spectrum.TryGetMatel
    for (const MetalEtaK& m : kMetals)
        if (name == m.name) return &m;
    return nullptr;
}

struct Loader {
    const JVal& root;
    std::string dir;
    Built& B;
    std::map<std::string, std::pair<TriModel, std::vector<std::pair<std::string, TriModel>>>> models;
    std::map<std::string, Image> images;
    int global_medium = -1;

    int add_const_tex(float r, float g, float b) {
        rd_texture_desc t{};
        t.value[0] = r; t.value[1] = g; t.value[2] = b;
        B.texs.push_back(t);
        return (int)B.texs.size() - 1;
    }
    // ReadTex2D (RendererBuilder.cs:239-275)
    int read_tex(const JVal* o, const char* key, float dr, float dg, float db) {
        const JVal* v = sub(o, key);
        if (!v) return add_const_tex(dr, dg, db);
        ParamType pt = param_type(*v);
        if (pt == PVec3) return add_const_tex(num_f(v->a[0], key), num_f(v->a[1], key), num_f(v->a[2], key));
        if (pt == PNumber) { float f = num_f(*v, key); return add_const_tex(f, f, f); }
        if (pt != PObject || !v->get("type")) bad(std::string("texture \"") + key + "\": no key type");
        std::string t = read_s(v, "type", "");
        if (t != "image") unsupported("texture type " + t);
        if (!v->get("name")) bad("image texture without a name");
        std::string name = read_s(v, "name", "");
        auto it = images.find(name);
        if (it == images.end()) bad("no image named " + name);
        const Image& im = it->second;
        rd_texture_desc td{};
        td.width = im.w;
        td.height = im.h;
        const JVal* smp = sub(v, "sampler");   // CreateTextureSampler (:551-567)
        std::string wrap = smp ? read_s(smp, "wrap", "Clamp") : "Clamp";
        std::string filt = smp ? read_s(smp, "filter", "Nearest") : "Nearest";
        if (wrap != "Clamp" && wrap != "Repeat") bad("unknown wrap mode " + wrap);
        if (filt != "Nearest" && filt != "Linear") bad("unknown filter mode " + filt);
        td.wrap = wrap == "Repeat";
        td.filter = filt == "Linear";
        // ColorBuffer.GetRGB (Radiantium.Core/ColorBuffer.cs:44-62) of every texel
        std::vector<float> rgb((size_t)im.w * im.h * 3, 0.0f);
        size_t total = im.data.size();
        for (size_t q = 0; q < (size_t)im.w * im.h; q++) {
            size_t start = q * im.channels;
            for (int c = 0; c < 3; c++) {
                if (im.channels >= 3) rgb[3 * q + c] = im.data[start + c];
                else rgb[3 * q + c] = im.channels >= c ? 0.0f : (start + c < total ? im.data[start + c] : 0.0f);
            }
        }
        B.texels.push_back(std::move(rgb));
        td.data = B.texels.back().data();
        B.texs.push_back(td);
        return (int)B.texs.size() - 1;
    }

    int create_material(const JVal* m) {   // the material builders (RendererBuilder.cs:93-210)
        if (!m->get("type")) bad("material: no key type");
        std::string t = read_s(m, "type", "");
        rd_material_desc d{};
        d.eta_a = read_f(m, "etaA", 1.000277f);
        d.eta_b = read_f(m, "etaB", 1.5046f);
        d.kd = 0.5f; d.ks = 0.5f;
        const MetalEtaK* cu = find_metal("Cu");
        for (int k = 0; k < 3; k++) { d.eta[k] = cu->eta[k]; d.k[k] = cu->k[k]; }
        auto dist = [&](const char* def) {
            std::string s = read_s(m, "dist", def);
            if (s == "Beckmann") return 0;
            if (s == "GGX") return 1;
            bad("unknown microfacet distribution " + s);
        };
        if (t == "diffuse") {
            d.type = RD_MAT_DIFFUSE;
            d.tex[0] = read_tex(m, "kd", 0.5f, 0.5f, 0.5f);
            d.two_sided = read_b(m, "is_two_side", false);
        } else if (t == "perfect_glass") {
            d.type = RD_MAT_PERFECT_GLASS;
            d.tex[0] = read_tex(m, "r", 1, 1, 1);
            d.tex[1] = read_tex(m, "t", 1, 1, 1);
        } else if (t == "perfect_mirror") {
            d.type = RD_MAT_PERFECT_MIRROR;
            d.tex[0] = read_tex(m, "r", 1, 1, 1);
            d.two_sided = read_b(m, "is_two_side", false);
        } else if (t == "rough_plastic") {
            d.type = RD_MAT_ROUGH_PLASTIC;
            d.tex[0] = read_tex(m, "r", 0.5f, 0.5f, 0.5f);
            d.dist = dist("Beckmann");
            d.tex[1] = read_tex(m, "roughness", 0.3f, 0.3f, 0.3f);
            d.tex[2] = read_tex(m, "anisotropic", 0, 0, 0);
            d.kd = read_f(m, "kd", 0.5f);
            d.ks = read_f(m, "ks", 0.5f);
            d.two_sided = read_b(m, "is_two_side", false);
        } else if (t == "rough_metal") {
            d.type = RD_MAT_ROUGH_METAL;
            d.dist = dist("GGX");
            d.tex[0] = read_tex(m, "roughness", 0.3f, 0.3f, 0.3f);
            d.tex[1] = read_tex(m, "anisotropic", 0, 0, 0);
            if (const JVal* mt = m->get("metal_type")) {
                if (mt->kind == JVal::Str) {
                    const MetalEtaK* me = find_metal(mt->s);
                    if (!me) bad("no metal type named " + mt->s);
                    for (int k = 0; k < 3; k++) { d.eta[k] = me->eta[k]; d.k[k] = me->k[k]; }
                } else {
                    // as written, "eta" / "k" are read from the material, not from the metal_type object
                    V3f e = read_v3(m, "eta", {cu->eta[0], cu->eta[1], cu->eta[2]});
                    V3f k = read_v3(m, "k", {cu->k[0], cu->k[1], cu->k[2]});
                    d.eta[0] = e.x; d.eta[1] = e.y; d.eta[2] = e.z;
                    d.k[0] = k.x; d.k[1] = k.y; d.k[2] = k.z;
                }
            }
            d.two_sided = read_b(m, "is_two_side", false);
        } else if (t == "rough_glass") {
            d.type = RD_MAT_ROUGH_GLASS;
            d.tex[0] = read_tex(m, "r", 1, 1, 1);
            d.tex[1] = read_tex(m, "t", 1, 1, 1);
            d.tex[2] = read_tex(m, "roughness", 0.3f, 0.3f, 0.3f);
            d.tex[3] = read_tex(m, "anisotropic", 0, 0, 0);
            d.dist = dist("GGX");
        } else if (t == "disney") {
            d.type = RD_MAT_DISNEY;
            static const char* names[15] = {"base_color", "metallic", "roughness", "eta", "specular", "specular_tint",
                                            "anisotropic", "sheen", "sheen_tint", "clearcoat", "clearcoat_gloss",
                                            "scatting_distance", "transmission", "transmission_roughness", "flatness"};
            static const float defs[15] = {0.5f, 0, 1, 1.5f, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int k = 0; k < 15; k++) {
                if (k == 12) d.thin = read_b(m, "is_thin", false);
                d.tex[k] = read_tex(m, names[k], defs[k], defs[k], defs[k]);
            }
        } else if (t == "subsurface") {   // RendererBuilder.cs:206-215
            d.type = RD_MAT_SUBSURFACE;
            d.tex[0] = read_tex(m, "r", 1, 1, 1);
            d.tex[1] = read_tex(m, "t", 1, 1, 1);
            d.tex[2] = read_tex(m, "a", 1, 1, 1);
            d.tex[3] = read_tex(m, "scatting_distance", 0, 0, 0);
        } else {
            unsupported("material type " + t);
        }
        B.mats.push_back(d);   // unused texture slots stay 0 (a valid index: every material adds one)
        return (int)B.mats.size() - 1;
    }

    // CreateMedium (:569-581): a new medium per call, or the global one
    int create_medium(const JVal* m) {
        if (!m) return -1;
        if (m->get("is_global") && read_b(m, "is_global", false)) return global_medium;
        if (!m->get("type")) bad("medium: no key type");
        std::string t = read_s(m, "type", "");
        if (t != "homogeneous") unsupported("medium type " + t);
        rd_medium_desc d{};
        V3f sa = read_v3(m, "sigma_a", {0, 0, 0}), ss = read_v3(m, "sigma_s", {0, 0, 0});
        d.sigma_a[0] = sa.x; d.sigma_a[1] = sa.y; d.sigma_a[2] = sa.z;
        d.sigma_s[0] = ss.x; d.sigma_s[1] = ss.y; d.sigma_s[2] = ss.z;
        d.g = read_f(m, "g", 0.0f);
        d.scale = read_f(m, "scale", 1.0f);
        B.media.push_back(d);
        return (int)B.media.size() - 1;
    }

    // Triangles of a model through m (TriangleMesh ctor, Shapes/Triangle.cs:15-45):
    // returns the first new triangle index
    int add_mesh(const TriModel& tm, const M4& m) {
        bool ident = m.is_identity();
        int v0 = (int)B.pos.size() / 3;
        size_t nv = tm.p.size();
        for (size_t i = 0; i < nv; i++) {
            V3f p = ident ? tm.p[i] : xform_point(tm.p[i], m);
            B.pos.insert(B.pos.end(), {p.x, p.y, p.z});
            V3f n{0, 0, 0};
            if (!tm.n.empty()) n = ident ? tm.n[i] : xform_normal(tm.n[i], m);
            B.nrm.insert(B.nrm.end(), {n.x, n.y, n.z});
            B.uv.push_back(tm.uv.empty() ? 0.0f : tm.uv[2 * i]);
            B.uv.push_back(tm.uv.empty() ? 0.0f : tm.uv[2 * i + 1]);
        }
        if (!tm.n.empty()) B.any_n = true;
        if (!tm.uv.empty()) B.any_uv = true;
        uint8_t fl = (tm.n.empty() ? 0 : 1) | (tm.uv.empty() ? 0 : 2);
        int t0 = (int)B.tflags.size();
        for (size_t k = 0; k < tm.idx.size(); k++) B.idx.push_back(v0 + tm.idx[k]);
        for (size_t k = 0; k < tm.idx.size() / 3; k++) B.tflags.push_back(fl);
        return t0;
    }

    const TriModel& find_model(const JVal* shape) {   // FindTriangleModel (:847-865)
        if (!shape->get("name")) bad("model shape: no key name");
        std::string name = read_s(shape, "name", "");
        auto it = models.find(name);
        if (it == models.end()) bad("no model named " + name);
        if (const JVal* c = shape->get("child")) {
            if (c->kind != JVal::Str) bad("\"child\" is not a string");
            for (auto& ch : it->second.second) if (ch.first == c->s) return ch.second;
            bad("can't find model " + name + "$" + c->s);
        }
        return it->second.first;
    }

    void load_assets() {
        const JVal* assets = root.get("assets");
        if (!assets) return;
        if (assets->kind != JVal::Arr) bad("\"assets\" is not an array");
        for (const JVal& a : assets->a) {
            if (!a.get("type")) bad("asset: no key type");
            std::string t = read_s(&a, "type", "");
            if (t == "model") {
                if (!a.get("name")) bad("model asset: no key name");
                if (!a.get("location")) bad("model asset: no key location");
                std::string name = read_s(&a, "name", ""), loc = read_s(&a, "location", "");
                bool subm = read_b(&a, "is_store_sub_model", false), gen = read_b(&a, "will_gen_normal", false);
                ObjFile obj;
                obj.parse(read_file(resolve(loc, dir)));
                std::vector<int> all(obj.faces.size());
                for (size_t k = 0; k < all.size(); k++) all[k] = (int)k;
                auto& entry = models[name];
                entry.first = obj.to_model(all, gen);
                entry.second.clear();
                if (subm)
                    for (auto& o : obj.objs) {
                        // ToModel(objName) walks Faces.Count entries of the object's
                        // face list (WavefrontObjReader.cs:449-452): only an object
                        // holding every face of the file can be read
                        if (o.second.size() != obj.faces.size())
                            bad("is_store_sub_model on " + loc + ": object " + o.first + " does not hold every face");
                        entry.second.emplace_back(o.first, obj.to_model(o.second, gen));
                    }
            } else if (t == "image") {
                if (!a.get("name")) bad("image asset: no key name");
                if (!a.get("location")) bad("image asset: no key location");
                std::string name = read_s(&a, "name", ""), loc = read_s(&a, "location", "");
                bool lin = read_b(&a, "is_cast_to_linear", !ext_is_hdr(loc));
                images[name] = load_image(resolve(loc, dir), name, read_i(&a, "channel", -1),
                                          read_b(&a, "is_flip_y", true), lin);
            } else {
                bad("unknown asset type " + t);
            }
        }
    }

    void build(rd_render_params* params) {
        // renderer: the one the integrator names (CreateRenderer, RendererBuilder.cs:514-526;
        // block_based / bdpt read "spp", adjoint_particle "particle_count", :29-48)
        int spp = 256;
        int64_t particles = 16384;
        const JVal* rnd = root.get("renderer");
        if (rnd) {
            spp = read_i(rnd, "spp", 256);
            particles = read_i(rnd, "particle_count", 16384);
        }
        // integrator
        rd_integrator_desc I{};
        I.type = RD_INTEGRATOR_PATH; I.max_depth = -1; I.min_depth = 3; I.rr_threshold = 0.99f;
        I.method = RD_METHOD_MIS; I.strategy = RD_STRATEGY_UNIFORM;
        if (const JVal* it = root.get("integrator")) {
            if (!it->get("type")) bad("integrator: no key type");
            std::string t = read_s(it, "type", "");
            if (t != "path" && t != "vol_path" && t != "adjoint_particle" && t != "bdpt")
                unsupported("integrator " + t + " is not on this path");
            I.type = t == "path" ? RD_INTEGRATOR_PATH
                                 : (t == "vol_path" ? RD_INTEGRATOR_VOL_PATH
                                                    : (t == "bdpt" ? RD_INTEGRATOR_BDPT : RD_INTEGRATOR_PARTICLE));
            // without a renderer block the builder makes a BlockBasedRenderer,
            // which casts the integrator to MonteCarloIntegrator (:518-519)
            if (!rnd && (I.type == RD_INTEGRATOR_PARTICLE || I.type == RD_INTEGRATOR_BDPT))
                bad("integrator " + t + " needs a renderer block");
            // bdpt: max_depth 7, min_depth 5 (:80-85); the others -1, 3 and rr 0.99
            I.max_depth = read_i(it, "max_depth", t == "bdpt" ? 7 : -1);
            I.min_depth = read_i(it, "min_depth", t == "bdpt" ? 5 : 3);
            I.rr_threshold = read_f(it, "rr_threshold", 0.99f);
            if (t == "path") {
                std::string m = read_s(it, "method", "Mis");
                if (m == "Bxdf") I.method = RD_METHOD_BXDF;            // PathSampleMethod (PathTracer.cs:12-17)
                else if (m == "Nee") I.method = RD_METHOD_NEE;
                else if (m == "Mis") I.method = RD_METHOD_MIS;
                else bad("unknown path method " + m);
            }
            std::string s = read_s(it, "strategy", "Uniform");
            if (s == "Uniform") I.strategy = RD_STRATEGY_UNIFORM;
            else if (s == "All") I.strategy = RD_STRATEGY_ALL;
            else bad("unknown light strategy " + s);
        } else {
            bad("config has no integrator");
        }
        // camera (CreateCamera, :479-513) and its medium as the global medium
        const JVal* cam = root.get("camera");
        if (!cam) bad("config has no camera");
        if (!cam->get("type")) bad("camera: no key type");
        if (read_s(cam, "type", "") != "perspective") unsupported("camera type " + read_s(cam, "type", ""));
        rd_camera_desc C{};
        C.width = 1280; C.height = 720;
        if (const JVal* sc = cam->get("screen")) {
            const JVal& a = need_arr(sc, "screen", 2);
            C.width = num_i(a.a[0], "screen");
            C.height = num_i(a.a[1], "screen");
        }
        C.fov = read_f(cam, "fov", 60.0f);
        C.near_z = read_f(cam, "near", 0.001f);
        C.far_z = read_f(cam, "far", 1000.0f);
        if (cam->get("origin")) {
            C.use_matrix = 0;
            V3f o = read_v3(cam, "origin", {0, 0, 3}), t = read_v3(cam, "target", {0, 0, 0}), u = read_v3(cam, "up", {0, 1, 0});
            C.origin[0] = o.x; C.origin[1] = o.y; C.origin[2] = o.z;
            C.target[0] = t.x; C.target[1] = t.y; C.target[2] = t.z;
            C.up[0] = u.x; C.up[1] = u.y; C.up[2] = u.z;
        } else {
            C.use_matrix = 1;
            M4 m = M4::identity();
            if (const JVal* tw = cam->get("to_world")) read_m4(tw, "to_world", m.m);
            std::memcpy(C.to_world, m.m, sizeof m.m);
        }
        // scene accelerator (RendererBuilder.cs:86-87): the SAH BVH or the loose octree
        int max_prim = 1;
        int accel = RD_ACCEL_BVH, oct_depth = 10, oct_count = 10;
        float oct_out = 2.0f;
        if (const JVal* ac = root.get("accel")) {
            if (!ac->get("type")) bad("accel: no key type");
            std::string t = read_s(ac, "type", "");
            if (t == "bvh") {
                max_prim = read_i(ac, "max_prim", 1);
                std::string split = read_s(ac, "split", "SAH");
                if (split != "SAH") unsupported("BVH split method " + split);
            } else if (t == "octree") {
                accel = RD_ACCEL_OCTREE;
                oct_depth = read_i(ac, "max_depth", 10);
                oct_count = read_i(ac, "max_count", 10);
                oct_out = read_f(ac, "out_bound", 2.0f);
            } else {
                bad("unknown accelerator " + t);
            }
        }
        load_assets();
        // the global medium object is created before the entities (:923)
        if (const JVal* gm = cam->get("medium")) global_medium = create_medium(gm);

        // instanced entries (Program.cs:76-95, RendererBuilder.cs:895-921)
        std::vector<int> entry_set;
        if (const JVal* ins = root.get("instanced")) {
            if (ins->kind != JVal::Arr) bad("\"instanced\" is not an array");
            for (const JVal& e : ins->a) {
                M4 local = read_transform(e.get("transform"));
                int mp = 1;
                if (const JVal* ac = e.get("accel")) {
                    if (!ac->get("type")) bad("instanced accel: no key type");
                    if (read_s(ac, "type", "") == "bvh") mp = read_i(ac, "max_prim", 1);
                }
                rd_instance_set_desc sd{};
                sd.first_triangle = (int)B.tflags.size();
                sd.bvh_max_prim = mp;
                if (const JVal* sh = e.get("shape")) {
                    if (!sh->get("type")) bad("invalid shape config");
                    std::string t = read_s(sh, "type", "");
                    if (t != "model") unsupported("instanced shape " + t + " (instance sets hold triangles)");
                    add_mesh(find_model(sh), local);
                }
                sd.num_triangles = (int)B.tflags.size() - sd.first_triangle;
                if (sd.num_triangles > 0) {
                    B.sets.push_back(sd);
                    entry_set.push_back((int)B.sets.size() - 1);
                } else {
                    entry_set.push_back(-1);   // an empty set: its instances hit nothing
                }
            }
        }

        // entities, breadth first (Program.cs:97-137)
        struct Ent { const JVal* cfg; int parent; M4 m2w; };
        std::vector<Ent> ents;
        std::deque<std::pair<const JVal*, int>> q;
        if (const JVal* sc = root.get("scene")) {
            if (sc->kind != JVal::Arr) bad("\"scene\" is not an array");
            for (const JVal& e : sc->a) q.emplace_back(&e, -1);
        }
        while (!q.empty()) {
            auto [cfg, parent] = q.front();
            q.pop_front();
            int index = (int)ents.size();
            ents.push_back({cfg, parent, M4::identity()});
            if (const JVal* ch = cfg->get("children")) {
                if (ch->kind != JVal::Arr) bad("\"children\" is not an array");
                for (const JVal& c : ch->a) q.emplace_back(&c, index);
            }
        }
        for (Ent& e : ents) {   // ModelToWorld = parent * (S * R * T) (:925-939)
            M4 local = read_transform(e.cfg->get("transform"));
            e.m2w = e.parent < 0 ? mul(M4::identity(), local) : mul(ents[e.parent].m2w, local);
        }
        for (Ent& e : ents) {   // primitives and lights (:942-999)
            const JVal* mat_cfg = e.cfg->get("material");
            const JVal* med_cfg = e.cfg->get("medium");
            const JVal* light_cfg = e.cfg->get("light");
            const JVal* shape = e.cfg->get("shape");
            if (!mat_cfg && !med_cfg) {
                if (!light_cfg) continue;
                if (!light_cfg->get("type")) bad("light: no key type");
                std::string t = read_s(light_cfg, "type", "");
                if (t != "infinite") bad("no infinite light type " + t);
                rd_light_desc L{};
                L.type = RD_LIGHT_INFINITE;
                L.prim = -1;
                L.tex = read_tex(light_cfg, "le", 0, 0, 0);
                std::memcpy(L.to_world, e.m2w.m, sizeof e.m2w.m);
                B.lights.push_back(L);
                continue;
            }
            int mat = mat_cfg ? create_material(mat_cfg) : -1;
            // CreateMediumAdapter (:583-602)
            int med_in = -1, med_out = -1;
            if (med_cfg) {
                if (const JVal* same = med_cfg->get("same")) {
                    med_in = med_out = create_medium(same);
                } else {
                    med_out = create_medium(med_cfg->get("outside"));
                    med_in = create_medium(med_cfg->get("inside"));
                }
            }
            V3f le{0, 0, 0};
            bool emissive = false;
            if (light_cfg) {
                if (!light_cfg->get("type")) bad("light: no key type");
                std::string t = read_s(light_cfg, "type", "");
                if (t != "diffuse_area") bad("no area light type " + t);
                le = read_v3(light_cfg, "le", {1, 1, 1});
                emissive = true;
            }
            auto add_prim = [&](int kind, int index) {
                rd_prim_desc pd{};
                pd.kind = kind; pd.index = index; pd.material = mat; pd.light = -1;
                pd.medium_inside = med_in; pd.medium_outside = med_out;
                int p = (int)B.prims.size();
                if (emissive && kind != RD_PRIM_INSTANCE) {   // InstancedPrimitive has no light
                    rd_light_desc L{};
                    L.type = RD_LIGHT_AREA;
                    L.prim = p;
                    L.le[0] = le.x; L.le[1] = le.y; L.le[2] = le.z;
                    L.tex = -1;
                    std::memcpy(L.to_world, M4::identity().m, sizeof(L.to_world));
                    pd.light = (int)B.lights.size();
                    B.lights.push_back(L);
                }
                B.prims.push_back(pd);
            };
            if (shape) {
                if (!shape->get("type")) bad("invalid shape config");
                std::string t = read_s(shape, "type", "");
                if (t == "sphere") {   // the sphere builder (:49-54)
                    V3f c = xform_point(read_v3(shape, "center", {0, 0, 0}), e.m2w);
                    float r = read_f(shape, "radius", 0.5f);
                    float inv[16];
                    if (!rd::invert4(e.m2w.m, inv)) bad("invalid model matrix");
                    int k = (int)B.spheres.size() / 4;
                    B.spheres.insert(B.spheres.end(), {c.x, c.y, c.z, r});
                    for (int rr = 0; rr < 3; rr++)
                        for (int cc = 0; cc < 3; cc++) B.sphere_w2m.push_back(inv[4 * rr + cc]);
                    if (!e.m2w.is_identity()) B.any_sphere_rot = true;
                    add_prim(RD_PRIM_SPHERE, k);
                } else if (t == "model") {
                    const TriModel& tm = find_model(shape);
                    int t0 = add_mesh(tm, e.m2w);
                    int nt = (int)tm.idx.size() / 3;
                    for (int k = 0; k < nt; k++) add_prim(RD_PRIM_TRIANGLE, t0 + k);
                } else {
                    unsupported("shape type " + t);
                }
            }
            if (const JVal* ins = e.cfg->get("instanced")) {
                if (!ins->get("id")) bad("invalid instanced config");
                int id = read_i(ins, "id", -1);
                if (id < 0 || id >= (int)entry_set.size()) bad("instanced id out of range");
                if (entry_set[id] >= 0) {
                    rd_instance_desc idd{};
                    idd.set = entry_set[id];
                    std::memcpy(idd.to_world, e.m2w.m, sizeof e.m2w.m);
                    B.insts.push_back(idd);
                    add_prim(RD_PRIM_INSTANCE, (int)B.insts.size() - 1);
                }
            }
        }

        // assemble the description
        rd_scene_desc& d = B.desc;
        std::memset(&d, 0, sizeof d);
        d.num_vertices = (int)B.pos.size() / 3;
        d.positions = B.pos.data();
        d.normals = B.any_n ? B.nrm.data() : nullptr;
        d.uvs = B.any_uv ? B.uv.data() : nullptr;
        d.num_triangles = (int)B.tflags.size();
        d.indices = B.idx.data();
        d.tri_flags = B.tflags.data();
        d.num_spheres = (int)B.spheres.size() / 4;
        d.spheres = B.spheres.data();
        d.sphere_w2m = B.any_sphere_rot ? B.sphere_w2m.data() : nullptr;
        d.num_prims = (int)B.prims.size();
        d.prims = B.prims.data();
        if (B.texs.empty()) add_const_tex(0, 0, 0);
        d.num_materials = (int)B.mats.size();
        d.materials = B.mats.data();
        d.num_textures = (int)B.texs.size();
        d.textures = B.texs.data();
        d.num_lights = (int)B.lights.size();
        d.lights = B.lights.data();
        d.num_media = (int)B.media.size();
        d.media = B.media.data();
        d.global_medium = global_medium;
        d.bvh_max_prim = max_prim;
        d.accel = accel;
        d.octree_max_depth = oct_depth;
        d.octree_max_count = oct_count;
        d.octree_out_bound = oct_out;
        if (accel == RD_ACCEL_OCTREE && !B.sets.empty()) {
            d.accel = RD_ACCEL_BVH;   // the device octree holds no instances
            B.notes.push_back("octree accelerator of an instanced scene replaced by the SAH BVH (max_prim 1)");
        }
        d.camera = C;
        d.integrator = I;
        d.num_instance_sets = (int)B.sets.size();
        d.instance_sets = B.sets.data();
        d.num_instances = (int)B.insts.size();
        d.instances = B.insts.data();
        if (params) {
            std::memset(params, 0, sizeof *params);
            params->spp = spp;
            params->particles = particles;
        }
    }
};

}  // namespace

struct rd_json_scene {
    Built built;
};

extern "C" {

int rd_json_load(const char* path, const char* search_dir, rd_json_scene** out, rd_render_params* params) {
    if (!path || !out) { rd::set_last_error("bad argument"); return RD_ERR_INVALID; }
    *out = nullptr;
    std::unique_ptr<rd_json_scene> js(new rd_json_scene());
    try {
        JVal root = parse_json(read_file(path));
        if (root.kind != JVal::Obj) bad("config JSON is not an object");
        std::string dir;
        if (search_dir) {
            dir = search_dir;
        } else {
            std::string p(path);
            size_t s = p.find_last_of("/\\");
            dir = s == std::string::npos ? "." : p.substr(0, s);
        }
        Loader L{root, dir, js->built};
        L.build(params);
    } catch (const LoadError& e) {
        rd::set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        rd::set_last_error(std::string("config: ") + e.what());
        return RD_ERR_INVALID;
    }
    *out = js.release();
    return RD_OK;
}

const rd_scene_desc* rd_json_scene_desc(const rd_json_scene* js) { return js ? &js->built.desc : nullptr; }

void rd_json_scene_free(rd_json_scene* js) { delete js; }

int rd_metal_eta_k(const char* name, float* eta, float* k) {
    const MetalEtaK* m = name ? find_metal(name) : nullptr;
    if (!m || !eta || !k) { rd::set_last_error(std::string("no metal type named ") + (name ? name : "(null)")); return RD_ERR_INVALID; }
    for (int c = 0; c < 3; c++) { eta[c] = m->eta[c]; k[c] = m->k[c]; }
    return RD_OK;
}

rd_scene* rd_scene_load_json(rd_ctx* ctx, const char* path, const char* search_dir, rd_render_params* params) {
    rd_json_scene* js = nullptr;
    if (rd_json_load(path, search_dir, &js, params) != RD_OK) return nullptr;
    rd_scene* s = rd_scene_create(ctx, &js->built.desc);
    rd_json_scene_free(js);
    return s;
}

}  // extern "C"
