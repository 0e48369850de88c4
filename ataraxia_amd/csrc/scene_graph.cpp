// scene_graph.cpp -- the reference's scene model on the host: SceneNode tree
// with TRS transforms, materials, lights, settings, camera; scene.json
// import/export; flattening to the sphere list the renderer consumes.
//
// Restates (paths relative to the reference repository):
//   Engine/include/SceneNode.h:11-66, Engine/src/SceneNode.cpp:42-59   nodes, local/global transform
//   Engine/src/Renderer.cu:67-96                                       traverseSceneGraph (flatten)
//   Engine/src/Utils.cpp:3-187                                         scene.json (de)serialisation
// The JSON reader/writer is this project's own (the reference uses the
// unvendored nlohmann/json).  Output follows that library's dump(4) layout:
// keys in lexicographic order, 4-space indent, one array element per line,
// numbers as the shortest decimal that round-trips the float's double value.
#include <errno.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "host_common.h"

namespace atx {

// ---------------------------------------------------------------------------
// JSON value + parser
// ---------------------------------------------------------------------------
struct Json {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double num = 0.0;
    bool is_int = false;   // written without a fraction or exponent
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    const Json* get(const char* k) const
    {
        if (type != Object) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

struct Parser {
    const char* p;
    const char* end;
    std::string err;

    void ws()
    {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool fail(const char* m)
    {
        if (err.empty()) err = m;
        return false;
    }
    bool lit(const char* s)
    {
        size_t n = strlen(s);
        if ((size_t)(end - p) < n || strncmp(p, s, n) != 0) return false;
        p += n;
        return true;
    }
    bool string(std::string* out)
    {
        if (p >= end || *p != '"') return fail("expected string");
        ++p;
        out->clear();
        while (p < end && *p != '"') {
            char c = *p++;
            if (c == '\\') {
                if (p >= end) return fail("bad escape");
                char e = *p++;
                switch (e) {
                case '"': out->push_back('"'); break;
                case '\\': out->push_back('\\'); break;
                case '/': out->push_back('/'); break;
                case 'b': out->push_back('\b'); break;
                case 'f': out->push_back('\f'); break;
                case 'n': out->push_back('\n'); break;
                case 'r': out->push_back('\r'); break;
                case 't': out->push_back('\t'); break;
                case 'u': {
                    if (end - p < 4) return fail("bad \\u escape");
                    unsigned v = 0;
                    for (int i = 0; i < 4; ++i) {
                        char h = *p++;
                        v <<= 4;
                        if (h >= '0' && h <= '9') v |= (unsigned)(h - '0');
                        else if (h >= 'a' && h <= 'f') v |= (unsigned)(h - 'a' + 10);
                        else if (h >= 'A' && h <= 'F') v |= (unsigned)(h - 'A' + 10);
                        else return fail("bad \\u escape");
                    }
                    if (v < 0x80) {
                        out->push_back((char)v);
                    } else if (v < 0x800) {
                        out->push_back((char)(0xC0 | (v >> 6)));
                        out->push_back((char)(0x80 | (v & 0x3F)));
                    } else {
                        out->push_back((char)(0xE0 | (v >> 12)));
                        out->push_back((char)(0x80 | ((v >> 6) & 0x3F)));
                        out->push_back((char)(0x80 | (v & 0x3F)));
                    }
                    break;
                }
                default: return fail("bad escape");
                }
            } else {
                out->push_back(c);
            }
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        return true;
    }
    bool value(Json* v, int depth)
    {
        if (depth > 200) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end of input");
        const char c = *p;
        if (c == '{') {
            ++p;
            v->type = Json::Object;
            ws();
            if (p < end && *p == '}') { ++p; return true; }
            for (;;) {
                ws();
                std::string k;
                if (!string(&k)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                ++p;
                Json child;
                if (!value(&child, depth + 1)) return false;
                v->obj[k] = std::move(child);
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == '}') { ++p; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++p;
            v->type = Json::Array;
            ws();
            if (p < end && *p == ']') { ++p; return true; }
            for (;;) {
                Json child;
                if (!value(&child, depth + 1)) return false;
                v->arr.push_back(std::move(child));
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == ']') { ++p; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v->type = Json::String;
            return string(&v->str);
        }
        if (lit("true")) { v->type = Json::Bool; v->b = true; return true; }
        if (lit("false")) { v->type = Json::Bool; v->b = false; return true; }
        if (lit("null")) { v->type = Json::Null; return true; }
        if (c == '-' || (c >= '0' && c <= '9')) {
            const char* s = p;
            bool integral = true;
            if (*p == '-') ++p;
            while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' ||
                               *p == '-')) {
                if (*p == '.' || *p == 'e' || *p == 'E') integral = false;
                ++p;
            }
            std::string tok(s, p);
            char* q = nullptr;
            errno = 0;
            const double d = strtod(tok.c_str(), &q);
            if (!q || *q != '\0') return fail("bad number");
            v->type = Json::Number;
            v->num = d;
            v->is_int = integral;
            return true;
        }
        return fail("unexpected character");
    }
};

// ---------------------------------------------------------------------------
// JSON writer (dump with 4-space indent, sorted keys)
// ---------------------------------------------------------------------------
// Shortest decimal digits that round-trip the double, laid out in fixed
// notation for decimal exponents in [-5, 15] and with an exponent otherwise.
static std::string fmt_double(double d)
{
    if (d == 0.0) return std::signbit(d) ? "-0.0" : "0.0";
    if (!std::isfinite(d)) return "null";
    char buf[64];
    int prec = 1;
    for (; prec <= 17; ++prec) {
        snprintf(buf, sizeof(buf), "%.*e", prec - 1, d);
        if (strtod(buf, nullptr) == d) break;
    }
    // buf = [-]d.ddde[+-]XX
    std::string s(buf);
    bool neg = s[0] == '-';
    if (neg) s = s.substr(1);
    const size_t epos = s.find('e');
    const int exp10 = atoi(s.c_str() + epos + 1);
    std::string digits;
    for (size_t i = 0; i < epos; ++i)
        if (s[i] != '.') digits.push_back(s[i]);
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    std::string out;
    if (exp10 >= -5 && exp10 <= 15) {
        if (exp10 >= 0) {
            std::string ip = digits.substr(0, std::min<size_t>(digits.size(), (size_t)exp10 + 1));
            while ((int)ip.size() < exp10 + 1) ip.push_back('0');
            std::string fp = digits.size() > (size_t)exp10 + 1 ? digits.substr((size_t)exp10 + 1) : "0";
            out = ip + "." + fp;
        } else {
            out = "0." + std::string((size_t)(-exp10 - 1), '0') + digits;
        }
    } else {
        out = digits.substr(0, 1);
        if (digits.size() > 1) out += "." + digits.substr(1);
        char e[16];
        snprintf(e, sizeof(e), "e%c%02d", exp10 < 0 ? '-' : '+', exp10 < 0 ? -exp10 : exp10);
        out += e;
    }
    return neg ? "-" + out : out;
}

static void esc(const std::string& s, std::string* out)
{
    out->push_back('"');
    for (char c : s) {
        switch (c) {
        case '"': *out += "\\\""; break;
        case '\\': *out += "\\\\"; break;
        case '\n': *out += "\\n"; break;
        case '\r': *out += "\\r"; break;
        case '\t': *out += "\\t"; break;
        default:
            if ((unsigned char)c < 0x20) {
                char b[8];
                snprintf(b, sizeof(b), "\\u%04x", (unsigned)c);
                *out += b;
            } else {
                out->push_back(c);
            }
        }
    }
    out->push_back('"');
}

static void dump(const Json& v, int indent, std::string* out)
{
    const std::string pad((size_t)indent * 4, ' '), pad2((size_t)(indent + 1) * 4, ' ');
    switch (v.type) {
    case Json::Null: *out += "null"; break;
    case Json::Bool: *out += v.b ? "true" : "false"; break;
    case Json::Number:
        if (v.is_int) {
            char b[32];
            snprintf(b, sizeof(b), "%lld", (long long)v.num);
            *out += b;
        } else {
            *out += fmt_double(v.num);
        }
        break;
    case Json::String: esc(v.str, out); break;
    case Json::Array:
        if (v.arr.empty()) { *out += "[]"; break; }
        *out += "[\n";
        for (size_t i = 0; i < v.arr.size(); ++i) {
            *out += pad2;
            dump(v.arr[i], indent + 1, out);
            *out += i + 1 < v.arr.size() ? ",\n" : "\n";
        }
        *out += pad + "]";
        break;
    case Json::Object: {
        if (v.obj.empty()) { *out += "{}"; break; }
        *out += "{\n";
        size_t i = 0;
        for (auto& kv : v.obj) {
            *out += pad2;
            esc(kv.first, out);
            *out += ": ";
            dump(kv.second, indent + 1, out);
            *out += ++i < v.obj.size() ? ",\n" : "\n";
        }
        *out += pad + "}";
        break;
    }
    }
}

static Json jnum(float f) { Json j; j.type = Json::Number; j.num = (double)f; return j; }
static Json jint(long long i) { Json j; j.type = Json::Number; j.num = (double)i; j.is_int = true; return j; }
static Json jbool(bool b) { Json j; j.type = Json::Bool; j.b = b; return j; }
static Json jstr(const std::string& s) { Json j; j.type = Json::String; j.str = s; return j; }
static Json jarr(std::initializer_list<float> xs)
{
    Json j;
    j.type = Json::Array;
    for (float x : xs) j.arr.push_back(jnum(x));
    return j;
}
static Json jobj() { Json j; j.type = Json::Object; return j; }

// ---------------------------------------------------------------------------
// Scene graph
// ---------------------------------------------------------------------------
struct Node {
    std::string name;
    uint32_t parent = 0;
    float pos[3] = {0, 0, 0};
    float rot[4] = {0, 0, 0, 0};   // x, y, z, w; a value-initialised glm::quat is all zeros
    float scl[3] = {1, 1, 1};
    std::vector<atx_sphere> spheres;
    std::vector<uint32_t> children;
};

struct SceneGraph {
    std::vector<Node> nodes;   // nodes[0] is the root
    std::vector<atx_material> materials;
    std::vector<atx_light> lights;
    atx_settings settings{1, 0, {0, 0}, 15};
    float cam_pos[3] = {0, 0, 0};
    float cam_dir[3] = {0, 0, 0};
    float fov = 45.0f;

    SceneGraph()
    {
        nodes.emplace_back();
        nodes[0].name = "Scene";
    }
};

// ---- transforms (column-major, m[c*4+r]) ----
static void m_identity(float m[16])
{
    memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}

// mat4 * mat4: column c = ((A0*b0 + A1*b1) + A2*b2) + A3*b3
static void m_mul(const float* A, const float* B, float* out)
{
    float r[16];
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row)
            r[c * 4 + row] = ((A[0 * 4 + row] * B[c * 4 + 0] + A[1 * 4 + row] * B[c * 4 + 1]) +
                              A[2 * 4 + row] * B[c * 4 + 2]) + A[3 * 4 + row] * B[c * 4 + 3];
    memcpy(out, r, sizeof(r));
}

// translate(I, p) * mat4_cast(q) * scale(I, s)  (SceneNode.cpp:46)
static void local_transform(const Node& n, float out[16])
{
    float T[16], R[16], S[16], TR[16];
    m_identity(T);
    for (int row = 0; row < 4; ++row) {
        const float i0 = row == 0 ? 1.0f : 0.0f, i1 = row == 1 ? 1.0f : 0.0f;
        const float i2 = row == 2 ? 1.0f : 0.0f, i3 = row == 3 ? 1.0f : 0.0f;
        T[12 + row] = ((i0 * n.pos[0] + i1 * n.pos[1]) + i2 * n.pos[2]) + i3;
    }
    const float qx = n.rot[0], qy = n.rot[1], qz = n.rot[2], qw = n.rot[3];
    const float qxx = qx * qx, qyy = qy * qy, qzz = qz * qz;
    const float qxz = qx * qz, qxy = qx * qy, qyz = qy * qz;
    const float qwx = qw * qx, qwy = qw * qy, qwz = qw * qz;
    m_identity(R);
    R[0] = 1.0f - 2.0f * (qyy + qzz);
    R[1] = 2.0f * (qxy + qwz);
    R[2] = 2.0f * (qxz - qwy);
    R[4] = 2.0f * (qxy - qwz);
    R[5] = 1.0f - 2.0f * (qxx + qzz);
    R[6] = 2.0f * (qyz + qwx);
    R[8] = 2.0f * (qxz + qwy);
    R[9] = 2.0f * (qyz - qwx);
    R[10] = 1.0f - 2.0f * (qxx + qyy);
    m_identity(S);
    S[0] = n.scl[0];
    S[5] = n.scl[1];
    S[10] = n.scl[2];
    m_mul(T, R, TR);
    m_mul(TR, S, out);
}

// Renderer::traverseSceneGraph: pre-order, a node's spheres before its children.
static void flatten(const SceneGraph& g, uint32_t node, const float parent[16], std::vector<atx_sphere>* out)
{
    const Node& n = g.nodes[node];
    float local[16], G[16];
    local_transform(n, local);
    m_mul(parent, local, G);
    const float sx = sqrtf((G[0] * G[0] + G[1] * G[1]) + G[2] * G[2]);
    const float sy = sqrtf((G[4] * G[4] + G[5] * G[5]) + G[6] * G[6]);
    const float sz = sqrtf((G[8] * G[8] + G[9] * G[9]) + G[10] * G[10]);
    const float uniform = (sx + sy + sz) / 3.0f;
    for (const atx_sphere& s : n.spheres) {
        float t[4];
        for (int r = 0; r < 4; ++r) {
            const float m0 = G[0 * 4 + r] * s.center[0], m1 = G[1 * 4 + r] * s.center[1];
            const float m2 = G[2 * 4 + r] * s.center[2], m3 = G[3 * 4 + r] * 1.0f;
            t[r] = (m0 + m1) + (m2 + m3);
        }
        atx_sphere o = s;
        o.center[0] = t[0] / t[3];
        o.center[1] = t[1] / t[3];
        o.center[2] = t[2] / t[3];
        o.radius = s.radius * uniform;
        out->push_back(o);
    }
    for (uint32_t c : n.children) flatten(g, c, G, out);
}

// ---- (de)serialisation (Utils.cpp:3-173) ----
// Numbers are read as doubles and narrowed to the reference's float and int
// fields.  A value outside the target's range is rejected: converting it
// would be undefined behaviour, and a float that overflowed to infinity
// would be written back as null (nlohmann's dump of a non-finite number),
// which no reader accepts.
static bool to_float(double d, float* out, std::string* err, const char* what)
{
    if (!(std::fabs(d) <= (double)FLT_MAX)) {
        *err = std::string("number out of the float range in ") + what;
        return false;
    }
    *out = (float)d;
    return true;
}

static bool to_int32(double d, int32_t* out, std::string* err, const char* what)
{
    if (!(d > -2147483649.0 && d < 2147483648.0)) {
        *err = std::string("number out of the int range in ") + what;
        return false;
    }
    *out = (int32_t)d;   // truncates, as a float-to-int conversion does
    return true;
}

static bool read_vec(const Json* j, int n, float* out, std::string* err, const char* what)
{
    if (!j || j->type != Json::Array || (int)j->arr.size() < n) {
        *err = std::string("missing or short array: ") + what;
        return false;
    }
    for (int i = 0; i < n; ++i) {
        if (j->arr[(size_t)i].type != Json::Number) {
            *err = std::string("non-numeric entry in ") + what;
            return false;
        }
        if (!to_float(j->arr[(size_t)i].num, &out[i], err, what)) return false;
    }
    return true;
}

static bool read_num(const Json* j, double* out, std::string* err, const char* what)
{
    if (!j || (j->type != Json::Number && j->type != Json::Bool)) {
        *err = std::string("missing number: ") + what;
        return false;
    }
    *out = j->type == Json::Bool ? (j->b ? 1.0 : 0.0) : j->num;
    return true;
}

static bool read_node(SceneGraph& g, const Json& j, uint32_t idx, std::string* err, int depth)
{
    if (depth > 1000) { *err = "scene graph too deep"; return false; }
    if (j.type != Json::Object) { *err = "node is not an object"; return false; }
    if (const Json* nm = j.get("name"))
        if (nm->type == Json::String) g.nodes[idx].name = nm->str;
    const Json* tr = j.get("transformation");
    if (!tr) { *err = "node without transformation"; return false; }
    float q[4];
    if (!read_vec(tr->get("position"), 3, g.nodes[idx].pos, err, "transformation.position") ||
        !read_vec(tr->get("rotation"), 4, q, err, "transformation.rotation") ||
        !read_vec(tr->get("scale"), 3, g.nodes[idx].scl, err, "transformation.scale"))
        return false;
    memcpy(g.nodes[idx].rot, q, sizeof(q));   // stored as x, y, z, w (Utils.cpp:72,145)
    if (const Json* sp = j.get("spheres")) {
        if (sp->type != Json::Array) { *err = "spheres is not an array"; return false; }
        for (const Json& s : sp->arr) {
            atx_sphere o{};
            double r = 0, id = 0;
            if (!read_vec(s.get("center"), 3, o.center, err, "sphere.center") ||
                !read_num(s.get("radius"), &r, err, "sphere.radius") ||
                !read_num(s.get("materialIndex"), &id, err, "sphere.materialIndex") ||
                !to_float(r, &o.radius, err, "sphere.radius") ||
                !to_int32(id, &o.material, err, "sphere.materialIndex"))
                return false;
            g.nodes[idx].spheres.push_back(o);
        }
    }
    if (const Json* ch = j.get("children")) {
        if (ch->type != Json::Array) { *err = "children is not an array"; return false; }
        for (const Json& c : ch->arr) {
            const uint32_t k = (uint32_t)g.nodes.size();
            g.nodes.emplace_back();
            g.nodes[idx].children.push_back(k);
            g.nodes[k].name = "Untitled";
            g.nodes[k].parent = idx;
            if (!read_node(g, c, k, err, depth + 1)) return false;
        }
    }
    return true;
}

static bool from_json(const Json& j, SceneGraph* g, std::string* err)
{
    if (j.type != Json::Object) { *err = "top level is not an object"; return false; }
    const Json* cam = j.get("camera");
    double fov = 45.0;
    if (!cam || !read_vec(cam->get("position"), 3, g->cam_pos, err, "camera.position") ||
        !read_vec(cam->get("direction"), 3, g->cam_dir, err, "camera.direction") ||
        !read_num(cam->get("fov"), &fov, err, "camera.fov") || !to_float(fov, &g->fov, err, "camera.fov")) {
        if (err->empty()) *err = "missing camera";
        return false;
    }
    if (const Json* sg = j.get("sceneGraph"))
        if (!read_node(*g, *sg, 0, err, 0)) return false;
    if (const Json* ms = j.get("materials")) {
        if (ms->type != Json::Array) { *err = "materials is not an array"; return false; }
        for (const Json& m : ms->arr) {
            atx_material o{};
            o.albedo[0] = o.albedo[1] = o.albedo[2] = 1.0f;
            o.F0[0] = o.F0[1] = o.F0[2] = 0.04f;
            double rough = 0, metal = 0, ei = 0;
            if (!read_vec(m.get("albedo"), 3, o.albedo, err, "material.albedo") ||
                !read_num(m.get("roughness"), &rough, err, "material.roughness") ||
                !read_num(m.get("metallic"), &metal, err, "material.metallic") ||
                !read_vec(m.get("F0"), 3, o.F0, err, "material.F0") ||
                !read_num(m.get("emissionIntensity"), &ei, err, "material.emissionIntensity") ||
                !read_vec(m.get("emissionColor"), 3, o.emission_color, err, "material.emissionColor") ||
                !to_float(rough, &o.roughness, err, "material.roughness") ||
                !to_float(metal, &o.metallic, err, "material.metallic") ||
                !to_float(ei, &o.emission_intensity, err, "material.emissionIntensity"))
                return false;
            o.id = 0;
            g->materials.push_back(o);
        }
    }
    if (const Json* ls = j.get("lights")) {
        if (ls->type != Json::Array) { *err = "lights is not an array"; return false; }
        for (const Json& l : ls->arr) {
            atx_light o{};
            double in = 0;
            if (!read_vec(l.get("position"), 3, o.position, err, "light.position") ||
                !read_num(l.get("intensity"), &in, err, "light.intensity") ||
                !read_vec(l.get("color"), 3, o.color, err, "light.color") ||
                !to_float(in, &o.intensity, err, "light.intensity"))
                return false;
            g->lights.push_back(o);
        }
    }
    const Json* st = j.get("settings");
    double mb = 15, sky = 0, acc = 1;
    if (!st || !read_num(st->get("maxBounces"), &mb, err, "settings.maxBounces") ||
        !read_num(st->get("skyLight"), &sky, err, "settings.skyLight") ||
        !read_num(st->get("accumulation"), &acc, err, "settings.accumulation")) {
        if (err->empty()) *err = "missing settings";
        return false;
    }
    if (!to_int32(mb, &g->settings.max_bounces, err, "settings.maxBounces")) return false;
    g->settings.sky_light = sky != 0.0;
    g->settings.accumulation = acc != 0.0;
    return true;
}

static Json node_json(const SceneGraph& g, uint32_t idx)
{
    const Node& n = g.nodes[idx];
    Json j = jobj();
    j.obj["name"] = jstr(n.name);
    Json t = jobj();
    t.obj["position"] = jarr({n.pos[0], n.pos[1], n.pos[2]});
    t.obj["rotation"] = jarr({n.rot[0], n.rot[1], n.rot[2], n.rot[3]});
    t.obj["scale"] = jarr({n.scl[0], n.scl[1], n.scl[2]});
    j.obj["transformation"] = t;
    if (!n.spheres.empty()) {
        Json sp;
        sp.type = Json::Array;
        for (const atx_sphere& s : n.spheres) {
            Json o = jobj();
            o.obj["center"] = jarr({s.center[0], s.center[1], s.center[2]});
            o.obj["radius"] = jnum(s.radius);
            o.obj["materialIndex"] = jint(s.material);
            sp.arr.push_back(o);
        }
        j.obj["spheres"] = sp;
    }
    if (!n.children.empty()) {
        Json ch;
        ch.type = Json::Array;
        for (uint32_t c : n.children) ch.arr.push_back(node_json(g, c));
        j.obj["children"] = ch;
    }
    return j;
}

static Json to_json(const SceneGraph& g)
{
    Json j = jobj();
    Json cam = jobj();
    cam.obj["position"] = jarr({g.cam_pos[0], g.cam_pos[1], g.cam_pos[2]});
    cam.obj["direction"] = jarr({g.cam_dir[0], g.cam_dir[1], g.cam_dir[2]});
    cam.obj["fov"] = jnum(g.fov);
    j.obj["camera"] = cam;
    j.obj["sceneGraph"] = node_json(g, 0);
    if (!g.materials.empty()) {
        Json ms;
        ms.type = Json::Array;
        for (const atx_material& m : g.materials) {
            Json o = jobj();
            o.obj["albedo"] = jarr({m.albedo[0], m.albedo[1], m.albedo[2]});
            o.obj["roughness"] = jnum(m.roughness);
            o.obj["metallic"] = jnum(m.metallic);
            o.obj["F0"] = jarr({m.F0[0], m.F0[1], m.F0[2]});
            o.obj["emissionIntensity"] = jnum(m.emission_intensity);
            o.obj["emissionColor"] = jarr({m.emission_color[0], m.emission_color[1], m.emission_color[2]});
            ms.arr.push_back(o);
        }
        j.obj["materials"] = ms;
    }
    if (!g.lights.empty()) {
        Json ls;
        ls.type = Json::Array;
        for (const atx_light& l : g.lights) {
            Json o = jobj();
            o.obj["position"] = jarr({l.position[0], l.position[1], l.position[2]});
            o.obj["intensity"] = jnum(l.intensity);
            o.obj["color"] = jarr({l.color[0], l.color[1], l.color[2]});
            ls.arr.push_back(o);
        }
        j.obj["lights"] = ls;
    }
    Json st = jobj();
    st.obj["maxBounces"] = jint(g.settings.max_bounces);
    st.obj["skyLight"] = jbool(g.settings.sky_light != 0);
    st.obj["accumulation"] = jbool(g.settings.accumulation != 0);
    j.obj["settings"] = st;
    return j;
}

static SceneGraph* cast(atx_scene_graph* h) { return reinterpret_cast<SceneGraph*>(h); }
static const SceneGraph* cast(const atx_scene_graph* h) { return reinterpret_cast<const SceneGraph*>(h); }

static int parse_into(const char* text, size_t len, SceneGraph* g)
{
    Parser ps{text, text + len, {}};
    Json j;
    if (!ps.value(&j, 0)) {
        set_error("scene.json parse error at byte %zu: %s", (size_t)(ps.p - text), ps.err.c_str());
        return ATX_ERR_IO;
    }
    ps.ws();
    if (ps.p != ps.end) {
        set_error("scene.json: trailing characters at byte %zu", (size_t)(ps.p - text));
        return ATX_ERR_IO;
    }
    std::string err;
    if (!from_json(j, g, &err)) {
        set_error("scene.json: %s", err.c_str());
        return ATX_ERR_IO;
    }
    return ATX_OK;
}

}  // namespace atx

using namespace atx;

extern "C" {

int atx_scene_create(atx_scene_graph** out)
{
    ATX_CHECK_ARG(out, "out is NULL");
    SceneGraph* g = new (std::nothrow) SceneGraph();
    if (!g) { set_error("out of host memory"); return ATX_ERR_NOMEM; }
    *out = reinterpret_cast<atx_scene_graph*>(g);
    return ATX_OK;
}

int atx_scene_destroy(atx_scene_graph* g)
{
    delete cast(g);
    return ATX_OK;
}

int atx_scene_parse_json(const char* text, uint64_t len, atx_scene_graph** out)
{
    ATX_CHECK_ARG(text && out, "NULL argument");
    *out = nullptr;
    std::unique_ptr<SceneGraph> g(new (std::nothrow) SceneGraph());
    if (!g) { set_error("out of host memory"); return ATX_ERR_NOMEM; }
    int rc = parse_into(text, (size_t)len, g.get());
    if (rc != ATX_OK) return rc;
    *out = reinterpret_cast<atx_scene_graph*>(g.release());
    return ATX_OK;
}

int atx_scene_load_json(const char* path, atx_scene_graph** out)
{
    ATX_CHECK_ARG(path && out, "NULL argument");
    *out = nullptr;
    FILE* f = fopen(path, "rb");
    if (!f)   // Utils::importScene returns a default Scene when the file cannot be opened
        return atx_scene_create(out);
    std::string text;
    char buf[65536];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    fclose(f);
    return atx_scene_parse_json(text.data(), text.size(), out);
}

int atx_scene_to_json(const atx_scene_graph* h, char* buf, uint64_t cap, uint64_t* len)
{
    const SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && len, "NULL argument");
    std::string s;
    dump(to_json(*g), 0, &s);
    *len = s.size();
    if (buf) {
        ATX_CHECK_ARG(cap >= s.size() + 1, "buffer holds %llu bytes, need %zu", (unsigned long long)cap,
                      s.size() + 1);
        memcpy(buf, s.c_str(), s.size() + 1);
    }
    return ATX_OK;
}

int atx_scene_save_json(const atx_scene_graph* h, const char* path)
{
    const SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && path, "NULL argument");
    std::string s;
    dump(to_json(*g), 0, &s);
    FILE* f = fopen(path, "wb");
    if (!f) { set_error("cannot open %s for writing", path); return ATX_ERR_IO; }
    const bool ok = fwrite(s.data(), 1, s.size(), f) == s.size();
    if (fclose(f) != 0 || !ok) { set_error("write to %s failed", path); return ATX_ERR_IO; }
    return ATX_OK;
}

int atx_scene_add_node(atx_scene_graph* h, uint32_t parent, const char* name, const float* position,
                       const float* rotation_xyzw, const float* scale, uint32_t* node)
{
    SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g, "scene is NULL");
    ATX_CHECK_ARG(parent < g->nodes.size(), "unknown parent node %u", parent);
    Node n;
    n.name = name ? name : "Untitled";
    n.parent = parent;
    if (position) memcpy(n.pos, position, sizeof(n.pos));
    if (rotation_xyzw) memcpy(n.rot, rotation_xyzw, sizeof(n.rot));
    if (scale) memcpy(n.scl, scale, sizeof(n.scl));
    const uint32_t k = (uint32_t)g->nodes.size();
    g->nodes.push_back(n);
    g->nodes[parent].children.push_back(k);
    if (node) *node = k;
    return ATX_OK;
}

int atx_scene_set_node_transform(atx_scene_graph* h, uint32_t node, const float* position,
                                 const float* rotation_xyzw, const float* scale)
{
    SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && node < g->nodes.size(), "unknown node");
    if (position) memcpy(g->nodes[node].pos, position, 3 * sizeof(float));
    if (rotation_xyzw) memcpy(g->nodes[node].rot, rotation_xyzw, 4 * sizeof(float));
    if (scale) memcpy(g->nodes[node].scl, scale, 3 * sizeof(float));
    return ATX_OK;
}

int atx_scene_add_sphere(atx_scene_graph* h, uint32_t node, const atx_sphere* s)
{
    SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && s, "NULL argument");
    ATX_CHECK_ARG(node < g->nodes.size(), "unknown node %u", node);
    g->nodes[node].spheres.push_back(*s);
    return ATX_OK;
}

int atx_scene_add_material(atx_scene_graph* h, const atx_material* m)
{
    SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && m, "NULL argument");
    g->materials.push_back(*m);
    return ATX_OK;
}

int atx_scene_add_light(atx_scene_graph* h, const atx_light* l)
{
    SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && l, "NULL argument");
    g->lights.push_back(*l);
    return ATX_OK;
}

int atx_scene_set_settings(atx_scene_graph* h, const atx_settings* s)
{
    SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && s, "NULL argument");
    g->settings = *s;
    return ATX_OK;
}

int atx_scene_get_settings(const atx_scene_graph* h, atx_settings* s)
{
    const SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && s, "NULL argument");
    *s = g->settings;
    return ATX_OK;
}

int atx_scene_set_camera(atx_scene_graph* h, const float* position, const float* direction, float fov_deg)
{
    SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && position && direction, "NULL argument");
    memcpy(g->cam_pos, position, sizeof(g->cam_pos));
    memcpy(g->cam_dir, direction, sizeof(g->cam_dir));
    g->fov = fov_deg;
    return ATX_OK;
}

int atx_scene_get_camera(const atx_scene_graph* h, float* position, float* direction, float* fov_deg)
{
    const SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && position && direction && fov_deg, "NULL argument");
    memcpy(position, g->cam_pos, sizeof(g->cam_pos));
    memcpy(direction, g->cam_dir, sizeof(g->cam_dir));
    *fov_deg = g->fov;
    return ATX_OK;
}

int atx_scene_counts(const atx_scene_graph* h, uint64_t* nodes, uint64_t* spheres, uint64_t* materials,
                     uint64_t* lights)
{
    const SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g, "scene is NULL");
    uint64_t ns = 0;
    for (const Node& n : g->nodes) ns += n.spheres.size();
    if (nodes) *nodes = g->nodes.size();
    if (spheres) *spheres = ns;
    if (materials) *materials = g->materials.size();
    if (lights) *lights = g->lights.size();
    return ATX_OK;
}

int atx_scene_get_node(const atx_scene_graph* h, uint32_t node, char* name, uint64_t name_cap, float* position,
                       float* rotation_xyzw, float* scale, uint32_t* parent, uint64_t* num_spheres)
{
    const SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && node < g->nodes.size(), "unknown node");
    const Node& n = g->nodes[node];
    if (name && name_cap > 0) {
        const size_t k = std::min<size_t>(n.name.size(), (size_t)name_cap - 1);
        memcpy(name, n.name.data(), k);
        name[k] = '\0';
    }
    if (position) memcpy(position, n.pos, sizeof(n.pos));
    if (rotation_xyzw) memcpy(rotation_xyzw, n.rot, sizeof(n.rot));
    if (scale) memcpy(scale, n.scl, sizeof(n.scl));
    if (parent) *parent = n.parent;
    if (num_spheres) *num_spheres = n.spheres.size();
    return ATX_OK;
}

int atx_scene_get_node_spheres(const atx_scene_graph* h, uint32_t node, atx_sphere* out, uint64_t cap,
                               uint64_t* count)
{
    const SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && count && node < g->nodes.size(), "bad argument");
    const Node& n = g->nodes[node];
    *count = n.spheres.size();
    if (out) {
        ATX_CHECK_ARG(cap >= n.spheres.size(), "buffer too small");
        if (!n.spheres.empty()) memcpy(out, n.spheres.data(), n.spheres.size() * sizeof(atx_sphere));
    }
    return ATX_OK;
}

int atx_scene_flatten(const atx_scene_graph* h, atx_sphere* out, uint64_t cap, uint64_t* count)
{
    const SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && count, "NULL argument");
    std::vector<atx_sphere> v;
    float I[16];
    m_identity(I);
    flatten(*g, 0, I, &v);
    *count = v.size();
    if (out) {
        ATX_CHECK_ARG(cap >= v.size(), "buffer holds %llu spheres, need %zu", (unsigned long long)cap, v.size());
        if (!v.empty()) memcpy(out, v.data(), v.size() * sizeof(atx_sphere));
    }
    return ATX_OK;
}

int atx_scene_get_materials(const atx_scene_graph* h, atx_material* out, uint64_t cap, uint64_t* count)
{
    const SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && count, "NULL argument");
    *count = g->materials.size();
    if (out) {
        ATX_CHECK_ARG(cap >= g->materials.size(), "buffer too small");
        if (!g->materials.empty()) memcpy(out, g->materials.data(), g->materials.size() * sizeof(atx_material));
    }
    return ATX_OK;
}

int atx_scene_get_lights(const atx_scene_graph* h, atx_light* out, uint64_t cap, uint64_t* count)
{
    const SceneGraph* g = cast(h);
    ATX_CHECK_ARG(g && count, "NULL argument");
    *count = g->lights.size();
    if (out) {
        ATX_CHECK_ARG(cap >= g->lights.size(), "buffer too small");
        if (!g->lights.empty()) memcpy(out, g->lights.data(), g->lights.size() * sizeof(atx_light));
    }
    return ATX_OK;
}

}  // extern "C"
