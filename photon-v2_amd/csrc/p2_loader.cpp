// .p2 scene ingestion (phg_scene_parse_p2 / phg_scene_load_p2): the subset of Photon's scene
// description language that a BNEEPT render on this path consumes, flattened into a
// phg_scene_host plus the render settings.
//
// Syntax (SDL 1.x; the reference's reader is SDL/SdlCommandParser.cpp:100-700 and
// SDL/SdlInlinePacketInterface.cpp:40-140): commands end with ';', '//' starts a comment,
//   category(type) @name = [type name value] ... ;            creates a resource
//   category(type).executor(@name) = [...] ;  executor(@name) = [...] ;   runs an executor
//   #version 1.1.0;                                            directive (ignored)
// A clause value is a bare token, a "quoted string", an @reference or a {braced block}.
//
// Resources understood (field names as the reference's SDL classes declare them):
//   option(single-frame-render-session)  frame-size, visualizer, observer, sample-source
//     (EngineEnv/Session/SingleFrameRenderSession.h)
//   observer(single-lens)  pos, dir, up-axis, fov-degrees, sensor-width-mm, sensor-offset-mm
//     (EngineEnv/Observer/OrientedRasterObserver.h, SingleLensObserver.h, .cpp:68-84)
//   sample-source(stratified | halton | uniform-random)  samples
//   visualizer(path-tracing)  estimator, sample-filter (EngineEnv/Visualizer/PathTracingVisualizer.h),
//     plus max-path-length (this path's bounce cap; the reference has none)
//   geometry(rectangle | cuboid | sphere | triangle | triangle-mesh)
//     (Actor/Geometry/GRectangle.cpp:67-101, GCuboid.h, GSphere.h, GTriangle.h, GTriangleMesh.h)
//   image(constant)  values, color-space
//   material(matte-opaque)  albedo                      -> Lambertian (Actor/Material/MatteOpaque.h)
//   material(abraded-opaque)  roughness, f0, fresnel, roughness-to-alpha, ior-inner-n/k
//                                                       -> GGX conductor (AbradedOpaque.h)
//   material(abraded-translucent)  roughness, ior-outer, ior-inner -> GGX dielectric
//   actor(model)  geometry, material;  actor(rectangle-light)  color, watts, width, height
//   executors translate / rotate / scale on actors (Actor/PhysicalActor.cpp:6-38): the
//     transform is T * R * S, rotations compose as new * old
//     (Math/Transform/TDecomposedTransform.h:40-80, 158-170).
// Spheres are tessellated (icosphere, 20480 triangles) because this path intersects
// triangles only.  Anything else is skipped and counted in phg_p2_settings.warnings.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/phg.h"
#include "color_energy.h"
#include "scene_buf.h"

namespace {

using namespace phg::sb;

struct Clause {
    std::string type, name, value;   // value: quotes and braces removed, '@' kept for references
};

struct Resource {
    std::string category, type, name;
    std::vector<Clause> clauses;
    // actors: decomposed transform
    F3 pos{0, 0, 0}, scale{1, 1, 1};
    double rot[4] = {1, 0, 0, 0};    // quaternion w, x, y, z
    const Clause* find(const std::string& n) const {
        for (const Clause& c : clauses)
            if (c.name == n) return &c;
        return nullptr;
    }
};

struct ParseError {
    std::string msg;
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// Remove // comments (outside quotes) and split into commands at ';' (outside quotes/braces).
std::vector<std::string> split_commands(const std::string& text) {
    std::vector<std::string> out;
    std::string cur;
    bool in_quote = false;
    int brace = 0;
    for (size_t i = 0; i < text.size(); ++i) {
        const char ch = text[i];
        if (!in_quote && brace == 0 && ch == '/' && i + 1 < text.size() && text[i + 1] == '/') {
            while (i < text.size() && text[i] != '\n') ++i;
            cur.push_back('\n');
            continue;
        }
        if (ch == '"' && brace == 0) in_quote = !in_quote;
        if (!in_quote) {
            if (ch == '{') ++brace;
            if (ch == '}') brace = std::max(0, brace - 1);
            if (ch == ';' && brace == 0) {
                const std::string t = trim(cur);
                if (!t.empty()) out.push_back(t);
                cur.clear();
                continue;
            }
        }
        cur.push_back(ch);
    }
    const std::string t = trim(cur);
    if (!t.empty()) out.push_back(t);
    return out;
}

// Position of the first '=' outside quotes.
size_t find_assign(const std::string& s) {
    bool q = false;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '"') q = !q;
        if (!q && s[i] == '=') return i;
    }
    return std::string::npos;
}

// Clauses: [type name value]; value may be "quoted", {braced} or a bare token.
std::vector<Clause> parse_clauses(const std::string& data) {
    std::vector<Clause> out;
    size_t i = 0;
    while (i < data.size()) {
        while (i < data.size() && data[i] != '[') {
            if (!std::isspace((unsigned char)data[i])) throw ParseError{"unexpected text outside a clause"};
            ++i;
        }
        if (i >= data.size()) break;
        size_t j = i + 1;
        bool q = false;
        int br = 0;
        for (; j < data.size(); ++j) {
            const char ch = data[j];
            if (ch == '"' && br == 0) q = !q;
            else if (!q && ch == '{') ++br;
            else if (!q && ch == '}') --br;
            else if (!q && br == 0 && ch == ']') break;
        }
        if (j >= data.size()) throw ParseError{"unterminated clause"};
        const std::string body = data.substr(i + 1, j - i - 1);
        i = j + 1;
        // tokens: type, name, value
        size_t k = 0;
        auto next_token = [&](bool rest) {
            while (k < body.size() && std::isspace((unsigned char)body[k])) ++k;
            if (k >= body.size()) return std::string();
            if (rest) {
                std::string v = trim(body.substr(k));
                k = body.size();
                if (v.size() >= 2 && ((v.front() == '"' && v.back() == '"') || (v.front() == '{' && v.back() == '}')))
                    v = v.substr(1, v.size() - 2);
                return v;
            }
            size_t e = k;
            while (e < body.size() && !std::isspace((unsigned char)body[e])) ++e;
            std::string t = body.substr(k, e - k);
            k = e;
            return t;
        };
        Clause c;
        c.type = next_token(false);
        c.name = next_token(false);
        c.value = next_token(true);
        if (c.type.empty() || c.name.empty()) throw ParseError{"clause needs a type and a name"};
        out.push_back(std::move(c));
    }
    return out;
}

std::vector<double> numbers(const std::string& v) {
    std::vector<double> out;
    std::string s = v;
    for (char& ch : s)
        if (ch == ',' || ch == '"' || ch == '{' || ch == '}' || ch == '(' || ch == ')') ch = ' ';
    std::istringstream is(s);
    double x;
    while (is >> x) out.push_back(x);
    return out;
}

std::string ref_name(const std::string& v) {
    std::string t = trim(v);
    if (!t.empty() && t[0] == '@') t = trim(t.substr(1));
    if (t.size() >= 2 && t.front() == '"' && t.back() == '"') t = t.substr(1, t.size() - 2);
    return t;
}

struct Header {
    enum Kind { Load, Exec, Directive, Other } kind = Other;
    std::string category, type, name, executor;
};

// "category(type) @name", "category(type).exec(@name)", "exec(@name)", "#..."
Header parse_header(const std::string& h) {
    Header out;
    const std::string s = trim(h);
    if (s.empty()) return out;
    if (s[0] == '#') { out.kind = Header::Directive; return out; }
    auto ident = [&](size_t& i) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
        size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '-' || s[i] == '_')) ++i;
        return s.substr(b, i - b);
    };
    auto expect = [&](size_t& i, char ch) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
        if (i >= s.size() || s[i] != ch) return false;
        ++i;
        return true;
    };
    size_t i = 0;
    const std::string a = ident(i);
    if (a.empty() || !expect(i, '(')) return out;
    size_t save = i;
    while (save < s.size() && std::isspace((unsigned char)s[save])) ++save;
    if (save < s.size() && s[save] == '@') {
        // exec(@name)
        const size_t close = s.rfind(')');
        if (close == std::string::npos || close < save) return out;
        out.kind = Header::Exec;
        out.executor = a;
        out.name = ref_name(s.substr(save, close - save));
        return out;
    }
    const std::string b = ident(i);
    if (b.empty() || !expect(i, ')')) return out;
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    if (i < s.size() && s[i] == '@') {
        out.kind = Header::Load;
        out.category = a; out.type = b;
        out.name = ref_name(s.substr(i));
        return out;
    }
    if (i < s.size() && s[i] == '.') {
        ++i;
        const std::string ex = ident(i);
        if (ex.empty() || !expect(i, '(')) return out;
        const size_t close = s.rfind(')');
        if (close == std::string::npos || close < i) return out;
        out.kind = Header::Exec;
        out.category = a; out.type = b; out.executor = ex;
        out.name = ref_name(s.substr(i, close - i));
        return out;
    }
    return out;
}

// quaternion helpers (w, x, y, z), double precision
void qmul(const double* a, const double* b, double* o) {
    const double w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    const double x = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    const double y = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    const double z = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    o[0] = w; o[1] = x; o[2] = y; o[3] = z;
}
void qnorm(double* q) {
    const double l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int k = 0; k < 4; ++k) q[k] /= l;
}
F3 qrot(const double* q, F3 v) {
    const double p[4] = {0, v.x, v.y, v.z};
    const double qc[4] = {q[0], -q[1], -q[2], -q[3]};
    double t[4], r[4];
    qmul(q, p, t);
    qmul(t, qc, r);
    return f3((float)r[1], (float)r[2], (float)r[3]);
}

float srgb_to_linear(float c) {
    return c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f);
}

struct Loader {
    std::map<std::string, Resource> res;
    std::vector<std::string> actor_order;
    phg_p2_settings st{};
    std::string session;

    void warn() { ++st.warnings; }

    void run(const std::string& text) {
        for (const std::string& cmd : split_commands(text)) {
            if (cmd[0] == '#') continue;
            const size_t eq = find_assign(cmd);
            if (eq == std::string::npos) { warn(); continue; }
            const Header h = parse_header(cmd.substr(0, eq));
            if (h.kind == Header::Directive) continue;
            if (h.kind == Header::Other) { warn(); continue; }
            const std::vector<Clause> clauses = parse_clauses(cmd.substr(eq + 1));
            if (h.kind == Header::Load) {
                Resource r;
                r.category = h.category; r.type = h.type; r.name = h.name; r.clauses = clauses;
                if (r.category == "actor") actor_order.push_back(r.name);
                if (r.category == "option" && r.type == "single-frame-render-session") session = r.name;
                res[r.name] = std::move(r);
            } else {
                auto it = res.find(h.name);
                if (it == res.end() || it->second.category != "actor") { warn(); continue; }
                execute(it->second, h.executor, clauses);
            }
        }
    }

    static const Clause* find(const std::vector<Clause>& cs, const char* n) {
        for (const Clause& c : cs)
            if (c.name == n) return &c;
        return nullptr;
    }

    void execute(Resource& a, const std::string& ex, const std::vector<Clause>& cs) {
        if (ex == "translate") {
            const Clause* c = find(cs, "amount");
            const std::vector<double> v = c ? numbers(c->value) : std::vector<double>();
            if (v.size() != 3) throw ParseError{"translate needs a vector3 amount"};
            a.pos = add(a.pos, f3((float)v[0], (float)v[1], (float)v[2]));
        } else if (ex == "rotate") {
            const Clause* ax = find(cs, "axis");
            const Clause* dg = find(cs, "degrees");
            const std::vector<double> v = ax ? numbers(ax->value) : std::vector<double>();
            const std::vector<double> d = dg ? numbers(dg->value) : std::vector<double>();
            if (v.size() != 3 || d.size() != 1) throw ParseError{"rotate needs a vector3 axis and real degrees"};
            const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            if (!(l > 0)) throw ParseError{"rotate axis is zero"};
            const double h = d[0] * 3.14159265358979323846 / 360.0;
            double q[4] = {std::cos(h), std::sin(h) * v[0] / l, std::sin(h) * v[1] / l, std::sin(h) * v[2] / l};
            qnorm(q);
            double t[4];
            qmul(q, a.rot, t);
            qnorm(t);
            std::memcpy(a.rot, t, sizeof(t));
        } else if (ex == "scale") {
            const Clause* c = find(cs, "amount");
            const std::vector<double> v = c ? numbers(c->value) : std::vector<double>();
            if (v.size() == 1) a.scale = mul(a.scale, (float)v[0]);
            else if (v.size() == 3) a.scale = f3(a.scale.x * (float)v[0], a.scale.y * (float)v[1], a.scale.z * (float)v[2]);
            else throw ParseError{"scale needs a real or vector3 amount"};
        } else {
            warn();
        }
    }

    const Resource* get(const std::string& ref) const {
        auto it = res.find(ref_name(ref));
        return it == res.end() ? nullptr : &it->second;
    }

    static double real_or(const Resource& r, const char* n, double dflt) {
        const Clause* c = r.find(n);
        if (!c) return dflt;
        const std::vector<double> v = numbers(c->value);
        return v.empty() ? dflt : v[0];
    }
    static bool vec3(const Resource& r, const char* n, F3* out) {
        const Clause* c = r.find(n);
        if (!c) return false;
        const std::vector<double> v = numbers(c->value);
        if (v.size() == 1) { *out = f3((float)v[0], (float)v[0], (float)v[0]); return true; }
        if (v.size() != 3) throw ParseError{std::string("field ") + n + " needs 3 numbers"};
        *out = f3((float)v[0], (float)v[1], (float)v[2]);
        return true;
    }

    // linear RGB of an image reference, an inline value list or a spectrum clause
    F3 color_of(const Clause* c, F3 dflt) {
        if (!c) return dflt;
        const std::string v = trim(c->value);
        if (!v.empty() && v[0] == '@') {
            const Resource* img = get(v);
            if (!img || img->category != "image") throw ParseError{"albedo references no image"};
            if (img->type != "constant") { warn(); return dflt; }
            const Clause* vals = img->find("values");
            F3 rgb = dflt;
            if (vals) {
                const std::vector<double> n = numbers(vals->value);
                if (n.size() == 1) rgb = f3((float)n[0], (float)n[0], (float)n[0]);
                else if (n.size() >= 3) rgb = f3((float)n[0], (float)n[1], (float)n[2]);
            }
            const Clause* cs = img->find("color-space");
            if (cs && (cs->value == "SRGB" || cs->value == "sRGB"))
                rgb = f3(srgb_to_linear(rgb.x), srgb_to_linear(rgb.y), srgb_to_linear(rgb.z));
            else if (cs && cs->value != "LSRGB" && cs->value != "Linear-sRGB") warn();
            return rgb;
        }
        const std::vector<double> n = numbers(v);
        if (n.size() == 1) return f3((float)n[0], (float)n[0], (float)n[0]);
        if (n.size() >= 3) return f3((float)n[0], (float)n[1], (float)n[2]);
        warn();
        return dflt;
    }

    uint32_t material_of(SceneBuf& s, std::map<std::string, uint32_t>& cache, const Clause* mref) {
        const std::string key = mref ? ref_name(mref->value) : std::string();
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
        uint32_t id;
        const Resource* m = mref ? get(mref->value) : nullptr;
        if (!m) {
            id = s.add_material(PHG_MAT_LAMBERT, f3(0.5f, 0.5f, 0.5f), 0, 1, 1);  // default matte
        } else if (m->type == "matte-opaque") {
            if (m->find("sigma-degrees")) warn();  // Oren-Nayar roughness is not on this path
            id = s.add_material(PHG_MAT_LAMBERT, color_of(m->find("albedo"), f3(0.5f, 0.5f, 0.5f)), 0, 1, 1);
        } else if (m->type == "abraded-opaque" || m->type == "abraded-translucent") {
            const Clause* ms = m->find("microsurface");
            if (ms && ms->value != "ggx" && ms->value != "trowbridge-reitz") warn();
            double r = real_or(*m, "roughness", 0.5);
            if (m->find("roughness-v")) warn();  // anisotropic roughness: isotropic u used
            const Clause* map = m->find("roughness-to-alpha");
            // the device squares roughness (the reference's default mapping); express others
            // as the roughness whose square is their alpha
            if (map && map->value == "equaled") r = std::sqrt(std::max(r, 0.0));
            else if (map && map->value == "pbrt-v3") {
                const double x = std::log(std::max(r, 0.001));
                const double alpha = 1.62142 + 0.819955 * x + 0.1734 * x * x + 0.0171201 * x * x * x +
                                     0.000640711 * x * x * x * x;
                r = std::sqrt(std::max(alpha, 0.0));
            } else if (map && map->value != "squared") warn();
            if (m->type == "abraded-opaque") {
                // ConductiveInterfaceInfo::genFresnelEffect (ConductiveInterfaceInfo.cpp): with
                // the complex inner index given, `fresnel exact` selects ExactConductorFresnel
                // and `schlick` (the default) Schlick with the F0 of that index; without it,
                // Schlick with f0 (an `exact` request then falls back, with a warning)
                F3 f0 = f3(1.0f, 1.0f, 1.0f);
                const Clause* fr = m->find("fresnel");
                const bool exact = fr && fr->value == "exact";
                if (fr && fr->value != "exact" && fr->value != "schlick") warn();
                if (m->find("ior-inner-n") && m->find("ior-inner-k")) {
                    F3 n{}, k{};
                    vec3(*m, "ior-inner-n", &n);
                    vec3(*m, "ior-inner-k", &k);
                    const float eo = (float)real_or(*m, "ior-outer", 1.0);
                    if (exact) {
                        id = s.add_exact_conductor(n, k, (float)r, eo);
                    } else {
                        // SchlickApproxConductorFresnel(iorOuter, n, k): normal-incidence reflectance
                        auto F0 = [&](float nn, float kk) {
                            return ((nn - eo) * (nn - eo) + kk * kk) / ((nn + eo) * (nn + eo) + kk * kk);
                        };
                        id = s.add_material(PHG_MAT_CONDUCTOR, f3(F0(n.x, k.x), F0(n.y, k.y), F0(n.z, k.z)), (float)r, 1,
                                            1);
                    }
                } else {
                    if (m->find("f0")) f0 = color_of(m->find("f0"), f0);
                    if (exact) warn();
                    id = s.add_material(PHG_MAT_CONDUCTOR, f0, (float)r, 1, 1);
                }
            } else {
                const float eo = (float)real_or(*m, "ior-outer", 1.0);
                const float ei = (float)real_or(*m, "ior-inner", 1.5);
                const Clause* fr = m->find("fresnel");
                if (fr && fr->value != "exact") warn();
                id = s.add_material(PHG_MAT_DIELECTRIC, f3(1, 1, 1), (float)r, eo, ei);
            }
        } else {
            warn();
            id = s.add_material(PHG_MAT_LAMBERT, f3(0.5f, 0.5f, 0.5f), 0, 1, 1);
        }
        cache[key] = id;
        return id;
    }

    // local-space triangles of a geometry resource (9 floats each)
    std::vector<F3> triangles_of(const Resource& g) {
        std::vector<F3> t;
        auto quad = [&](F3 a, F3 b, F3 c, F3 d) {  // (a, b, d), (b, c, d) as GRectangle splits
            t.push_back(a); t.push_back(b); t.push_back(d);
            t.push_back(b); t.push_back(c); t.push_back(d);
        };
        if (g.type == "rectangle") {
            const float w = (float)real_or(g, "width", 1.0), h = (float)real_or(g, "height", 1.0);
            if (!(w > 0 && h > 0)) throw ParseError{"rectangle needs positive width and height"};
            quad(f3(-w / 2, h / 2, 0), f3(-w / 2, -h / 2, 0), f3(w / 2, -h / 2, 0), f3(w / 2, h / 2, 0));
        } else if (g.type == "cuboid") {
            F3 size = f3(1, 1, 1), off = f3(0, 0, 0);
            vec3(g, "size", &size);
            vec3(g, "offset", &off);
            const F3 lo = sub(off, mul(size, 0.5f)), hi = add(off, mul(size, 0.5f));
            auto P = [&](int x, int y, int z) { return f3(x ? hi.x : lo.x, y ? hi.y : lo.y, z ? hi.z : lo.z); };
            // faces counter-clockwise seen from outside
            quad(P(1, 1, 1), P(1, 0, 1), P(1, 0, 0), P(1, 1, 0));  // +x
            quad(P(0, 1, 0), P(0, 0, 0), P(0, 0, 1), P(0, 1, 1));  // -x
            quad(P(0, 1, 0), P(0, 1, 1), P(1, 1, 1), P(1, 1, 0));  // +y
            quad(P(0, 0, 1), P(0, 0, 0), P(1, 0, 0), P(1, 0, 1));  // -y
            quad(P(0, 1, 1), P(0, 0, 1), P(1, 0, 1), P(1, 1, 1));  // +z
            quad(P(1, 1, 0), P(1, 0, 0), P(0, 0, 0), P(0, 1, 0));  // -z
        } else if (g.type == "sphere") {
            const float r = (float)real_or(g, "radius", 1.0);
            if (!(r > 0)) throw ParseError{"sphere needs a positive radius"};
            static const Icosphere ico = make_icosphere(5);
            for (size_t i = 0; i < ico.f.size(); i += 3)
                for (int k = 0; k < 3; ++k) t.push_back(mul(ico.v[ico.f[i + k]], r));
        } else if (g.type == "triangle") {
            F3 a{}, b{}, c{};
            if (!vec3(g, "v-A", &a) || !vec3(g, "v-B", &b) || !vec3(g, "v-C", &c))
                throw ParseError{"triangle needs v-A, v-B and v-C"};
            t.push_back(a); t.push_back(b); t.push_back(c);
        } else if (g.type == "triangle-mesh") {
            const Clause* p = g.find("positions");
            const std::vector<double> v = p ? numbers(p->value) : std::vector<double>();
            if (v.empty() || v.size() % 9 != 0) throw ParseError{"triangle-mesh positions must hold 3 vertices per triangle"};
            for (size_t i = 0; i < v.size(); i += 3) t.push_back(f3((float)v[i], (float)v[i + 1], (float)v[i + 2]));
        } else {
            warn();
        }
        return t;
    }

    F3 to_world(const Resource& a, F3 p) const {
        const F3 s = f3(p.x * a.scale.x, p.y * a.scale.y, p.z * a.scale.z);
        return add(qrot(a.rot, s), a.pos);
    }

    void build(SceneBuf& s) {
        // settings from the session option (SingleFrameRenderSession), with defaults
        st.width = 512; st.height = 512; st.spp = 16; st.max_bounces = 10000;
        st.estimator = PHG_ESTIMATOR_BNEEPT; st.filter = PHG_FILTER_BOX;
        const Resource* sess = session.empty() ? nullptr : get(session);
        const Resource* obs = nullptr;
        const Resource* vis = nullptr;
        const Resource* smp = nullptr;
        if (sess) {
            const Clause* fs = sess->find("frame-size");
            if (fs) {
                const std::vector<double> v = numbers(fs->value);
                if (v.size() == 2 && v[0] >= 1 && v[1] >= 1) { st.width = (uint32_t)v[0]; st.height = (uint32_t)v[1]; }
            }
            if (const Clause* c = sess->find("observer")) obs = get(c->value);
            if (const Clause* c = sess->find("visualizer")) vis = get(c->value);
            if (const Clause* c = sess->find("sample-source")) smp = get(c->value);
        }
        for (const auto& kv : res) {  // without a session: the first of each kind
            const Resource& r = kv.second;
            if (!obs && r.category == "observer") obs = &r;
            if (!vis && r.category == "visualizer") vis = &r;
            if (!smp && r.category == "sample-source") smp = &r;
        }
        if (smp) st.spp = (uint32_t)std::max(1.0, real_or(*smp, "samples", 16));
        // stratified (SGStratified), Halton (SGHalton; HaltonSampleSource.h: permutation
        // default fixed, sequence default original) and uniform-random (SGUniformRandom)
        st.sampler = PHG_SAMPLER_INDEPENDENT;
        if (smp && smp->type == "stratified") st.sampler = PHG_SAMPLER_STRATIFIED;
        if (smp && smp->type == "halton") {
            st.sampler = PHG_SAMPLER_HALTON;
            if (const Clause* c = smp->find("permutation")) {
                if (c->value == "none") st.sampler |= PHG_HALTON_PERM_NONE;
                else if (c->value == "per-digit") st.sampler |= PHG_HALTON_PERM_PER_DIGIT;
                else if (c->value == "owen") st.sampler |= PHG_HALTON_PERM_OWEN;
                else if (!c->value.empty() && c->value != "fixed") warn();
            }
            if (const Clause* c = smp->find("sequence")) {
                if (c->value == "leap") st.sampler |= PHG_HALTON_SEQ_LEAP;
                else if (c->value == "random-leap") st.sampler |= PHG_HALTON_SEQ_RANDOM_LEAP;
                else if (c->value == "random-start") st.sampler |= PHG_HALTON_SEQ_RANDOM_START;
                else if (!c->value.empty() && c->value != "original") warn();
            }
        }
        if (vis) {
            if (vis->type != "path-tracing") warn();
            const Clause* e = vis->find("estimator");
            if (e) {
                if (e->value == "bneept") st.estimator = PHG_ESTIMATOR_BNEEPT;
                else if (e->value == "bvpt") st.estimator = PHG_ESTIMATOR_BVPT;
                else if (e->value == "bvptdl") st.estimator = PHG_ESTIMATOR_BVPTDL;
                else throw ParseError{"estimator " + e->value + " is not on this path (bneept, bvpt, bvptdl)"};
            }
            const Clause* f = vis->find("sample-filter");
            if (f) {
                if (f->value == "box") st.filter = PHG_FILTER_BOX;
                else if (f->value == "gaussian") st.filter = PHG_FILTER_GAUSSIAN;
                else if (f->value == "mn" || f->value == "mitchell-netravali") st.filter = PHG_FILTER_MN;
                else if (f->value == "bh" || f->value == "blackman-harris") st.filter = PHG_FILTER_BH;
                else warn();
            }
            const double mb = real_or(*vis, "max-path-length", -1.0);
            if (mb >= 1) st.max_bounces = (uint32_t)std::min(mb, 10000.0);
        }
        // camera
        F3 pos = f3(0, 0, 0), dir = f3(0, 0, -1), up = f3(0, 1, 0);
        double tan_half = 0.5;  // sensor width 36 mm at 36 mm offset
        if (obs) {
            if (obs->type != "single-lens") warn();
            if (!vec3(*obs, "pos", &pos)) vec3(*obs, "position", &pos);
            if (!vec3(*obs, "dir", &dir) && !vec3(*obs, "direction", &dir) && obs->find("yaw-pitch-row-degrees")) warn();
            vec3(*obs, "up-axis", &up);
            if (real_or(*obs, "lens-radius-mm", 0.0) > 0.0) warn();  // thin lens: pinhole used
            const double fov = real_or(*obs, "fov-degrees", -1.0);
            if (fov > 0) tan_half = std::tan(fov * 3.14159265358979323846 / 360.0);
            else tan_half = 0.5 * real_or(*obs, "sensor-width-mm", 36.0) / real_or(*obs, "sensor-offset-mm", 36.0);
        }
        if (!(dot3(dir, dir) > 0)) throw ParseError{"observer direction is zero"};
        s.camera(pos, add(pos, dir), up, (float)(2.0 * std::atan(tan_half) * 180.0 / 3.14159265358979323846));
        // actors in creation order
        std::map<std::string, uint32_t> mats;
        uint32_t absorber = UINT32_MAX;
        for (const std::string& name : actor_order) {
            const Resource& a = res.at(name);
            if (a.type == "model") {
                const Clause* gref = a.find("geometry");
                const Resource* g = gref ? get(gref->value) : nullptr;
                if (!g) throw ParseError{"model " + name + " has no geometry"};
                const std::vector<F3> tri = triangles_of(*g);
                if (tri.empty()) continue;
                const uint32_t m = material_of(s, mats, a.find("material"));
                for (size_t i = 0; i < tri.size(); i += 3)
                    s.tri(to_world(a, tri[i]), to_world(a, tri[i + 1]), to_world(a, tri[i + 2]), m);
            } else if (a.type == "rectangle-light") {
                const float w = (float)real_or(a, "width", 1.0), h = (float)real_or(a, "height", 1.0);
                if (!(w > 0 && h > 0)) throw ParseError{"rectangle-light needs positive width and height"};
                F3 color = f3(1, 1, 1);
                if (const Clause* c = a.find("color")) color = color_of(c, color);
                const double watts = real_or(a, "watts", 100.0);
                const F3 v[4] = {to_world(a, f3(-w / 2, h / 2, 0)), to_world(a, f3(-w / 2, -h / 2, 0)),
                                 to_world(a, f3(w / 2, -h / 2, 0)), to_world(a, f3(w / 2, h / 2, 0))};
                auto tri_area = [](F3 p, F3 q, F3 r) {
                    const F3 c = crs(sub(q, p), sub(r, p));
                    return 0.5 * std::sqrt((double)dot3(c, c));
                };
                const double area = tri_area(v[0], v[1], v[3]) + tri_area(v[1], v[2], v[3]);
                // AAreaLight::buildEmitter (AAreaLight.cpp:55-58): the color put to the given
                // energy (Spectrum::putEnergy: divided by its colour energy, times the watts),
                // spread over the area and the cosine-weighted hemisphere
                const double energy = phg::color_energy(color.x, color.y, color.z);
                if (!(energy > 0) || !(area > 0)) throw ParseError{"rectangle-light has no power or area"};
                const float k = (float)(watts / (energy * area * 3.14159265358979323846));
                if (absorber == UINT32_MAX) absorber = s.add_material(PHG_MAT_ABSORBER, f3(0, 0, 0), 0, 1, 1);
                const int32_t e = (int32_t)s.emits.size();
                phg_emitter em{};
                em.radiance[0] = color.x * k; em.radiance[1] = color.y * k; em.radiance[2] = color.z * k;
                em.tri_begin = (uint32_t)s.emit_tris.size();
                em.tri_count = 2;
                s.emits.push_back(em);
                s.tri(v[0], v[1], v[3], absorber, e);
                s.tri(v[1], v[2], v[3], absorber, e);
            } else {
                warn();
            }
        }
        if (s.tri_mat.empty()) throw ParseError{"the scene has no triangles"};
    }
};

phg_status fail(char* err, uint32_t err_len, const std::string& msg, phg_status st) {
    if (err && err_len) {
        std::snprintf(err, err_len, "%s", msg.c_str());
    }
    return st;
}

}  // namespace

extern "C" phg_status phg_scene_parse_p2(const char* text, phg_p2_settings* settings, phg_scene_host** out, char* err,
                                         uint32_t err_len) {
    if (!text || !out) return PHG_ERR_INVALID_ARG;
    *out = nullptr;
    SceneBuf* s = new (std::nothrow) SceneBuf();
    if (!s) return PHG_ERR_OUT_OF_MEMORY;
    std::memset(&s->host, 0, sizeof(s->host));
    try {
        Loader ld;
        ld.run(text);
        ld.build(*s);
        if (settings) *settings = ld.st;
    } catch (const ParseError& e) {
        delete s;
        return fail(err, err_len, e.msg, PHG_ERR_INVALID_ARG);
    } catch (const std::bad_alloc&) {
        delete s;
        return fail(err, err_len, "out of memory", PHG_ERR_OUT_OF_MEMORY);
    } catch (const std::exception& e) {
        delete s;
        return fail(err, err_len, e.what(), PHG_ERR_INVALID_ARG);
    }
    s->finish();
    *out = &s->host;  // host is the first member: phg_scene_free converts it back
    return PHG_OK;
}

extern "C" phg_status phg_scene_load_p2(const char* path, phg_p2_settings* settings, phg_scene_host** out, char* err,
                                        uint32_t err_len) {
    if (!path || !out) return PHG_ERR_INVALID_ARG;
    std::string text;
    try {  // nothing throws across the C ABI
        std::ifstream f(path, std::ios::binary);
        if (!f) return fail(err, err_len, std::string("cannot open ") + path, PHG_ERR_INVALID_ARG);
        std::stringstream ss;
        ss << f.rdbuf();
        text = ss.str();
    } catch (const std::bad_alloc&) {
        return fail(err, err_len, "out of memory", PHG_ERR_OUT_OF_MEMORY);
    }
    return phg_scene_parse_p2(text.c_str(), settings, out, err, err_len);
}
