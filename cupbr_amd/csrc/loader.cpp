// Scene file loading: a small XML reader, the OBJ subset the reference reads,
// and a Radiance RGBE (.hdr) reader for environment maps.
//
// SceneLoader::loadFromFile (lib/src/Scene/SceneLoader.cpp:140-454) uses
// tinyxml2 and stb_image, which are empty submodules in the reference; the
// readers here are written for the documents the reference accepts: several
// top-level elements (<header>, <geometry>, <lights>, <volume>,
// <environment>), child elements with text, comments and declarations.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "scene.h"

namespace cupbr {
namespace {

struct XmlNode {
    std::string name;
    std::string text;
    std::vector<std::unique_ptr<XmlNode>> children;
    const XmlNode* child(const char* n) const {  // tinyxml2 FirstChildElement
        for (const auto& c : children)
            if (c->name == n) return c.get();
        return nullptr;
    }
};

class XmlReader {
public:
    explicit XmlReader(const std::string& s) : src(s) {}
    bool parse(XmlNode& root, std::string& err) {
        pos = 0;
        if (!content(root, err)) return false;
        if (pos < src.size()) { err = "unexpected closing tag at offset " + std::to_string(pos); return false; }
        return true;
    }

private:
    const std::string& src;
    size_t pos = 0;

    bool starts(const char* s) const { return src.compare(pos, strlen(s), s) == 0; }
    bool skip_until(const char* s) {
        size_t p = src.find(s, pos);
        if (p == std::string::npos) return false;
        pos = p + strlen(s);
        return true;
    }
    // parses children and text of `parent` until its closing tag (or end of input for the root)
    bool content(XmlNode& parent, std::string& err) {
        while (pos < src.size()) {
            if (src[pos] != '<') {
                size_t p = src.find('<', pos);
                if (p == std::string::npos) p = src.size();
                parent.text += decode(src.substr(pos, p - pos));
                pos = p;
                continue;
            }
            if (starts("<!--")) { if (!skip_until("-->")) { err = "unterminated comment"; return false; } continue; }
            if (starts("<?")) { if (!skip_until("?>")) { err = "unterminated declaration"; return false; } continue; }
            if (starts("<![CDATA[")) {
                size_t p = src.find("]]>", pos);
                if (p == std::string::npos) { err = "unterminated CDATA"; return false; }
                parent.text += src.substr(pos + 9, p - pos - 9);
                pos = p + 3;
                continue;
            }
            if (starts("<!")) { if (!skip_until(">")) { err = "unterminated declaration"; return false; } continue; }
            if (starts("</")) return true;  // the caller checks the name
            // element
            ++pos;
            size_t ns = pos;
            while (pos < src.size() && !isspace((unsigned char)src[pos]) && src[pos] != '>' && src[pos] != '/') ++pos;
            auto node = std::make_unique<XmlNode>();
            node->name = src.substr(ns, pos - ns);
            if (node->name.empty()) { err = "empty element name at offset " + std::to_string(ns); return false; }
            size_t gt = src.find('>', pos);
            if (gt == std::string::npos) { err = "unterminated tag <" + node->name; return false; }
            bool self_closing = gt > 0 && src[gt - 1] == '/';
            pos = gt + 1;
            if (!self_closing) {
                if (!content(*node, err)) return false;
                if (!starts("</")) { err = "missing closing tag for <" + node->name + ">"; return false; }
                pos += 2;
                size_t ce = src.find('>', pos);
                if (ce == std::string::npos) { err = "unterminated closing tag"; return false; }
                std::string cname = trim(src.substr(pos, ce - pos));
                if (cname != node->name) { err = "mismatched </" + cname + "> for <" + node->name + ">"; return false; }
                pos = ce + 1;
            }
            parent.children.push_back(std::move(node));
        }
        return true;
    }
    static std::string decode(const std::string& s) {
        if (s.find('&') == std::string::npos) return s;
        std::string out;
        for (size_t i = 0; i < s.size(); ++i) {
            if (s[i] == '&') {
                static const char* ents[][2] = {{"&lt;", "<"}, {"&gt;", ">"}, {"&amp;", "&"}, {"&quot;", "\""}, {"&apos;", "'"}};
                bool done = false;
                for (auto& e : ents)
                    if (s.compare(i, strlen(e[0]), e[0]) == 0) { out += e[1]; i += strlen(e[0]) - 1; done = true; break; }
                if (done) continue;
            }
            out += s[i];
        }
        return out;
    }

public:
    static std::string trim(const std::string& s) {
        size_t a = 0, b = s.size();
        while (a < b && isspace((unsigned char)s[a])) ++a;
        while (b > a && isspace((unsigned char)s[b - 1])) --b;
        return s.substr(a, b - a);
    }
};

struct LoadError {
    int code;
    std::string msg;
};

const XmlNode& need(const XmlNode* parent, const char* name, const std::string& where) {
    const XmlNode* c = parent ? parent->child(name) : nullptr;
    if (!c) throw LoadError{CUPBR_ERR_PARSE, "missing <" + std::string(name) + "> in " + where};
    return *c;
}

float to_float(const std::string& s, const std::string& where) {  // std::stof
    const char* b = s.c_str();
    char* e = nullptr;
    float v = strtof(b, &e);
    if (e == b) throw LoadError{CUPBR_ERR_PARSE, "not a number: '" + s + "' in " + where};
    return v;
}

void to_vec(const std::string& s, float out[3], const std::string& where) {  // string2vector, :20-33
    std::stringstream ss(s);
    std::string item;
    int k = 0;
    while (std::getline(ss, item, ',') && k < 3) out[k++] = to_float(item, where);
    if (k < 3) throw LoadError{CUPBR_ERR_PARSE, "expected three comma-separated values in " + where};
}

// detail::loadMaterial, SceneLoader.cpp:35-137
cupbr_material load_material(const XmlNode& m, const std::string& where) {
    cupbr_material mat = default_material();
    std::string type = XmlReader::trim(need(&m, "type", where).text);
    static const char* names[] = {"LAMBERT", "PHONG", "MIRROR", "GLASS", "GGX", "VOLUME"};
    int t = -1;
    for (int i = 0; i < 6; ++i)
        if (type == names[i]) t = i;
    if (t < 0) throw LoadError{CUPBR_ERR_MATERIAL_TYPE, "Error while loading material: " + type + " is not a valid material type"};
    mat.type = t;
    if (auto* c = m.child("albedo_e")) to_vec(c->text, mat.albedo_e, where);
    if (auto* c = m.child("albedo_d")) to_vec(c->text, mat.albedo_d, where);
    if (auto* c = m.child("albedo_s")) to_vec(c->text, mat.albedo_s, where);
    if (auto* c = m.child("shininess")) mat.shininess = to_float(c->text, where);
    if (auto* c = m.child("roughness")) mat.roughness = std::max(1e-4f, to_float(c->text, where));
    if (auto* c = m.child("eta")) mat.eta = to_float(c->text, where);
    if (auto* c = m.child("sigma_s")) to_vec(c->text, mat.volume.sigma_s, where);
    if (auto* c = m.child("sigma_a")) to_vec(c->text, mat.volume.sigma_a, where);
    if (auto* c = m.child("g")) mat.volume.g = to_float(c->text, where);
    if (auto* c = m.child("interface"))
        if (XmlReader::trim(c->text) == "GLASS") mat.volume.interface_type = CUPBR_INTERFACE_GLASS;
    return mat;
}

std::string resolve(const std::string& p, const std::string& xml_dir) {
    if (p.empty() || p[0] == '/') return p;
    if (std::ifstream(p).good()) return p;  // the reference opens paths relative to the working directory
    std::string d = xml_dir;
    while (!d.empty()) {  // fall back to the XML file's directory and its ancestors
        std::string cand = d + "/" + p;
        if (std::ifstream(cand).good()) return cand;
        size_t s = d.find_last_of('/');
        if (s == std::string::npos) break;
        d = d.substr(0, s);
    }
    return p;
}

// ObjLoader::loadObj, lib/src/Scene/ObjLoader.cpp:13-87: 'v x y z' lines and 'f a/b/c d/e/f g/h/i'
// faces (vertex slots 0, 3, 6); the box spans every vertex of the file.
void load_obj(const std::string& path, const float pos[3], const float scale[3], cupbr_scene* s, cupbr_object& o) {
    std::ifstream f(path);
    if (!f.good()) throw LoadError{CUPBR_ERR_FILE, "Failed to open OBJ file: " + path};
    std::vector<float> verts;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<uint32_t> faces;
    std::string line;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.size() >= 2 && line[0] == 'v' && line[1] == ' ') {
            std::stringstream ss(line.substr(2));
            std::string item;
            float r[3];
            int k = 0;
            while (std::getline(ss, item, ' ') && k < 3)
                if (!item.empty()) r[k++] = to_float(item, path);
            if (k < 3) throw LoadError{CUPBR_ERR_PARSE, "bad vertex line in " + path};
            for (int a = 0; a < 3; ++a) {
                float v = scale[a] * r[a] + pos[a];
                verts.push_back(v);
                mx[a] = v > mx[a] ? v : mx[a];
                mn[a] = v < mn[a] ? v : mn[a];
            }
        } else if (line.size() >= 2 && line[0] == 'f' && line[1] == ' ') {
            std::string t = line.substr(2);
            for (char& c : t)
                if (c == ' ') c = '/';
            std::stringstream ss(t);
            std::string item;
            std::vector<long> idx;
            while (std::getline(ss, item, '/'))
                if (!item.empty()) idx.push_back(strtol(item.c_str(), nullptr, 10));
            if (idx.size() < 7) throw LoadError{CUPBR_ERR_PARSE, "face line is not 'f a/b/c d/e/f g/h/i' in " + path};
            faces.push_back((uint32_t)(idx[0] - 1));
            faces.push_back((uint32_t)(idx[3] - 1));
            faces.push_back((uint32_t)(idx[6] - 1));
        }
    }
    const size_t nv = verts.size() / 3;
    o.first_triangle = (uint32_t)(s->triangles.size() / 9);
    o.triangle_count = (uint32_t)(faces.size() / 3);
    for (size_t i = 0; i < faces.size(); ++i) {
        if (faces[i] >= nv) throw LoadError{CUPBR_ERR_PARSE, "face index out of range in " + path};
        for (int a = 0; a < 3; ++a) s->triangles.push_back(verts[3 * faces[i] + a]);
    }
    for (int a = 0; a < 3; ++a) { o.params[a] = mn[a]; o.params[3 + a] = mx[a]; }
}

}  // namespace

// Radiance RGBE: "#?RADIANCE"/"#?RGBE" header, "-Y h +X w" resolution, flat or
// new-style run-length scanlines. Decoding per Ward's format: c = m * 2^(e - 136), 0 when e == 0.
int load_rgbe(const std::string& path, std::vector<float>& rgb, uint32_t& w, uint32_t& h) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return set_error(CUPBR_ERR_FILE, "Failed to open environment map: " + path, true);
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, fclose);
    char line[512];
    if (!fgets(line, sizeof line, f) || (strncmp(line, "#?RADIANCE", 10) && strncmp(line, "#?RGBE", 6)))
        return set_error(CUPBR_ERR_PARSE, "not a Radiance RGBE file: " + path, true);
    bool fmt_ok = true;
    for (;;) {
        if (!fgets(line, sizeof line, f)) return set_error(CUPBR_ERR_PARSE, "truncated RGBE header: " + path, true);
        if (line[0] == '\n' || line[0] == '\r') break;
        if (!strncmp(line, "FORMAT=", 7) && strncmp(line + 7, "32-bit_rle_rgbe", 15)) fmt_ok = false;
    }
    if (!fmt_ok) return set_error(CUPBR_ERR_PARSE, "unsupported RGBE format: " + path, true);
    int iw = 0, ih = 0;
    if (!fgets(line, sizeof line, f) || sscanf(line, "-Y %d +X %d", &ih, &iw) != 2 || iw <= 0 || ih <= 0)
        return set_error(CUPBR_ERR_PARSE, "unsupported RGBE orientation (expected -Y h +X w): " + path, true);
    w = (uint32_t)iw;
    h = (uint32_t)ih;
    rgb.assign((size_t)w * h * 3, 0.0f);
    std::vector<unsigned char> sl((size_t)w * 4);
    auto decode = [&](size_t row) {
        for (uint32_t x = 0; x < w; ++x) {
            unsigned char* p = &sl[4 * (size_t)x];
            float* o = &rgb[3 * (row * w + x)];
            if (p[3] == 0) { o[0] = o[1] = o[2] = 0.0f; continue; }
            float fct = std::ldexp(1.0f, (int)p[3] - (128 + 8));
            o[0] = p[0] * fct; o[1] = p[1] * fct; o[2] = p[2] * fct;
        }
    };
    for (uint32_t y = 0; y < h; ++y) {
        unsigned char hd[4];
        if (fread(hd, 1, 4, f) != 4) return set_error(CUPBR_ERR_PARSE, "truncated RGBE data: " + path, true);
        if (w < 8 || w > 0x7fff || hd[0] != 2 || hd[1] != 2 || (hd[2] & 0x80)) {
            // flat scanline (this pixel already read)
            memcpy(&sl[0], hd, 4);
            if (w > 1 && fread(&sl[4], 1, 4 * (size_t)(w - 1), f) != 4 * (size_t)(w - 1))
                return set_error(CUPBR_ERR_PARSE, "truncated RGBE data: " + path, true);
            decode(y);
            continue;
        }
        if ((uint32_t)((hd[2] << 8) | hd[3]) != w) return set_error(CUPBR_ERR_PARSE, "bad RGBE scanline: " + path, true);
        for (int c = 0; c < 4; ++c) {
            uint32_t x = 0;
            while (x < w) {
                int cnt = fgetc(f);
                if (cnt == EOF) return set_error(CUPBR_ERR_PARSE, "truncated RGBE data: " + path, true);
                if (cnt > 128) {
                    cnt -= 128;
                    int v = fgetc(f);
                    if (v == EOF || x + cnt > w) return set_error(CUPBR_ERR_PARSE, "bad RGBE run: " + path, true);
                    while (cnt--) sl[4 * (size_t)x++ + c] = (unsigned char)v;
                } else {
                    if (cnt == 0 || x + cnt > w) return set_error(CUPBR_ERR_PARSE, "bad RGBE run: " + path, true);
                    while (cnt--) {
                        int v = fgetc(f);
                        if (v == EOF) return set_error(CUPBR_ERR_PARSE, "truncated RGBE data: " + path, true);
                        sl[4 * (size_t)x++ + c] = (unsigned char)v;
                    }
                }
            }
        }
        decode(y);
    }
    return CUPBR_OK;
}

// SceneLoader::loadFromFile, lib/src/Scene/SceneLoader.cpp:140-454
int load_scene_xml(const std::string& path, cupbr_scene* s) {
    std::ifstream f(path);
    if (!f.good()) return set_error(CUPBR_ERR_FILE, "Failed to load XML file: " + path + ". Error code: file not found", true);
    std::stringstream buf;
    buf << f.rdbuf();
    std::string text = buf.str();
    XmlNode root;
    std::string err;
    XmlReader rd(text);
    if (!rd.parse(root, err)) return set_error(CUPBR_ERR_FILE, "Failed to load XML file: " + path + ". Error: " + err, true);
    std::string dir = path.find('/') == std::string::npos ? std::string(".") : path.substr(0, path.find_last_of('/'));
    try {
        const XmlNode& header = need(&root, "header", path);
        long scene_size = strtol(need(&header, "scene_size", path).text.c_str(), nullptr, 10);
        long light_count = strtol(need(&header, "light_count", path).text.c_str(), nullptr, 10);
        if (scene_size < 0 || light_count < 0) throw LoadError{CUPBR_ERR_PARSE, "negative counts in <header>"};
        const XmlNode* geo = root.child("geometry");
        if (scene_size > 0 && !geo) throw LoadError{CUPBR_ERR_PARSE, "missing <geometry> in " + path};
        for (long i = 0; i < scene_size; ++i) {
            std::string name = "geometry" + std::to_string(i + 1);
            const XmlNode& g = need(geo, name.c_str(), path);
            std::string where = path + ":" + name;
            std::string type = XmlReader::trim(need(&g, "type", where).text);
            cupbr_object o;
            memset(&o, 0, sizeof o);
            if (type == "PLANE") {
                o.type = CUPBR_PLANE;
                to_vec(need(&g, "position", where).text, o.params, where);
                to_vec(need(&g, "normal", where).text, o.params + 3, where);
            } else if (type == "SPHERE") {
                o.type = CUPBR_SPHERE;
                to_vec(need(&g, "position", where).text, o.params, where);
                o.params[3] = to_float(need(&g, "radius", where).text, where);
            } else if (type == "QUAD") {
                o.type = CUPBR_QUAD;
                to_vec(need(&g, "position", where).text, o.params, where);
                to_vec(need(&g, "normal", where).text, o.params + 3, where);
                to_vec(need(&g, "extend1", where).text, o.params + 6, where);
                to_vec(need(&g, "extend2", where).text, o.params + 9, where);
            } else if (type == "TRIANGLE") {
                o.type = CUPBR_TRIANGLE;
                to_vec(need(&g, "vertex1", where).text, o.params, where);
                to_vec(need(&g, "vertex2", where).text, o.params + 3, where);
                to_vec(need(&g, "vertex3", where).text, o.params + 6, where);
            } else if (type == "MESH") {
                o.type = CUPBR_MESH;
                std::string mp = XmlReader::trim(need(&g, "path", where).text);
                float pos[3] = {0, 0, 0}, sc[3] = {1, 1, 1};
                if (auto* c = g.child("position")) to_vec(c->text, pos, where);
                if (auto* c = g.child("scale")) to_vec(c->text, sc, where);
                load_obj(resolve(mp, dir), pos, sc, s, o);
            } else {
                throw LoadError{CUPBR_ERR_GEOMETRY_TYPE, "Error while loading scene: " + type + " is not a valid geometry type!"};
            }
            o.material = load_material(need(&g, "material", where), where);
            s->objects.push_back(o);
        }
        const XmlNode* lh = root.child("lights");
        if (light_count > 0 && !lh) throw LoadError{CUPBR_ERR_PARSE, "missing <lights> in " + path};
        for (long i = 0; i < light_count; ++i) {
            std::string name = "light" + std::to_string(i + 1);
            const XmlNode& l = need(lh, name.c_str(), path);
            std::string where = path + ":" + name;
            std::string type = XmlReader::trim(need(&l, "type", where).text);
            cupbr_light L;
            memset(&L, 0, sizeof L);
            if (type == "POINT") L.type = CUPBR_POINT;
            else if (type == "AREA") L.type = CUPBR_AREA;
            else throw LoadError{CUPBR_ERR_LIGHT_TYPE, "Error while loading scene: " + type + " is not a valid light type!"};
            if (auto* c = l.child("position")) to_vec(c->text, L.position, where);
            if (auto* c = l.child("intensity")) to_vec(c->text, L.intensity, where);
            if (auto* c = l.child("radiance")) to_vec(c->text, L.radiance, where);
            if (auto* c = l.child("extend1")) to_vec(c->text, L.half_extend1, where);
            if (auto* c = l.child("extend2")) to_vec(c->text, L.half_extend2, where);
            s->lights.push_back(L);
        }
        s->volume = default_volume();
        if (const XmlNode* v = root.child("volume")) {  // :330-344
            to_vec(need(v, "sigma_s", path).text, s->volume.sigma_s, path);
            to_vec(need(v, "sigma_a", path).text, s->volume.sigma_a, path);
            s->volume.g = to_float(need(v, "g", path).text, path);
        }
        if (const XmlNode* e = root.child("environment")) {  // :346-368
            std::string ep = resolve(XmlReader::trim(need(e, "path", path).text), dir);
            int rc = load_rgbe(ep, s->env, s->env_w, s->env_h);
            if (rc != CUPBR_OK) return rc;
        }
    } catch (const LoadError& e) {
        return set_error(e.code, e.msg, true);
    }
    return CUPBR_OK;
}

}  // namespace cupbr
