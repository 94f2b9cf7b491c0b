// Scene files in the reference's XML format (Scene::load and its helpers,
// src/Scene.cpp:13-303, src/Material.cpp:5-70): integrator film size, camera
// (position, angle or lookAt, FOV, lens radius, focal distance) and the model
// instances -- Wavefront OBJ meshes with a transform, an optional material
// override, or an emitter of given total power spread over its area.
//
// The reference parses XML with pugixml and meshes with assimp; neither is in
// this build, so both readers are written here for the subset the format uses:
//  * XML: elements, attributes (single or double quotes, the five predefined
//    entities), self-closing tags, comments, the declaration and DOCTYPE; text
//    content is ignored.
//  * OBJ: v / vt / vn / f (any polygon, fan-triangulated; v, v/vt, v//vn,
//    v/vt/vn; negative indices), o / g / usemtl start a new mesh, mtllib + MTL
//    newmtl / Kd (baseColor) / map_Kd (an albedo texture with the linear filter,
//    Resource.cpp:160-171). Each face corner is its own vertex (no vertex joining), UVs are
//    flipped (v -> 1 - v), meshes without normals get smooth normals (the
//    normalised sum of the unit normals of the faces sharing a position).
//    Faces before any usemtl use a default material (grey 0.6).
// Attributes the reference leaves uninitialised when absent (a missing
// translate / rotate / scale) read as 0 / 0 / 1 here.
#include "Scene.h"
#include "restirpt_host.h"
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <tuple>

namespace {

// ---- XML ------------------------------------------------------------------------------------
struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XmlNode> children;
    const XmlNode* child(const char* n) const {
        for (const XmlNode& c : children)
            if (c.name == n) return &c;
        return nullptr;
    }
    const char* attr(const char* n) const {
        for (const auto& a : attrs)
            if (a.first == n) return a.second.c_str();
        return nullptr;
    }
    std::string attrOr(const char* n, const char* dflt = "") const {
        const char* a = attr(n);
        return a ? a : dflt;
    }
};

class XmlParser {
public:
    explicit XmlParser(const std::string& text) : s(text) {}
    XmlNode parse() {
        XmlNode root;
        root.name = "#document";
        parseContent(root);
        return root;
    }

private:
    const std::string& s;
    size_t i = 0;

    [[noreturn]] void fail(const std::string& what) const {
        throw std::runtime_error("scene XML: " + what + " at offset " + std::to_string(i));
    }
    bool starts(const char* t) const { return s.compare(i, std::strlen(t), t) == 0; }
    void skipSpace() { while (i < s.size() && std::isspace((unsigned char)s[i])) i++; }
    void skipPast(const char* t) {
        size_t k = s.find(t, i);
        if (k == std::string::npos) fail(std::string("missing '") + t + "'");
        i = k + std::strlen(t);
    }
    std::string name() {
        size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == ':' || s[i] == '.')) i++;
        if (i == b) fail("expected a name");
        return s.substr(b, i - b);
    }
    static std::string unescape(const std::string& v) {
        std::string o;
        for (size_t k = 0; k < v.size(); k++) {
            if (v[k] != '&') { o += v[k]; continue; }
            static const std::pair<const char*, char> ents[] = {{"&lt;", '<'}, {"&gt;", '>'}, {"&amp;", '&'},
                                                                {"&quot;", '"'}, {"&apos;", '\''}};
            bool done = false;
            for (const auto& e : ents) {
                if (v.compare(k, std::strlen(e.first), e.first) == 0) { o += e.second; k += std::strlen(e.first) - 1; done = true; break; }
            }
            if (!done) o += '&';
        }
        return o;
    }
    // children of `parent` up to its end tag (or the end of the document for the root)
    void parseContent(XmlNode& parent) {
        for (;;) {
            size_t lt = s.find('<', i);
            if (lt == std::string::npos) {
                if (parent.name != "#document") fail("unterminated <" + parent.name + ">");
                i = s.size();
                return;
            }
            i = lt;
            if (starts("<!--")) { skipPast("-->"); continue; }
            if (starts("<?")) { skipPast("?>"); continue; }
            if (starts("<!")) { skipPast(">"); continue; }
            if (starts("</")) {
                i += 2;
                std::string n = name();
                skipSpace();
                if (i >= s.size() || s[i] != '>') fail("malformed end tag");
                i++;
                if (n != parent.name) fail("end tag </" + n + "> does not close <" + parent.name + ">");
                return;
            }
            i++;  // '<'
            XmlNode node;
            node.name = name();
            for (;;) {
                skipSpace();
                if (i >= s.size()) fail("unterminated tag");
                if (s[i] == '/') {
                    if (i + 1 >= s.size() || s[i + 1] != '>') fail("malformed tag");
                    i += 2;
                    parent.children.push_back(std::move(node));
                    break;
                }
                if (s[i] == '>') {
                    i++;
                    parseContent(node);
                    parent.children.push_back(std::move(node));
                    break;
                }
                std::string an = name();
                skipSpace();
                if (i >= s.size() || s[i] != '=') fail("expected '=' after attribute " + an);
                i++;
                skipSpace();
                if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) fail("expected a quoted value");
                char q = s[i++];
                size_t e = s.find(q, i);
                if (e == std::string::npos) fail("unterminated attribute value");
                node.attrs.emplace_back(an, unescape(s.substr(i, e - i)));
                i = e + 1;
            }
        }
    }
};

// whitespace-separated numbers as `std::stringstream >> x >> y >> z` reads them
void readFloats(const std::string& text, float* out, int n) {
    std::stringstream ss(text);
    for (int k = 0; k < n; k++) ss >> out[k];
}
float readFloat(const XmlNode* node, float dflt) {
    if (!node || !node->attr("value")) return dflt;
    float v = dflt;
    readFloats(node->attr("value"), &v, 1);
    return v;
}

// ---- glm-style 4x4 column-major matrices ----------------------------------------------------
struct M4 { float m[16]; };
M4 identity() { M4 r{}; r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f; return r; }
M4 mul(const M4& a, const M4& b) {
    M4 r{};
    for (int c = 0; c < 4; c++)
        for (int row = 0; row < 4; row++) {
            float v = 0.0f;
            for (int k = 0; k < 4; k++) v += a.m[k * 4 + row] * b.m[c * 4 + k];
            r.m[c * 4 + row] = v;
        }
    return r;
}
M4 translation(const float t[3]) { M4 r = identity(); r.m[12] = t[0]; r.m[13] = t[1]; r.m[14] = t[2]; return r; }
M4 scaling(float x, float y, float z) { M4 r = identity(); r.m[0] = x; r.m[5] = y; r.m[10] = z; return r; }
// glm::rotate(mat4(1), angle, axis): Rodrigues' formula on the normalised axis
M4 rotation(float radians, float ax, float ay, float az) {
    const float len = std::sqrt(ax * ax + ay * ay + az * az);
    ax /= len; ay /= len; az /= len;
    const float c = std::cos(radians), s = std::sin(radians);
    const float tx = (1.0f - c) * ax, ty = (1.0f - c) * ay, tz = (1.0f - c) * az;
    M4 r = identity();
    r.m[0] = c + tx * ax;      r.m[1] = tx * ay + s * az; r.m[2] = tx * az - s * ay;
    r.m[4] = ty * ax - s * az; r.m[5] = c + ty * ay;      r.m[6] = ty * az + s * ax;
    r.m[8] = tz * ax + s * ay; r.m[9] = tz * ay - s * ax; r.m[10] = c + tz * az;
    return r;
}
float radians(float deg) { return deg * 0.01745329251994329577f; }

// ModelInstance::modelMatrix (src/Model.cpp:11-21)
M4 modelMatrix(const float pos[3], const float rot[3], const float scale[3]) {
    M4 m = mul(identity(), translation(pos));
    m = mul(m, rotation(radians(rot[0]), 0.0f, 0.0f, 1.0f));
    m = mul(m, rotation(radians(rot[1] + 90.0f), 1.0f, 0.0f, 0.0f));
    m = mul(m, rotation(radians(rot[2]), 0.0f, 1.0f, 0.0f));
    return mul(m, scaling(scale[0], scale[2], scale[1]));
}

Vec3 xformPoint(const M4& m, const float p[3]) {
    Vec3 r;
    r.x = m.m[0] * p[0] + m.m[4] * p[1] + m.m[8] * p[2] + m.m[12];
    r.y = m.m[1] * p[0] + m.m[5] * p[1] + m.m[9] * p[2] + m.m[13];
    r.z = m.m[2] * p[0] + m.m[6] * p[1] + m.m[10] * p[2] + m.m[14];
    return r;
}

// ---- OBJ / MTL ------------------------------------------------------------------------------
struct ObjMaterial { std::string name; float kd[3] = {0.6f, 0.6f, 0.6f}; std::string map; };
struct ObjMesh {
    uint32_t material = 0;   // into ObjModel::materials (0 = the default material)
    std::vector<rt_mesh_vertex> vertices;
    std::vector<uint32_t> indices;  // local to the mesh
    bool hasNormals = true;
};
struct ObjModel { std::vector<ObjMaterial> materials; std::vector<ObjMesh> meshes; };

std::string dirOf(const std::string& path) {
    size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

// Resource::addImage (src/Resource.cpp:20-35): one texture per file path, loaded once.
// Only PNG files are read (readPng); a file that cannot be read is an error here (the
// reference leaves such a material untextured).
uint32_t addTexture(Scene& S, std::map<std::string, uint32_t>& index, const std::string& path, uint32_t filter) {
    auto it = index.find(path);
    if (it != index.end()) return it->second;
    Texture t = readPng(path);
    t.filter = filter;
    S.textures.push_back(std::move(t));
    const uint32_t k = (uint32_t)S.textures.size() - 1;
    index[path] = k;
    return k;
}

void readMtl(const std::string& path, std::vector<ObjMaterial>& mats) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("scene: cannot open material library " + path);
    std::string line;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "newmtl") {
            ObjMaterial m;
            ss >> m.name;
            mats.push_back(m);
        } else if (key == "Kd" && !mats.empty()) {
            ss >> mats.back().kd[0] >> mats.back().kd[1] >> mats.back().kd[2];
        } else if (key == "map_Kd" && !mats.empty()) {
            std::string file;
            ss >> file;
            mats.back().map = file.empty() || file[0] == '/' ? file : dirOf(path) + file;
        }
    }
}

void smoothNormals(ObjMesh& mesh) {
    std::map<std::tuple<float, float, float>, std::array<float, 3>> acc;
    auto key = [](const rt_mesh_vertex& v) { return std::make_tuple(v.pos[0], v.pos[1], v.pos[2]); };
    std::vector<std::array<float, 3>> faceN(mesh.indices.size() / 3);
    for (size_t t = 0; t + 2 < mesh.indices.size(); t += 3) {
        const float* a = mesh.vertices[mesh.indices[t]].pos;
        const float* b = mesh.vertices[mesh.indices[t + 1]].pos;
        const float* c = mesh.vertices[mesh.indices[t + 2]].pos;
        float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (l > 0.0f) { n[0] /= l; n[1] /= l; n[2] /= l; }
        for (int k = 0; k < 3; k++) {
            auto& s = acc[key(mesh.vertices[mesh.indices[t + k]])];
            s[0] += n[0]; s[1] += n[1]; s[2] += n[2];
        }
    }
    for (rt_mesh_vertex& v : mesh.vertices) {
        auto it = acc.find(key(v));
        if (it == acc.end()) continue;
        const auto& s = it->second;
        const float l = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        for (int k = 0; k < 3; k++) v.norm[k] = l > 0.0f ? s[k] / l : 0.0f;
    }
}

ObjModel readObj(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("scene: cannot open model " + path);
    ObjModel model;
    model.materials.push_back(ObjMaterial{"DefaultMaterial"});
    std::vector<std::array<float, 3>> P, N;
    std::vector<std::array<float, 2>> T;
    uint32_t curMat = 0;
    bool newMesh = true;
    auto resolve = [](long idx, size_t n) -> long { return idx < 0 ? (long)n + idx : idx - 1; };
    std::string line;
    size_t lineNo = 0;
    while (std::getline(f, line)) {
        lineNo++;
        std::stringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "v") { std::array<float, 3> p{}; ss >> p[0] >> p[1] >> p[2]; P.push_back(p); }
        else if (key == "vn") { std::array<float, 3> n{}; ss >> n[0] >> n[1] >> n[2]; N.push_back(n); }
        else if (key == "vt") { std::array<float, 2> t{}; ss >> t[0] >> t[1]; T.push_back(t); }
        else if (key == "o" || key == "g") newMesh = true;
        else if (key == "mtllib") {
            std::string lib;
            ss >> lib;
            readMtl(dirOf(path) + lib, model.materials);
        } else if (key == "usemtl") {
            std::string name;
            ss >> name;
            uint32_t m = 0;
            for (uint32_t k = 1; k < model.materials.size(); k++)
                if (model.materials[k].name == name) m = k;
            if (m != curMat) { curMat = m; newMesh = true; }
        } else if (key == "f") {
            std::vector<rt_mesh_vertex> corners;
            bool hasN = true;
            std::string tok;
            while (ss >> tok) {
                long vi = 0, ti = 0, ni = 0;
                int parts = 0;
                size_t a = tok.find('/');
                vi = std::stol(tok.substr(0, a));
                parts = 1;
                if (a != std::string::npos) {
                    size_t b = tok.find('/', a + 1);
                    std::string ts = tok.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1);
                    if (!ts.empty()) ti = std::stol(ts);
                    if (b != std::string::npos && b + 1 < tok.size()) { ni = std::stol(tok.substr(b + 1)); parts = 3; }
                }
                (void)parts;
                long pv = resolve(vi, P.size());
                if (pv < 0 || (size_t)pv >= P.size())
                    throw std::runtime_error("scene: vertex index out of range in " + path + ":" + std::to_string(lineNo));
                rt_mesh_vertex v{};
                for (int k = 0; k < 3; k++) v.pos[k] = P[(size_t)pv][k];
                if (ti != 0) {
                    long pt = resolve(ti, T.size());
                    if (pt >= 0 && (size_t)pt < T.size()) { v.uvx = T[(size_t)pt][0]; v.uvy = 1.0f - T[(size_t)pt][1]; }
                }
                if (ni != 0) {
                    long pn = resolve(ni, N.size());
                    if (pn < 0 || (size_t)pn >= N.size())
                        throw std::runtime_error("scene: normal index out of range in " + path + ":" + std::to_string(lineNo));
                    for (int k = 0; k < 3; k++) v.norm[k] = N[(size_t)pn][k];
                } else {
                    hasN = false;
                }
                corners.push_back(v);
            }
            if (corners.size() < 3) continue;
            if (newMesh || model.meshes.empty()) {
                model.meshes.push_back(ObjMesh{});
                model.meshes.back().material = curMat;
                newMesh = false;
            }
            ObjMesh& mesh = model.meshes.back();
            const uint32_t base = (uint32_t)mesh.vertices.size();
            mesh.vertices.insert(mesh.vertices.end(), corners.begin(), corners.end());
            for (uint32_t k = 1; k + 1 < corners.size(); k++) {
                mesh.indices.push_back(base);
                mesh.indices.push_back(base + k);
                mesh.indices.push_back(base + k + 1);
            }
            mesh.hasNormals = mesh.hasNormals && hasN;
        }
    }
    for (ObjMesh& m : model.meshes)
        if (!m.hasNormals) smoothNormals(m);
    return model;
}

// loadMaterialNoBaseColor (src/Material.cpp:5-70) into rt_material; false for type "default"
bool loadMaterial(const XmlNode& node, rt_material& mat) {
    const std::string type = node.attrOr("type");
    mat = rt_material{};
    mat.baseColor[0] = mat.baseColor[1] = mat.baseColor[2] = 1.0f;
    mat.type = 1;  // Lambertian
    mat.textureIdx = 0xffffffffu;
    mat.metallic = 0.0f; mat.roughness = 1.0f; mat.ior = 1.5f;
    auto f = [&](const char* n, float& v) { v = readFloat(node.child(n), v); };
    if (type == "default") return false;
    if (type == "metalWorkflow") { f("metallic", mat.metallic); f("roughness", mat.roughness); mat.type = 2; }
    else if (type == "metal") { f("roughness", mat.roughness); f("ior", mat.ior); mat.type = 3; }
    else if (type == "dielectric") { f("ior", mat.ior); f("roughness", mat.roughness); mat.type = 4; }
    else if (type == "thinDielectric") { f("ior", mat.ior); mat.type = 5; }
    else if (type == "lambertian") mat.type = 1;
    else if (type == "fake") mat.type = 6;
    return true;
}

}  // namespace

// Scene::load (src/Scene.cpp:104-303)
std::unique_ptr<Scene> loadSceneFile(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("scene: cannot open " + path);
    std::stringstream buf;
    buf << f.rdbuf();
    const std::string text = buf.str();
    XmlNode doc = XmlParser(text).parse();
    const XmlNode* sceneNode = doc.child("scene");
    if (!sceneNode) throw std::runtime_error("Scene: failed to load");
    auto S = std::make_unique<Scene>();
    std::map<std::string, uint32_t> textureIndex;
    // Resource() starts with an empty magenta material (src/Resource.cpp:36-40)
    rt_material empty{};
    empty.baseColor[0] = 1.0f; empty.baseColor[1] = 0.0f; empty.baseColor[2] = 1.0f;
    empty.type = 1; empty.textureIdx = 0xffffffffu; empty.metallic = 0.0f; empty.roughness = 1.0f; empty.ior = 1.5f;
    S->materials.push_back(empty);

    // integrator (Scene::loadIntegrator): film size
    uint32_t W = 1, H = 1;
    if (const XmlNode* integ = sceneNode->child("integrator"))
        if (const XmlNode* size = integ->child("size")) {
            W = (uint32_t)std::max(1, std::atoi(size->attrOr("width", "1").c_str()));
            H = (uint32_t)std::max(1, std::atoi(size->attrOr("height", "1").c_str()));
        }
    // camera (Scene::loadCamera, src/Scene.cpp:154-189): position, then angle if present, else
    // lookAt; FOV, lens radius and focal distance read as 0 when absent. A file without a camera
    // keeps the default camera (the reference would look at its own position).
    Camera c;
    c.setFilmSize(W, H);
    if (const XmlNode* cam = sceneNode->child("camera")) {
        float v[3] = {0, 0, 0};
        if (const XmlNode* p = cam->child("position")) readFloats(p->attrOr("value"), v, 3);
        const Vec3 camPos{v[0], v[1], v[2]};
        c.setPos(camPos);
        float q[3] = {0, 0, 0};
        if (const XmlNode* a = cam->child("angle")) {
            readFloats(a->attrOr("value"), q, 3);
            c.setAngle({q[0], q[1], q[2]});
        } else {
            if (const XmlNode* l = cam->child("lookAt")) readFloats(l->attrOr("value"), q, 3);
            if (q[0] == camPos.x && q[1] == camPos.y && q[2] == camPos.z)
                throw std::runtime_error("scene: camera needs an angle or a lookAt point away from its position");
            c.lookAt({q[0], q[1], q[2]});
        }
        c.setFOV(readFloat(cam->child("fov"), 0.0f));
        S->lensRadius = readFloat(cam->child("lensRadius"), 0.0f);
        S->focalDist = readFloat(cam->child("focalDistance"), 0.0f);
        c.setLens(S->lensRadius, S->focalDist);
    }
    S->camera = c;
    S->fov = c.FOV();
    S->filmSize[0] = W;
    S->filmSize[1] = H;
    // a still camera for every frame index, with per-frame seeds
    uint64_t state = 0x853c49e6748fea9bull;
    for (uint32_t k = 0; k < 1024; k++) {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        S->cameraPath.push_back({c.pos(), c.angle(), (uint32_t)(state >> 32)});
    }

    // model instances (Scene::loadModels, loadModelInstance)
    const std::string dir = dirOf(path);
    std::vector<rt_triangle_light> lights;
    if (const XmlNode* models = sceneNode->child("modelInstances")) {
        for (const XmlNode& m : models->children) {
            const bool isLight = m.attrOr("type") == "light";
            float power[3] = {0, 0, 0};
            if (isLight) {
                readFloats(m.child("radiance") ? m.child("radiance")->attrOr("value") : "", power, 3);
                // a zero-power light is filed with the objects by the reference, at an index offset
                // into the wrong mesh table (src/Scene.cpp:199-215): rejected here
                if (!(power[0] * power[0] + power[1] * power[1] + power[2] * power[2] > 0.0f))
                    throw std::runtime_error("scene: light model '" + m.attrOr("name") + "' has no power");
            }
            float pos[3] = {0, 0, 0}, scl[3] = {1, 1, 1}, rot[3] = {0, 0, 0};
            if (const XmlNode* t = m.child("transform")) {
                if (t->attr("translate")) readFloats(t->attr("translate"), pos, 3);
                if (t->attr("scale")) readFloats(t->attr("scale"), scl, 3);
                if (t->attr("rotate")) readFloats(t->attr("rotate"), rot, 3);
            }
            const M4 M = modelMatrix(pos, rot, scl);
            const ObjModel obj = readObj(dir + m.attrOr("path"));
            if (isLight) {
                // light triangles: world vertices, geometric normal (flipped on request),
                // radiance = power / total area (src/Scene.cpp:217-287). Every triangle is
                // filled; the reference's per-thread split leaves the last (count mod threads)
                // triangles of a light zero when the count exceeds and is not a multiple of its
                // hardware thread count.
                const bool flip = m.attrOr("flip") == "true";
                const size_t first = lights.size();
                float sumArea = 0.0f;
                for (const ObjMesh& mesh : obj.meshes) {
                    for (size_t t = 0; t + 2 < mesh.indices.size(); t += 3) {
                        rt_triangle_light L{};
                        const Vec3 a = xformPoint(M, mesh.vertices[mesh.indices[t]].pos);
                        const Vec3 b = xformPoint(M, mesh.vertices[mesh.indices[t + 1]].pos);
                        const Vec3 cc = xformPoint(M, mesh.vertices[mesh.indices[t + 2]].pos);
                        L.v0[0] = a.x; L.v0[1] = a.y; L.v0[2] = a.z;
                        L.v1[0] = b.x; L.v1[1] = b.y; L.v1[2] = b.z;
                        L.v2[0] = cc.x; L.v2[1] = cc.y; L.v2[2] = cc.z;
                        const float e1[3] = {b.x - a.x, b.y - a.y, b.z - a.z}, e2[3] = {cc.x - a.x, cc.y - a.y, cc.z - a.z};
                        float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                      e1[0] * e2[1] - e1[1] * e2[0]};
                        const float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                        L.area = 0.5f * len;
                        for (int k = 0; k < 3; k++) n[k] = len > 0.0f ? n[k] / len : 0.0f;
                        if (flip) for (int k = 0; k < 3; k++) n[k] = -n[k];
                        L.nx = n[0]; L.ny = n[1]; L.nz = n[2];
                        sumArea += L.area;
                        lights.push_back(L);
                    }
                }
                for (size_t k = first; k < lights.size(); k++)
                    for (int ch = 0; ch < 3; ch++) lights[k].radiance[ch] = power[ch] / sumArea;
                continue;
            }
            // object: vertices and indices appended (indices absolute), one material per OBJ
            // material, per-triangle material indices, one instance
            const uint32_t materialOffset = (uint32_t)S->materials.size();
            for (const ObjMaterial& om : obj.materials) {
                rt_material mat = empty;
                for (int k = 0; k < 3; k++) mat.baseColor[k] = om.kd[k];
                if (!om.map.empty()) mat.textureIdx = addTexture(*S, textureIndex, om.map, 1u);
                S->materials.push_back(mat);
            }
            rt_object_instance inst{};
            inst.indexOffset = (uint32_t)S->indices.size();
            for (const ObjMesh& mesh : obj.meshes) {
                const uint32_t vbase = (uint32_t)S->vertices.size();
                S->vertices.insert(S->vertices.end(), mesh.vertices.begin(), mesh.vertices.end());
                for (uint32_t idx : mesh.indices) S->indices.push_back(idx + vbase);
                for (size_t t = 0; t < mesh.indices.size() / 3; t++)
                    S->materialIndices.push_back((int32_t)(materialOffset + mesh.material));
            }
            inst.indexCount = (uint32_t)S->indices.size() - inst.indexOffset;
            makeInstanceTransform(inst, M.m);
            S->objectInstances.push_back(inst);
            // <material> overrides the model's mesh materials (src/Scene.cpp:59-98)
            if (const XmlNode* matNode = m.child("material")) {
                rt_material over;
                if (loadMaterial(*matNode, over)) {
                    bool overrideColor = false;
                    float bc[3] = {1, 1, 1};
                    uint32_t texIdx = 0xffffffffu;
                    if (const XmlNode* b = matNode->child("baseColor")) {
                        if (b->attr("value")) readFloats(b->attr("value"), bc, 3);
                        if (const char* img = b->attr("image")) {
                            const char* filt = b->attr("filter");
                            const uint32_t filter = (filt && std::string(filt) == "nearest") ? 0u : 1u;
                            texIdx = addTexture(*S, textureIndex, dir + img, filter);
                        }
                        overrideColor = true;
                    }
                    for (const ObjMesh& mesh : obj.meshes) {
                        const uint32_t idx = materialOffset + mesh.material;
                        rt_material mm = over;
                        mm.textureIdx = texIdx != 0xffffffffu ? texIdx : S->materials[idx].textureIdx;
                        for (int k = 0; k < 3; k++) mm.baseColor[k] = overrideColor ? bc[k] : S->materials[idx].baseColor[k];
                        S->materials[idx] = mm;
                    }
                }
            }
        }
    }
    S->triangleLights = lights;
    S->buildLightDataStructure();
    return S;
}
