// Scene.cpp -- reference-shaped scene API (TransformComponent, RTModel/loadModel, GameObject,
// RaytraceScene) over the C ABI.  Reference: RaytracerGPU_MastersProject/VulkanWrapper/.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>

#include "rtgpu/Scene.hpp"

namespace rtgpu {

Context::Context(int device) { check(rt_create(&ctx_, device), "rt_create"); }
Context::~Context() {
    if (ctx_) rt_destroy(ctx_);
}
void Context::check(int status, const char *what) const {
    if (status != RT_OK) throw std::runtime_error(std::string(what) + ": " + (ctx_ ? rt_last_error(ctx_) : "no context"));
}

}  // namespace rtgpu

using rtgpu::vec3;

// TransformComponent.cpp:9-45 -- translate * Ry * Rx * Rz * scale (Tait-Bryan Y, X, Z), column-major
rtgpu::mat4 TransformComponent::mat4() const {
    const float c3 = std::cos(rotation.z), s3 = std::sin(rotation.z);
    const float c2 = std::cos(rotation.x), s2 = std::sin(rotation.x);
    const float c1 = std::cos(rotation.y), s1 = std::sin(rotation.y);
    rtgpu::mat4 r;
    r.m[0][0] = scale.x * (c1 * c3 + s1 * s2 * s3);
    r.m[0][1] = scale.x * (c2 * s3);
    r.m[0][2] = scale.x * (c1 * s2 * s3 - c3 * s1);
    r.m[0][3] = 0.0f;
    r.m[1][0] = scale.y * (c3 * s1 * s2 - c1 * s3);
    r.m[1][1] = scale.y * (c2 * c3);
    r.m[1][2] = scale.y * (c1 * c3 * s2 + s1 * s3);
    r.m[1][3] = 0.0f;
    r.m[2][0] = scale.z * (c2 * s1);
    r.m[2][1] = scale.z * (-s2);
    r.m[2][2] = scale.z * (c1 * c2);
    r.m[2][3] = 0.0f;
    r.m[3][0] = translation.x;
    r.m[3][1] = translation.y;
    r.m[3][2] = translation.z;
    r.m[3][3] = 1.0f;
    return r;
}

// TransformComponent.cpp:61-84 -- (R S)^-T = R S^-1
rtgpu::mat3 TransformComponent::normalMatrix() const {
    const float c3 = std::cos(rotation.z), s3 = std::sin(rotation.z);
    const float c2 = std::cos(rotation.x), s2 = std::sin(rotation.x);
    const float c1 = std::cos(rotation.y), s1 = std::sin(rotation.y);
    const vec3 inv{1.0f / scale.x, 1.0f / scale.y, 1.0f / scale.z};
    rtgpu::mat3 r;
    r.m[0][0] = inv.x * (c1 * c3 + s1 * s2 * s3);
    r.m[0][1] = inv.x * (c2 * s3);
    r.m[0][2] = inv.x * (c1 * s2 * s3 - c3 * s1);
    r.m[1][0] = inv.y * (c3 * s1 * s2 - c1 * s3);
    r.m[1][1] = inv.y * (c2 * c3);
    r.m[1][2] = inv.y * (c1 * c3 * s2 + s1 * s3);
    r.m[2][0] = inv.z * (c2 * s1);
    r.m[2][1] = inv.z * (-s2);
    r.m[2][2] = inv.z * (c1 * c2);
    return r;
}

// ---------------- RTModel ----------------
RTModel_Triangles::RTModel_Triangles(std::vector<SceneTypes::CPU::Triangle> &&t, SceneTypes::GPU::Material mat)
    : triangles(std::move(t)), material(mat) {}
const std::vector<SceneTypes::CPU::Triangle> &RTModel_Triangles::getTriangles() const { return triangles; }
SceneTypes::GPU::Material RTModel_Triangles::getMaterialType() const { return material; }

RTModel_Sphere::RTModel_Sphere(f32 r, SceneTypes::GPU::Material mat) : center(0.0f), radius(r), material(mat) {}
const vec3 &RTModel_Sphere::getCenter() const { return center; }
f32 RTModel_Sphere::getRadius() const { return radius; }
SceneTypes::GPU::Material RTModel_Sphere::getMaterialType() const { return material; }

namespace {
std::mutex g_model_dir_mu;
std::string g_model_dir;
}  // namespace

void setModelDirectory(const std::string &dir) {
    std::lock_guard<std::mutex> lk(g_model_dir_mu);
    g_model_dir = dir;
}

std::string modelPath(const std::string &filepath) {
    if (!filepath.empty() && filepath[0] == '/') return filepath;
    std::string dir;
    {
        std::lock_guard<std::mutex> lk(g_model_dir_mu);
        dir = g_model_dir;
    }
    if (dir.empty()) {
        if (const char *e = std::getenv("RTGPU_MODEL_DIR")) dir = e;
    }
    if (dir.empty()) return filepath;  // relative to the working directory, as the reference
    // "models/quad.obj" in the reference's Scenes.cpp: strip a leading "models/" when a directory is set
    std::string rel = filepath.rfind("models/", 0) == 0 ? filepath.substr(7) : filepath;
    return dir + "/" + rel;
}

// OBJ parsing: `v x y z [w]` and `f` records with v, v/vt, v/vt/vn or v//vn references (1-based,
// negative = relative to the end).  The reference loads meshes with tinyobjloader's LoadObj, which
// triangulates by default (RTModel.cpp:53); its triangulation of a quad (exportGroupsToShape in the
// library's 2.x releases) splits along the shorter diagonal, with the squared lengths summed in
// float: (0,1,2),(0,2,3) when |v2-v0|^2 < |v3-v1|^2, else (0,1,3),(1,2,3).  Larger polygons, which
// that library ear-clips in a projection plane, are split into fans (v0, vi, vi+1) here.
std::vector<SceneTypes::CPU::Triangle> loadObjTriangles(const std::string &filepath) {
    const std::string path = modelPath(filepath);
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open OBJ file " + path);
    std::vector<vec3> verts;
    std::vector<SceneTypes::CPU::Triangle> tris;
    std::string line;
    size_t lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag) || tag[0] == '#') continue;
        if (tag == "v") {
            float x, y, z;
            if (!(ss >> x >> y >> z)) throw std::runtime_error(path + ":" + std::to_string(lineno) + ": bad vertex");
            verts.push_back(vec3{x, y, z});
        } else if (tag == "f") {
            std::vector<long> idx;
            std::string tok;
            while (ss >> tok) {
                long v = std::strtol(tok.c_str(), nullptr, 10);
                if (v < 0) v = (long)verts.size() + v + 1;
                if (v < 1 || v > (long)verts.size())
                    throw std::runtime_error(path + ":" + std::to_string(lineno) + ": vertex index out of range");
                idx.push_back(v - 1);
            }
            if (idx.size() == 4) {
                const vec3 &a = verts[idx[0]], &b = verts[idx[1]], &c = verts[idx[2]], &d = verts[idx[3]];
                const f32 ex = c.x - a.x, ey = c.y - a.y, ez = c.z - a.z;
                const f32 fx = d.x - b.x, fy = d.y - b.y, fz = d.z - b.z;
                const f32 sqr02 = ex * ex + ey * ey + ez * ez, sqr13 = fx * fx + fy * fy + fz * fz;
                if (sqr02 < sqr13) {
                    tris.push_back(SceneTypes::CPU::Triangle{a, b, c});
                    tris.push_back(SceneTypes::CPU::Triangle{a, c, d});
                } else {
                    tris.push_back(SceneTypes::CPU::Triangle{a, b, d});
                    tris.push_back(SceneTypes::CPU::Triangle{b, c, d});
                }
                continue;
            }
            for (size_t k = 1; k + 1 < idx.size(); ++k)
                tris.push_back(SceneTypes::CPU::Triangle{verts[idx[0]], verts[idx[k]], verts[idx[k + 1]]});
        }
    }
    return tris;
}

std::unique_ptr<RTModel> loadModel(const std::string &filepath, vec3 color) {
    return loadModel(filepath, SceneTypes::GPU::Material(color, SceneTypes::MaterialType::DIFFUSE));
}
std::unique_ptr<RTModel> loadModel(const std::string &filepath, SceneTypes::GPU::Material mat) {
    return std::make_unique<RTModel>(RTModel_Triangles(loadObjTriangles(filepath), mat));
}
std::unique_ptr<RTModel> loadModel(f32 radius, SceneTypes::GPU::Material mat) {
    return std::make_unique<RTModel>(RTModel_Sphere(radius, mat));
}

// ---------------- GameObject ----------------
namespace {
GameObjectId next_id() {
    static std::mutex mu;
    static GameObjectId current = 1;
    std::lock_guard<std::mutex> lk(mu);
    return current++;
}
}  // namespace

GameObject GameObject::createGameObject() { return GameObject{next_id()}; }
void GameObject::setModel(std::shared_ptr<RTModel> m, bool triangular) {
    if (!m) throw std::invalid_argument("setModel: null model");
    if (triangular != std::holds_alternative<RTModel_Triangles>(*m))
        throw std::invalid_argument("setModel: model kind does not match `triangular`");
    isTriangleModel = triangular;
    isSphereModel = !triangular;
    model = std::move(m);
}
std::shared_ptr<RTModel> GameObject::getModel() const { return model; }
GameObjectId GameObject::getId() const { return id; }
GameObject::GameObject(const GameObject &o)
    : id(next_id()), model(o.model), color(o.color), transform(o.transform), isTriangleModel(o.isTriangleModel),
      isSphereModel(o.isSphereModel) {}

// ---------------- RaytraceScene ----------------
RaytraceScene::RaytraceScene(rtgpu::Context &ctx) : context(ctx) {}
RaytraceScene::~RaytraceScene() { releaseBuffers(); }

void RaytraceScene::setMaxRaytraceDepth(u32 d) { maxRaytraceDepth = d; }
void RaytraceScene::setRaysPerPixel(u32 r) { raysPerPixel = r; }
u32 RaytraceScene::getMaxRaytraceDepth() const { return maxRaytraceDepth; }
u32 RaytraceScene::getRaysPerPixel() const { return raysPerPixel; }
CameraGameObject &RaytraceScene::getCamera() { return camera; }
void RaytraceScene::setBuildOptions(const rt_build_options &o) {
    buildOptions = o;
    dirty = true;
}

void RaytraceScene::addGameObject(GameObject &&g) {
    if (g.getModel() != nullptr) {  // RaytraceScene.cpp:46-52: objects without a model are ignored
        gameObjects.push_back(std::move(g));
        dirty = true;
    }
}

GameObject &RaytraceScene::getGameObject(GameObjectId id) {
    for (auto &g : gameObjects)
        if (g.getId() == id) return g;
    throw std::out_of_range("no game object with id " + std::to_string(id));
}

bool RaytraceScene::removeGameObject(GameObjectId id) {
    for (auto it = gameObjects.begin(); it != gameObjects.end(); ++it)
        if (it->getId() == id) {
            gameObjects.erase(it);
            dirty = true;
            return true;
        }
    return false;
}

// RaytraceScene.cpp:157-200
void RaytraceScene::moveGameObjectsToHostVectors() {
    models.clear();
    triangles.clear();
    spheres.clear();
    materials.clear();
    for (const GameObject &g : gameObjects) {
        const u32 modelIndex = (u32)models.size();
        models.push_back(SceneTypes::GPU::Model{g.transform.mat4()});
        if (const auto *tm = std::get_if<RTModel_Triangles>(g.getModel().get())) {
            const u32 materialIndex = (u32)materials.size();
            materials.push_back(tm->getMaterialType());
            for (const auto &t : tm->getTriangles()) {
                rt_triangle gt{};
                const vec3 vs[3] = {t.v0, t.v1, t.v2};
                float *dst[3] = {gt.v0, gt.v1, gt.v2};
                for (int k = 0; k < 3; ++k) {
                    dst[k][0] = vs[k].x; dst[k][1] = vs[k].y; dst[k][2] = vs[k].z; dst[k][3] = 1.0f;
                }
                gt.materialIndex = materialIndex;
                gt.modelIndex = modelIndex;
                triangles.push_back(gt);
            }
        } else if (const auto *sm = std::get_if<RTModel_Sphere>(g.getModel().get())) {
            const u32 materialIndex = (u32)materials.size();
            materials.push_back(sm->getMaterialType());
            rt_sphere s{};
            s.center[0] = sm->getCenter().x; s.center[1] = sm->getCenter().y; s.center[2] = sm->getCenter().z;
            s.center[3] = 1.0f;
            s.radius = sm->getRadius();
            s.materialIndex = materialIndex;
            s.modelIndex = modelIndex;
            spheres.push_back(s);
        }
    }
}

void RaytraceScene::releaseBuffers() {
    if (buffersCreated && !context.hostOnly()) {
        rt_scene_release(context.get(), sceneId);
        buffersCreated = false;
        sceneId = 0;
    }
}

void RaytraceScene::createBuffers() {
    releaseBuffers();
    if (context.hostOnly()) {  // flattening only
        buffersCreated = true;
        dirty = false;
        return;
    }
    std::vector<float> mats16(16 * models.size());
    for (size_t i = 0; i < models.size(); ++i) std::memcpy(&mats16[16 * i], models[i].modelMatrix.data(), 64);
    uint32_t id = 0;
    context.check(rt_scene_upload(context.get(), &id, triangles.data(), (uint32_t)triangles.size(), spheres.data(),
                                  (uint32_t)spheres.size(), reinterpret_cast<const rt_material *>(materials.data()),
                                  (uint32_t)materials.size(), mats16.data(), (uint32_t)models.size(), nullptr, 0,
                                  &buildOptions),
                  "rt_scene_upload");
    sceneId = id;
    buffersCreated = true;
    dirty = false;
}

void RaytraceScene::prepForRender() {
    moveGameObjectsToHostVectors();
    createBuffers();
}

void RaytraceScene::updateScene() {
    if (!buffersCreated) {
        prepForRender();
        return;
    }
    // the reference re-flattens and re-uploads every frame (RaytraceScene.cpp:78-113); here the
    // upload and device build run only when the flattened data changed
    auto oldModels = models;
    auto oldTris = triangles;
    auto oldSph = spheres;
    auto oldMats = materials;
    moveGameObjectsToHostVectors();
    auto same = [](const auto &a, const auto &b) {
        return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), sizeof(a[0]) * a.size()) == 0);
    };
    if (dirty || !same(oldModels, models) || !same(oldTris, triangles) || !same(oldSph, spheres) ||
        !same(oldMats, materials))
        createBuffers();
}
