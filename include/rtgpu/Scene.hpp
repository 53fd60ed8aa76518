// Scene.hpp -- the reference's scene API on top of the C ABI:
//   TransformComponent (VulkanWrapper/components/TransformComponent.{hpp,cpp}),
//   RTModel / loadModel (VulkanWrapper/RTModel.{hpp,cpp}),
//   GameObject (VulkanWrapper/GameObject.{hpp,cpp}),
//   RaytraceScene (VulkanWrapper/RaytraceScene.{hpp,cpp}),
//   Scenes (Scenes.{hpp,cpp}).
// A `Device&` in the reference becomes an `rtgpu::Context&` (one rt_context on one GPU).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include "SceneTypes.hpp"

namespace rtgpu {

// Owns one rt_context; throws std::runtime_error with rt_last_error() on failure.
class Context {
    rt_context *ctx_ = nullptr;

public:
    explicit Context(int device = 0);
    // host-only context: scenes flatten on the CPU but nothing is uploaded (tests without a GPU)
    explicit Context(std::nullptr_t) {}
    ~Context();
    bool hostOnly() const { return ctx_ == nullptr; }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    rt_context *get() const { return ctx_; }
    void check(int status, const char *what) const;
};

}  // namespace rtgpu

// ---------------- TransformComponent (TransformComponent.cpp:9-45) ----------------
struct TransformComponent {
    rtgpu::vec3 translation{};
    rtgpu::vec3 scale{1.0f, 1.0f, 1.0f};
    rtgpu::vec3 rotation{};  // radians, Tait-Bryan Y, X, Z
    TransformComponent() = default;
    TransformComponent(rtgpu::vec3 t, rtgpu::vec3 s, rtgpu::vec3 r) : translation(t), scale(s), rotation(r) {}
    // translate * Ry * Rx * Rz * scale, column-major
    rtgpu::mat4 mat4() const;
    rtgpu::mat3 normalMatrix() const;
};

// ---------------- RTModel (RTModel.hpp:56-85) ----------------
class RTModel_Triangles {
    std::vector<SceneTypes::CPU::Triangle> triangles;
    SceneTypes::GPU::Material material;

public:
    RTModel_Triangles(std::vector<SceneTypes::CPU::Triangle> &&t, SceneTypes::GPU::Material mat);
    const std::vector<SceneTypes::CPU::Triangle> &getTriangles() const;
    SceneTypes::GPU::Material getMaterialType() const;
};

class RTModel_Sphere {
    rtgpu::vec3 center;
    f32 radius;
    SceneTypes::GPU::Material material;

public:
    RTModel_Sphere(f32 radius, SceneTypes::GPU::Material mat);  // centred at the origin
    const rtgpu::vec3 &getCenter() const;
    f32 getRadius() const;
    SceneTypes::GPU::Material getMaterialType() const;
};

using RTModel = std::variant<RTModel_Triangles, RTModel_Sphere>;

// OBJ meshes are resolved against the model directory (set by setModelDirectory, else the
// RTGPU_MODEL_DIR environment variable, else the package's models/ directory).
void setModelDirectory(const std::string &dir);
std::string modelPath(const std::string &filepath);
// Triangulated OBJ: faces with more than three vertices become triangle fans (v0, vi, vi+1).
std::vector<SceneTypes::CPU::Triangle> loadObjTriangles(const std::string &filepath);
std::unique_ptr<RTModel> loadModel(const std::string &filepath, rtgpu::vec3 color);
std::unique_ptr<RTModel> loadModel(const std::string &filepath, SceneTypes::GPU::Material mat);
std::unique_ptr<RTModel> loadModel(f32 radius, SceneTypes::GPU::Material mat);

// ---------------- GameObject (GameObject.hpp:249-301) ----------------
using GameObjectId = u32;

struct GameObject {
protected:
    GameObjectId id;
    std::shared_ptr<RTModel> model{};
    explicit GameObject(GameObjectId objId) : id(objId) {}

public:
    static GameObject createGameObject();
    void setModel(std::shared_ptr<RTModel> model, bool triangular);
    std::shared_ptr<RTModel> getModel() const;
    GameObjectId getId() const;
    GameObject(const GameObject &);  // copies get a fresh id (GameObject.cpp:405-423)
    GameObject &operator=(const GameObject &) = delete;
    GameObject(GameObject &&) = default;
    GameObject &operator=(GameObject &&) = default;

    rtgpu::vec3 color{};
    TransformComponent transform{};
    bool isTriangleModel = false;
    bool isSphereModel = false;
};

// Camera state the ray tracer reads: the reference only forwards the vertical FOV
// (RaytracerBVH.hpp:344); the pose lives here instead of being hard-coded in the UBO.
struct CameraGameObject {
    f32 verticalFOV = 50.0f;
    rtgpu::vec3 position{275.0f, 275.0f, -800.0f};  // RaytracerBVH.hpp:341-343 (A8)
    rtgpu::vec3 lookAt{275.0f, 275.0f, 0.0f};
    rtgpu::vec3 up{0.0f, 1.0f, 0.0f};
    f32 getVerticalFOV() const { return verticalFOV; }
    void setVerticalFOV(f32 fovy) { verticalFOV = fovy; }
};

// ---------------- RaytraceScene (RaytraceScene.hpp:155-264) ----------------
class RaytraceScene {
    rtgpu::Context &context;
    CameraGameObject camera;
    std::vector<GameObject> gameObjects;
    std::vector<SceneTypes::GPU::Model> models;
    std::vector<SceneTypes::GPU::Triangle> triangles;
    std::vector<SceneTypes::GPU::Sphere> spheres;
    std::vector<SceneTypes::GPU::Material> materials;
    u32 maxRaytraceDepth = 8;
    u32 raysPerPixel = 1;
    bool buffersCreated = false;
    bool dirty = true;
    uint32_t sceneId = 0;
    rt_build_options buildOptions{0, 4, 1, 0};

    void moveGameObjectsToHostVectors();
    void createBuffers();
    void releaseBuffers();

public:
    explicit RaytraceScene(rtgpu::Context &context);
    ~RaytraceScene();
    RaytraceScene(const RaytraceScene &) = delete;
    RaytraceScene &operator=(const RaytraceScene &) = delete;

    void setMaxRaytraceDepth(u32 depth);
    void setRaysPerPixel(u32 rpp);
    u32 getMaxRaytraceDepth() const;
    u32 getRaysPerPixel() const;
    CameraGameObject &getCamera();
    void setBuildOptions(const rt_build_options &opts);

    void addGameObject(GameObject &&gameObject);
    GameObject &getGameObject(GameObjectId id);
    bool removeGameObject(GameObjectId id);

    // flatten game objects (insertion order: per object one model matrix and one material,
    // its triangles or its sphere) and upload + build the BVH on the device
    void prepForRender();
    // re-flatten and re-upload when objects changed (the reference re-uploads every frame)
    void updateScene();
    void markDirty() { dirty = true; }

    uint32_t getSceneId() const { return sceneId; }
    rtgpu::Context &getContext() { return context; }
    const std::vector<SceneTypes::GPU::Model> &getModels() const { return models; }
    const std::vector<SceneTypes::GPU::Triangle> &getTriangles() const { return triangles; }
    const std::vector<SceneTypes::GPU::Sphere> &getSpheres() const { return spheres; }
    const std::vector<SceneTypes::GPU::Material> &getMaterials() const { return materials; }
    u32 getModelCount() const { return (u32)models.size(); }
    u32 getTriangleCount() const { return (u32)triangles.size(); }
    u32 getSphereCount() const { return (u32)spheres.size(); }
    u32 getMaterialCount() const { return (u32)materials.size(); }
    // host-only flattening (no upload), for inspection and tests
    void flattenOnly() { moveGameObjectsToHostVectors(); }
};

// ---------------- Scenes (Scenes.hpp:9-17) ----------------
void randomSpheres(std::unique_ptr<RaytraceScene> &scene);
void cornellMixedScene(std::unique_ptr<RaytraceScene> &scene);
void cornellBoxScene(std::unique_ptr<RaytraceScene> &scene);
void simpleScene(std::unique_ptr<RaytraceScene> &scene);
void complexScene(std::unique_ptr<RaytraceScene> &scene);
