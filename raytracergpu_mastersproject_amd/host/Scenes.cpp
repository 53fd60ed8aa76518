// Scenes.cpp -- the reference's five scene descriptions (RaytracerGPU_MastersProject/Scenes.cpp:37-415):
// same objects, transforms, materials, depths, FOVs and samples per pixel.  Meshes are read from
// the model directory (quad.obj and cube.obj ship with this package; monkey.obj and
// smooth_vase.obj must be supplied by the user, as in the reference, whose .obj files are not
// part of its repository).
#include <random>

#include "rtgpu/Scene.hpp"

using SceneTypes::MaterialType;
using SceneTypes::GPU::Material;
using rtgpu::vec3;

namespace {

// randomColor/randomDouble (Scenes.cpp:26-35); seeded for reproducibility instead of std::random_device
struct SceneRandom {
    std::mt19937 gen{0};
    std::uniform_real_distribution<double> dist{0.0, 1.0};
    f32 real(f32 lo = 0.0f, f32 hi = 1.0f) { return lo + (hi - lo) * (f32)dist(gen); }
    vec3 color(f32 lo = 0.0f, f32 hi = 1.0f) { return vec3{real(lo, hi), real(lo, hi), real(lo, hi)}; }
};

}  // namespace

void randomSpheres(std::unique_ptr<RaytraceScene> &scene) {
    SceneRandom rnd;
    auto groundMaterial = Material(vec3{0.5f, 0.5f, 0.5f}, MaterialType::DIFFUSE);
    auto floor = GameObject::createGameObject();
    floor.setModel(loadModel(1000.0f, groundMaterial), false);
    floor.transform.translation = {0.0f, -1000.0f, 0.0f};
    scene->addGameObject(std::move(floor));

    auto mat1 = Material({1.0f, 1.0f, 1.0f}, MaterialType::LIGHT);
    auto mat2 = Material(rnd.color(), MaterialType::DIFFUSE);
    auto mat3 = Material(rnd.color(), MaterialType::METALLIC);
    auto mat4 = Material(rnd.color(), MaterialType::DIFFUSE);
    struct S { vec3 t; f32 r; Material m; };
    const S objs[] = {{{0.0f, 1.05f, 3.5f}, 1.0f, mat1}, {{-4, 1, 4}, 2.0f, mat2}, {{4, 1, 4}, 3.0f, mat3},
                      {{0, 1, -4}, 1.0f, mat4},         {{4, 1, -4}, 2.0f, mat4}, {{-4, 1, -4}, 3.0f, mat4},
                      {{4, 5, 4}, 1.0f, mat2},          {{-4, 5, 4}, 1.0f, mat2}};
    for (const S &s : objs) {
        auto o = GameObject::createGameObject();
        o.transform.translation = s.t;
        o.setModel(loadModel(s.r, s.m), false);
        scene->addGameObject(std::move(o));
    }
    auto triObj = GameObject::createGameObject();
    triObj.setModel(loadModel("models/quad.obj", vec3{0.3f, 0.5f, 0.7f}), true);
    triObj.transform.translation = {0.0f, 1.0f, 5.0f};
    triObj.transform.rotation = {4.0f, 0.0f, 0.0f};
    triObj.transform.scale = {5.0f, 1.0f, 3.0f};
    scene->addGameObject(std::move(triObj));

    scene->setMaxRaytraceDepth(5);
    scene->getCamera().setVerticalFOV(90.0f);
    scene->prepForRender();
}

void cornellMixedScene(std::unique_ptr<RaytraceScene> &scene) {
    std::shared_ptr<RTModel> quadRed = loadModel("models/quad.obj", vec3{1.0f, 0.1f, 0.1f});
    std::shared_ptr<RTModel> quadGreen = loadModel("models/quad.obj", vec3{0.1f, 1.0f, 0.1f});
    std::shared_ptr<RTModel> quadBlue = loadModel("models/quad.obj", vec3{0.1f, 0.1f, 1.0f});
    std::shared_ptr<RTModel> quadWhite = loadModel("models/quad.obj", vec3{1.0f, 1.0f, 1.0f});
    std::shared_ptr<RTModel> cubeWhite = loadModel("models/cube.obj", vec3{1.0f, 1.0f, 1.0f});
    Material sphereMat(vec3{1.0f, 1.0f, 1.0f}, MaterialType::LIGHT);
    const f32 pi = rtgpu::pi();

    auto add = [&](std::shared_ptr<RTModel> m, bool tri, vec3 t, vec3 s, vec3 r) {
        auto o = GameObject::createGameObject();
        o.setModel(m, tri);
        o.transform.translation = t;
        o.transform.scale = s;
        o.transform.rotation = r;
        scene->addGameObject(std::move(o));
    };
    add(quadGreen, true, {0, -4.0f, 10.0f}, {10.0f, 1.0f, 3.0f}, {0, 0, 0});                  // floor
    add(quadWhite, true, {-10.0f, 1.0f, 10.0f}, {5.0f, 1.0f, 3.0f}, {0.0f, 0.0f, pi / 2});     // left wall
    add(quadWhite, true, {10.0f, 1.0f, 10.0f}, {5.0f, 1.0f, 3.0f}, {0.0f, 0.0f, -pi / 2});     // right wall
    add(quadRed, true, {0.0f, 6.0f, 10.0f}, {10.0f, 1.0f, 3.0f}, {0, 0, 0});                   // roof
    add(quadBlue, true, {0.0f, 1.0f, 13.0f}, {10.0f, 1.0f, 6.0f}, {pi / 2, 0.0f, 0.0f});       // back wall
    // the vase object is created but not added (Scenes.cpp:174-191)
    add(std::shared_ptr<RTModel>(loadModel(0.5f, sphereMat)), false, {0.0f, 6.0f, 10.0f}, {1, 1, 1}, {0, 0, 0});
    add(cubeWhite, true, {-5.0f, -2.0f, 10.0f}, {2.0f, 2.0f, 2.0f}, {0, pi / 3, 0});
    add(cubeWhite, true, {4.0f, -2.0f, 10.0f}, {2.0f, 2.0f, 2.0f}, {0, -pi / 4, 0});
    auto mat = Material(vec3{0.3f, 0.5f, 0.7f}, MaterialType::DIFFUSE);
    add(std::shared_ptr<RTModel>(loadModel(2.0f, mat)), false, {-0.5f, 0.0f, 12.0f}, {1, 1, 1}, {0, 0, 0});

    scene->setMaxRaytraceDepth(100);
    scene->getCamera().setVerticalFOV(80.0f);
    scene->prepForRender();
}

namespace {

// the 275-scaled Cornell room pieces shared by cornellBoxScene, simpleScene and complexScene
GameObject roomPiece(std::shared_ptr<RTModel> model, vec3 translation, vec3 rotation,
                     vec3 scale = vec3{275.0f, 1.0f, 275.0f}) {
    GameObject o = GameObject::createGameObject();
    o.setModel(std::move(model), true);
    o.transform.translation = translation;
    o.transform.scale = scale;
    o.transform.rotation = rotation;
    return o;
}

}  // namespace

void cornellBoxScene(std::unique_ptr<RaytraceScene> &scene) {
    Material redDiff(vec3{0.65f, 0.05f, 0.05f}, MaterialType::DIFFUSE);
    Material whiteDiff(vec3{0.73f, 0.73f, 0.73f}, MaterialType::DIFFUSE);
    Material greenDiff(vec3{0.12f, 0.45f, 0.15f}, MaterialType::DIFFUSE);
    Material light(vec3{15.0f, 15.0f, 15.0f}, MaterialType::LIGHT);
    std::shared_ptr<RTModel> quadRed = loadModel("models/quad.obj", redDiff);
    std::shared_ptr<RTModel> quadGreen = loadModel("models/quad.obj", greenDiff);
    std::shared_ptr<RTModel> quadWhite = loadModel("models/quad.obj", whiteDiff);
    std::shared_ptr<RTModel> quadLight = loadModel("models/quad.obj", light);
    std::shared_ptr<RTModel> cubeWhite = loadModel("models/cube.obj", whiteDiff);
    const f32 pi = rtgpu::pi();

    scene->addGameObject(roomPiece(quadGreen, {550.0f, 275.0f, 275.0f}, {0, 0, pi / 2}));   // left wall
    scene->addGameObject(roomPiece(quadRed, {0.0f, 275.0f, 275.0f}, {0, 0, pi / 2}));       // right wall
    scene->addGameObject(roomPiece(quadLight, {275.0f, 549.0f, 300.0f}, {0, 0, 0}, {65.0f, 1.0f, 50.0f}));
    scene->addGameObject(roomPiece(quadWhite, {275.0f, 0.0f, 275.0f}, {0, 0, 0}));          // bottom
    scene->addGameObject(roomPiece(quadWhite, {275.0f, 550.0f, 275.0f}, {0, 0, 0}));        // top
    scene->addGameObject(roomPiece(quadWhite, {275.0f, 275.0f, 550.0f}, {pi / 2, 0, 0}));   // back
    scene->addGameObject(roomPiece(cubeWhite, {350.0f, 160.0f, 395.0f}, {0, rtgpu::radians(15.0f), 0},
                                   {80.0f, 160.0f, 80.0f}));
    scene->addGameObject(roomPiece(cubeWhite, {180.0f, 160.0f, 175.0f}, {0, rtgpu::radians(-18.0f), 0},
                                   {80.0f, 80.0f, 80.0f}));
    GameObject dummySphere = GameObject::createGameObject();  // the reference needs >= 1 sphere (A16)
    dummySphere.setModel(std::shared_ptr<RTModel>(loadModel(1.0f, whiteDiff)), false);
    dummySphere.transform.translation = {-200.0f, 215.0f, -50.0f};
    scene->addGameObject(std::move(dummySphere));

    scene->setRaysPerPixel(128);
    scene->setMaxRaytraceDepth(25);
    scene->getCamera().setVerticalFOV(40.0f);
    scene->prepForRender();
}

void simpleScene(std::unique_ptr<RaytraceScene> &scene) {
    Material light(vec3{15.0f, 15.0f, 15.0f}, MaterialType::LIGHT);
    Material whiteDiff(vec3{0.73f, 0.73f, 0.73f}, MaterialType::DIFFUSE);
    Material redDiff(vec3{0.65f, 0.05f, 0.05f}, MaterialType::DIFFUSE);
    std::shared_ptr<RTModel> quadLight = loadModel("models/quad.obj", light);
    std::shared_ptr<RTModel> quadWhite = loadModel("models/quad.obj", whiteDiff);
    std::shared_ptr<RTModel> quadRed = loadModel("models/quad.obj", redDiff);
    const f32 pi = rtgpu::pi();

    scene->addGameObject(roomPiece(quadLight, {275.0f, 549.0f, 300.0f}, {0, 0, 0}, {65.0f, 1.0f, 50.0f}));
    scene->addGameObject(roomPiece(quadWhite, {275.0f, 0.0f, 275.0f}, {0, 0, 0}));
    scene->addGameObject(roomPiece(quadRed, {275.0f, 275.0f, 550.0f}, {pi / 2, 0, 0}));
    GameObject dummySphere = GameObject::createGameObject();
    dummySphere.setModel(std::shared_ptr<RTModel>(loadModel(200.0f, whiteDiff)), false);
    dummySphere.transform.translation = {100.0f, 445.0f, 215.0f};
    scene->addGameObject(std::move(dummySphere));

    scene->setRaysPerPixel(16);
    scene->setMaxRaytraceDepth(8);
    scene->getCamera().setVerticalFOV(40.0f);
    scene->prepForRender();
}

void complexScene(std::unique_ptr<RaytraceScene> &scene) {
    Material light(vec3{15.0f, 15.0f, 15.0f}, MaterialType::LIGHT);
    Material whiteDiff(vec3{0.33f, 0.73f, 0.33f}, MaterialType::DIFFUSE);
    Material redDiff(vec3{0.65f, 0.05f, 0.05f}, MaterialType::DIFFUSE);
    Material blueDiff(vec3{0.12f, 0.15f, 0.45f}, MaterialType::DIFFUSE);
    std::shared_ptr<RTModel> quadLight = loadModel("models/quad.obj", light);
    std::shared_ptr<RTModel> quadWhite = loadModel("models/quad.obj", whiteDiff);
    std::shared_ptr<RTModel> monkeyBlue = loadModel("models/monkey.obj", blueDiff);
    const f32 pi = rtgpu::pi();

    scene->addGameObject(roomPiece(quadLight, {275.0f, 549.0f, 300.0f}, {0, 0, 0}, {65.0f, 1.0f, 50.0f}));
    scene->addGameObject(roomPiece(quadWhite, {275.0f, 0.0f, 275.0f}, {0, 0, 0}));
    scene->addGameObject(roomPiece(quadWhite, {275.0f, 275.0f, 550.0f}, {pi / 2, 0, 0}));
    // monkey1 (red) is created but not added (Scenes.cpp:392-407)
    scene->addGameObject(roomPiece(monkeyBlue, {175.0f, 125.0f, 275.0f},
                                   {rtgpu::radians(-36.0f), rtgpu::radians(180.0f), rtgpu::radians(21.0f)},
                                   {100.0f, 100.0f, 100.0f}));
    GameObject dummySphere = GameObject::createGameObject();
    dummySphere.setModel(std::shared_ptr<RTModel>(loadModel(40.0f, redDiff)), false);
    dummySphere.transform.translation = {100.0f, 215.0f, 50.0f};
    scene->addGameObject(std::move(dummySphere));

    scene->setRaysPerPixel(8);
    scene->setMaxRaytraceDepth(8);
    scene->getCamera().setVerticalFOV(40.0f);
    scene->prepForRender();
}
