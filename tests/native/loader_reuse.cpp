// One SL_SceneFileLoader loads two OBJ files in turn; the second has no
// mtllib, so its faces take the default material.  Every mesh of each load
// must name a material of that load's own table.  Prints "ok" or the first
// bad mesh and exits non-zero.
#include <cstdio>

#include "softlight_amd/SL_SceneFileLoader.hpp"

static int check(const SL_SceneFileData& d, const char* what)
{
    for (size_t i = 0; i < d.meshes.size(); ++i)
    {
        if (d.meshes[i].material >= d.materials.size())
        {
            std::printf("%s: mesh %zu names material %u of %zu\n", what, i, d.meshes[i].material, d.materials.size());
            return 1;
        }
    }
    return 0;
}

int main(int argc, char** argv)
{
    if (argc < 3)
    {
        std::fprintf(stderr, "usage: %s first.obj second.obj\n", argv[0]);
        return 2;
    }
    SL_SceneFileLoader loader;
    if (loader.load(argv[1]) != SL_SCENE_LOAD_OK)
    {
        std::printf("first load: %s\n", loader.error().c_str());
        return 3;
    }
    if (check(loader.data(), "first")) return 1;
    const size_t firstMaterials = loader.data().materials.size();
    if (loader.load(argv[2]) != SL_SCENE_LOAD_OK)
    {
        std::printf("second load: %s\n", loader.error().c_str());
        return 3;
    }
    if (check(loader.data(), "second")) return 1;
    std::printf("ok %zu %zu\n", firstMaterials, loader.data().materials.size());
    return 0;
}
