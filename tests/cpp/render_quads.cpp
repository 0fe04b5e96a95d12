// Renders a floor under a small quad light and a gradient sky through
// btrc_hip::WavefrontPathTracer (include/btrc_hip.hpp) and writes the RGBA
// film (float32, row-major, 48 x 32) to argv[1]. On a library error it
// prints the message and exits with status 2.
#include "../../include/btrc_hip.hpp"

#include <cstdio>

// counts the Reporter calls of a render
struct CountingReporter : btrc_hip::Reporter
{
    int stages = 0, completed = 0, progress_calls = 0;
    float last = -1.0f;
    void new_stage(const char *) override { ++stages; }
    void complete_stage() override { ++completed; }
    void progress(float pct) override
    {
        ++progress_calls;
        last = pct;
    }
};

int main(int argc, char **argv)
{
    if(argc < 2)
    {
        std::fprintf(stderr, "usage: render_quads out.f32\n");
        return 1;
    }
    try
    {
        // triangles (a, b, c); geometric normals: floor +y, light -y
        const float floor_tris[18] = { -1, 0, -1, -1, 0, 1, 1, 0, 1, -1, 0, -1, 1, 0, 1, 1, 0, -1 };
        const float light_tris[18] = { -0.3f, 1.5f, -0.3f, 0.3f, 1.5f, 0.3f, -0.3f, 1.5f, 0.3f,
                                       -0.3f, 1.5f, -0.3f, 0.3f, 1.5f, -0.3f, 0.3f, 1.5f, 0.3f };
        btrc_geometry_desc geos[2] = {};
        geos[0].triangle_count = 2;
        geos[0].positions = floor_tris;
        geos[1].triangle_count = 2;
        geos[1].positions = light_tris;

        btrc_material_desc mats[2] = {};
        mats[0].type = BTRC_MATERIAL_DIFFUSE;
        mats[0].shadow_terminator = 1;
        mats[0].color[0] = 0.6f; mats[0].color[1] = 0.5f; mats[0].color[2] = 0.4f;
        mats[1].type = BTRC_MATERIAL_DIFFUSE;
        mats[1].shadow_terminator = 1;

        btrc_instance_desc insts[2] = {};
        for(int i = 0; i < 2; ++i)
        {
            insts[i].geometry = i;
            insts[i].material = i;
            for(int k = 0; k < 4; ++k)
                insts[i].local_to_world[5 * k] = insts[i].world_to_local[5 * k] = 1.0f;
            insts[i].inner_medium = insts[i].outer_medium = -1;
        }
        insts[1].has_light = 1;
        insts[1].light_intensity[0] = insts[1].light_intensity[1] = insts[1].light_intensity[2] = 4.0f;

        btrc_scene_desc scene = {};
        scene.geometry_count = 2;
        scene.geometries = geos;
        scene.material_count = 2;
        scene.materials = mats;
        scene.instance_count = 2;
        scene.instances = insts;
        scene.envir_light.type = BTRC_ENVIR_GRADIENT_SKY;
        scene.envir_light.up[1] = 1.0f;
        scene.envir_light.lower[0] = scene.envir_light.lower[1] = scene.envir_light.lower[2] = 0.1f;
        scene.envir_light.upper[0] = 0.3f; scene.envir_light.upper[1] = 0.4f; scene.envir_light.upper[2] = 0.6f;
        scene.accel = BTRC_ACCEL_FLAT;

        btrc_camera_desc cam = {};
        cam.eye[1] = 1.0f; cam.eye[2] = 3.0f;
        cam.dst[1] = 0.3f;
        cam.up[1] = 1.0f;
        cam.fov_y_deg = 45.0f;

        btrc_hip::Context ctx(0);
        btrc_hip::WavefrontPathTracer::Params p;
        p.spp = 4;
        p.min_depth = 3;
        p.max_depth = 5;
        p.state_count = 1 << 14;
        btrc_hip::WavefrontPathTracer tracer(ctx, p);
        tracer.set_scene(scene);
        tracer.set_camera(cam);
        tracer.set_film(48, 32);
        CountingReporter rep;
        tracer.set_reporter(&rep);
        const btrc_hip::RenderResult r = tracer.render();
        tracer.set_reporter(nullptr);
        const std::vector<float> toned = tracer.tonemapped(1.0f);
        float toned_max = 0.0f;
        for(size_t k = 0; k < toned.size(); ++k)
            toned_max = toned[k] > toned_max ? toned[k] : toned_max;

        std::FILE *f = std::fopen(argv[1], "wb");
        if(!f || std::fwrite(r.color.data(), sizeof(float), r.color.size(), f) != r.color.size())
        {
            std::fprintf(stderr, "cannot write %s\n", argv[1]);
            return 1;
        }
        std::fclose(f);
        std::printf("rendered %d x %d, %llu extension rays, %llu shadow rays\n", r.width, r.height,
                    (unsigned long long)r.stats.extension_rays, (unsigned long long)r.stats.shadow_rays);
        std::printf("reporter: %d stage, %d completed, %d progress calls, last %.1f; tone mapped max %.3f\n",
                    rep.stages, rep.completed, rep.progress_calls, rep.last, toned_max);
        return 0;
    }
    catch(const btrc_hip::BtrcException &e)
    {
        std::fprintf(stderr, "BtrcException (status %d): %s\n", e.status(), e.what());
        return 2;
    }
}
