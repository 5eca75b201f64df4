// render_box.cpp -- C++ host example: the reference's UI flow on the crt library.
//
// Builds a small Cornell-style box in memory (the shape of asset_loader::model_data),
// loads it into a crt::scene, then drives crt::renderer as CRender's UI drives
// cr::renderer: construct (starts rendering), update() a setting (pause, mutate,
// start), wait for the target spp, read current_progress() and current_stats().
// Writes the display buffer as a binary PPM when a path is given.
//
// usage: render_box [spp] [out.ppm]
#include <crt.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace {

void quad(crt::model_data &m, const float p[4][3], uint32_t mat)
{
    const uint32_t base = (uint32_t)(m.vertices.size() / 3);
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 3; ++k) m.vertices.push_back(p[i][k]);
    const uint32_t idx[6] = {0, 1, 2, 0, 2, 3};
    for (uint32_t i : idx) m.vertex_indices.push_back(base + i);
    m.material_indices.push_back(mat);
    m.material_indices.push_back(mat);
}

crt_material mat(float r, float g, float b, float emission = 0.0f)
{
    crt_material m;
    crt_material_default(&m);
    m.colour[0] = r;
    m.colour[1] = g;
    m.colour[2] = b;
    m.emission  = emission;
    return m;
}

}  // namespace

int main(int argc, char **argv)
{
    const uint64_t spp = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 16;
    try {
        crt::context ctx(0);
        std::printf("device: %s\n", ctx.device_name().c_str());
        crt::model_data m;
        m.name      = "box";
        m.materials = {mat(0.63f, 0.065f, 0.05f), mat(0.14f, 0.45f, 0.091f), mat(0.725f, 0.71f, 0.68f),
                       mat(1, 1, 1, 15.0f)};
        const float floor_[4][3] = {{-1, 0, -1}, {-1, 0, 1}, {1, 0, 1}, {1, 0, -1}};
        const float ceil_[4][3]  = {{-1, 2, -1}, {1, 2, -1}, {1, 2, 1}, {-1, 2, 1}};
        const float back_[4][3]  = {{-1, 0, 1}, {-1, 2, 1}, {1, 2, 1}, {1, 0, 1}};
        const float left_[4][3]  = {{-1, 0, -1}, {-1, 2, -1}, {-1, 2, 1}, {-1, 0, 1}};
        const float right_[4][3] = {{1, 0, -1}, {1, 0, 1}, {1, 2, 1}, {1, 2, -1}};
        const float light_[4][3] = {{-0.25f, 1.98f, -0.25f}, {0.25f, 1.98f, -0.25f}, {0.25f, 1.98f, 0.25f},
                                    {-0.25f, 1.98f, 0.25f}};
        quad(m, floor_, 2);
        quad(m, ceil_, 2);
        quad(m, back_, 2);
        quad(m, left_, 0);
        quad(m, right_, 1);
        quad(m, light_, 3);

        crt::scene sc(ctx);
        sc.add_model(m);
        sc.set_sun_enabled(false);
        crt_camera cam;
        crt_camera_default(&cam);
        cam.position[0] = 0.0f;
        cam.position[1] = 1.0f;
        cam.position[2] = -3.4f;
        cam.fov         = 40.0f;
        crt_camera_update(&cam);
        sc.set_camera(cam);
        sc.commit();
        const crt_bvh_info bi = sc.bvh_info();
        std::printf("bvh: width %u, %u nodes, depth %u, built in %.3f ms\n", bi.width, bi.node_count, bi.depth,
                    bi.build_ms);
        const float eye[3] = {0.0f, 1.0f, -3.4f}, ahead[3] = {0.0f, 0.0f, 2.0f};
        const crt::intersection_record hit = sc.cast_ray(eye, ahead);   // the back wall at z = 1
        std::printf("cast_ray: distance %.4f material %u normal %.1f %.1f %.1f point %.2f %.2f %.2f\n", hit.distance,
                    hit.material, hit.normal[0], hit.normal[1], hit.normal[2], hit.intersection_point[0],
                    hit.intersection_point[1], hit.intersection_point[2]);

        crt::renderer r(ctx, sc, 256, 256, 8);   // starts rendering, like cr::renderer
        r.update([&] {
            r.set_resolution(256, 256);
            r.set_target_spp(spp);
            crt_render_options o{CRT_MODE_COMPAT, 3u, 1u, 0u};
            r.set_options(o);
        });
        r.wait();
        const crt::image img = r.current_progress();
        const auto       st  = r.current_stats();
        double           sum = 0.0;
        for (size_t i = 0; i < img.data.size(); i += 4) sum += img.data[i] + img.data[i + 1] + img.data[i + 2];
        std::printf("samples %llu, rays/s (reference counting) %llu, total rays %llu, mean display %.5f\n",
                    (unsigned long long)r.current_sample_count(), (unsigned long long)st.rays_per_second,
                    (unsigned long long)st.total_rays, sum / (3.0 * img.width * img.height));
        if (argc > 2) {
            FILE *f = std::fopen(argv[2], "wb");
            if (!f) return 2;
            std::fprintf(f, "P6\n%u %u\n255\n", img.width, img.height);
            for (size_t i = 0; i < img.data.size(); i += 4)
                for (int c = 0; c < 3; ++c) std::fputc((int)std::lround(std::fmin(1.0f, img.data[i + c]) * 255.0f), f);
            std::fclose(f);
        }
    } catch (const crt::error &e) {
        std::fprintf(stderr, "crt error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
