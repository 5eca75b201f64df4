// crt.hpp -- C++ host mirror of CRender's scene / renderer classes over the crt C ABI.
//
// Header-only.  Class and method names follow the reference so a caller of cr::scene and
// cr::renderer switches by changing the namespace and the constructor arguments:
//   crt::scene     ~ cr::scene    (src/render/scene.h:19-53)
//   crt::renderer  ~ cr::renderer (src/render/renderer.h:24-105)
// Errors throw crt::error (the C ABI itself never exits or throws).
#pragma once

#include <crt.h>

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace crt {

class error : public std::runtime_error {
public:
    error(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
    int code;
};

inline void check(int rc)
{
    if (rc != CRT_OK) throw error(rc, crt_last_error());
}

class context {
public:
    explicit context(int device = 0, uint32_t bvh_width = 0, uint32_t flags = 0)
    {
        crt_options o;
        crt_options_default(&o);
        o.device    = device;
        o.bvh_width = bvh_width;
        o.flags     = flags;
        check(crt_context_create(&o, &h_));
    }
    ~context() { crt_context_destroy(h_); }
    context(const context &) = delete;
    context &operator=(const context &) = delete;
    crt_context *get() const { return h_; }
    std::string device_name() const
    {
        char buf[256];
        check(crt_context_device_name(h_, buf, sizeof buf));
        return buf;
    }

private:
    crt_context *h_ = nullptr;
};

// asset_loader::model_data (asset_loader.h:16-30) with flat arrays
struct model_data {
    std::string               name;
    std::vector<float>        vertices;        // xyz per vertex
    std::vector<crt_material> materials;
    std::vector<float>        texture_coords;  // uv per coordinate
    std::vector<crt_image>    textures;        // caller keeps the texel memory alive until add_model returns
    std::vector<uint32_t>     vertex_indices;  // 3 per triangle
    std::vector<uint32_t>     material_indices;
    std::vector<uint32_t>     texture_indices; // 3 per triangle or empty
};

// cr::ray::intersection_record (ray.h:12-20); the material is an index into the scene's
// material table instead of a pointer.
struct intersection_record {
    uint32_t prim_id  = 0xFFFFFFFFu;
    float    distance = __builtin_huge_valf();
    uint32_t material = 0xFFFFFFFFu;
    float    uv[2]    = {0, 0};
    float    normal[3] = {0, 0, 0};
    float    intersection_point[3] = {0, 0, 0};
    uint32_t geom_id  = 0xFFFFFFFFu;   // model of the hit
};

class scene {
public:
    explicit scene(context &ctx) { check(crt_scene_create(ctx.get(), &h_)); }
    ~scene() { crt_scene_destroy(h_); }
    scene(const scene &) = delete;
    scene &operator=(const scene &) = delete;
    crt_scene *get() const { return h_; }

    // scene::add_model (scene.cpp:31-34)
    uint32_t add_model(const model_data &m)
    {
        crt_model_data md{};
        md.name               = m.name.c_str();
        md.vertices           = m.vertices.data();
        md.num_vertices       = (uint32_t)(m.vertices.size() / 3);
        md.texture_coords     = m.texture_coords.empty() ? nullptr : m.texture_coords.data();
        md.num_texture_coords = (uint32_t)(m.texture_coords.size() / 2);
        md.materials          = m.materials.data();
        md.num_materials      = (uint32_t)m.materials.size();
        md.textures           = m.textures.empty() ? nullptr : m.textures.data();
        md.num_textures       = (uint32_t)m.textures.size();
        md.vertex_indices     = m.vertex_indices.data();
        md.texture_indices    = m.texture_indices.empty() ? nullptr : m.texture_indices.data();
        md.material_indices   = m.material_indices.data();
        md.num_triangles      = (uint32_t)m.material_indices.size();
        uint32_t id = 0;
        check(crt_scene_add_model(h_, &md, &id));
        return id;
    }
    // asset_loader::load_model + scene::add_model (asset_loader.cpp:182-303, scene.cpp:31-34)
    uint32_t load_model(const std::string &file, const std::string &folder = "")
    {
        crt_obj *o = nullptr;
        check(crt_obj_load(file.c_str(), folder.empty() ? nullptr : folder.c_str(), &o));
        crt_model_data md{};
        uint32_t       id = 0;
        int            rc = crt_obj_model(o, &md);
        if (rc == CRT_OK) rc = crt_scene_add_model(h_, &md, &id);
        crt_obj_free(o);
        check(rc);
        return id;
    }
    // entity::instances (components.h:18-21): column-major 4x4 object->world transforms
    void set_instances(uint32_t model_id, const std::vector<float> &transforms)
    {
        check(crt_scene_set_instances(h_, model_id, transforms.data(), (uint32_t)(transforms.size() / 16)));
    }
    void set_materials(uint32_t model_id, const std::vector<crt_material> &m)
    {
        check(crt_scene_set_materials(h_, model_id, m.data(), (uint32_t)m.size()));
    }
    // scene::set_skybox (scene.cpp:36-60); rgba is copied
    void set_skybox(const float *rgba, uint32_t w, uint32_t h)
    {
        crt_image im{rgba, w, h};
        check(crt_scene_set_skybox(h_, rgba ? &im : nullptr));
    }
    void set_skybox_rotation(float x, float y) { check(crt_scene_set_skybox_rotation(h_, x, y)); }
    void set_sun_enabled(bool v) { check(crt_scene_set_sun_enabled(h_, v ? 1 : 0)); }
    void set_sun(const crt_sun &s) { check(crt_scene_set_sun(h_, &s)); }
    void set_camera(const crt_camera &c) { check(crt_scene_set_camera(h_, &c)); }
    void commit() { check(crt_scene_commit(h_)); }
    crt_bvh_info bvh_info() const
    {
        crt_bvh_info i;
        check(crt_scene_bvh_info(h_, &i));
        return i;
    }

    // scene::cast_ray (scene.cpp:79-98) for one ray; batches go through cast_rays() or, for
    // Embree-style hit records, intersect()
    intersection_record cast_ray(const float origin[3], const float direction[3])
    {
        crt_rays r{&origin[0], &origin[1], &origin[2], &direction[0], &direction[1], &direction[2], nullptr, nullptr};
        intersection_record out;
        crt_records         rec{&out.distance,
                        &out.prim_id,
                        &out.geom_id,
                        &out.material,
                        &out.uv[0],
                        &out.uv[1],
                        &out.normal[0],
                        &out.normal[1],
                        &out.normal[2],
                        &out.intersection_point[0],
                        &out.intersection_point[1],
                        &out.intersection_point[2]};
        check(crt_cast_rays(h_, &r, 1, &rec, CRT_MEM_HOST));
        return out;
    }
    void cast_rays(const crt_rays &rays, uint32_t n, crt_records &records, int memory = CRT_MEM_HOST)
    {
        check(crt_cast_rays(h_, &rays, n, &records, memory));
    }
    void intersect(const crt_rays &rays, uint32_t n, crt_hits &hits, int memory = CRT_MEM_HOST)
    {
        check(crt_intersect(h_, &rays, n, &hits, memory));
    }

private:
    crt_scene *h_ = nullptr;
};

// RGBA FP32 image returned by current_progress() etc. (cr::image, image.h:13-151)
struct image {
    uint32_t           width = 0, height = 0;
    std::vector<float> data;   // (x + y * width) * 4
};

class renderer {
public:
    // renderer.cpp:106-145: the reference starts its management thread immediately
    renderer(context &ctx, scene &sc, uint32_t res_x, uint32_t res_y, uint32_t bounces, bool autostart = true)
    {
        check(crt_renderer_create(ctx.get(), sc.get(), res_x, res_y, bounces, &h_));
        if (autostart) start();
    }
    ~renderer() { crt_renderer_destroy(h_); }
    renderer(const renderer &) = delete;
    renderer &operator=(const renderer &) = delete;

    // renderer.cpp:154-170: true if the renderer was paused (accumulation cleared)
    bool start()
    {
        const bool was = !running();
        check(crt_renderer_start(h_));
        return was;
    }
    // renderer.cpp:172-183: true if it was running; returns after the current pass
    bool pause()
    {
        const bool was = running();
        check(crt_renderer_pause(h_));
        return was;
    }
    // renderer.cpp:185-192
    void update(const std::function<void()> &fn)
    {
        pause();
        fn();
        start();
    }
    void set_resolution(int x, int y) { check(crt_renderer_set_resolution(h_, (uint32_t)x, (uint32_t)y)); }
    void set_max_bounces(int b) { check(crt_renderer_set_max_bounces(h_, (uint32_t)b)); }
    void set_target_spp(uint64_t t) { check(crt_renderer_set_target_spp(h_, t)); }
    void set_options(const crt_render_options &o) { check(crt_renderer_set_options(h_, &o)); }
    void wait() { check(crt_renderer_wait(h_)); }
    void render(uint32_t passes) { check(crt_renderer_render(h_, passes)); }

    struct renderer_stats {   // renderer.h:48-54
        uint64_t rays_per_second;
        uint64_t samples_per_second;
        uint64_t total_rays;
        double   running_time;
    };
    renderer_stats current_stats() const
    {
        crt_stats s;
        check(crt_renderer_current_stats(h_, &s));
        return {s.rays_per_second, s.samples_per_second, s.total_rays, s.running_time};
    }
    crt_stats current_stats_ext() const
    {
        crt_stats s;
        check(crt_renderer_current_stats(h_, &s));
        return s;
    }
    uint64_t current_sample_count() const
    {
        uint64_t n = 0;
        check(crt_renderer_current_sample_count(h_, &n));
        return n;
    }
    std::pair<int, int> current_resolution() const
    {
        uint32_t x = 0, y = 0;
        check(crt_renderer_current_resolution(h_, &x, &y));
        return {(int)x, (int)y};
    }
    image current_progress() const { return read(CRT_BUFFER_PROGRESS); }
    image current_normals() const { return read(CRT_BUFFER_NORMALS); }
    image current_albedos() const { return read(CRT_BUFFER_ALBEDOS); }
    image current_depths() const { return read(CRT_BUFFER_DEPTHS); }
    std::vector<float> raw_buffer() const
    {
        auto r = current_resolution();
        std::vector<float> out((size_t)r.first * r.second * 3);
        check(crt_renderer_read(h_, CRT_BUFFER_RAW, out.data()));
        return out;
    }

private:
    bool running() const
    {
        int r = 0;
        check(crt_renderer_is_running(h_, &r));
        return r != 0;
    }
    image read(int which) const
    {
        auto  r = current_resolution();
        image im;
        im.width  = (uint32_t)r.first;
        im.height = (uint32_t)r.second;
        im.data.resize((size_t)im.width * im.height * 4);
        check(crt_renderer_read(h_, which, im.data.data()));
        return im;
    }
    crt_renderer *h_ = nullptr;
};

// asset_loader::export_framebuffer (asset_loader.cpp:348-377): the caller names the file
inline void export_framebuffer(const image &im, const std::string &path, int format)
{
    check(crt_image_write(path.c_str(), format, im.data.data(), im.width, im.height));
}

// Renders many jobs together (crt_render_batch); returns each job's latency in ms.
inline std::vector<double> render_batch(context &ctx, std::vector<crt_job> &jobs, crt_batch_stats *stats = nullptr)
{
    std::vector<double> ms(jobs.size());
    crt_batch_options   o;
    crt_batch_options_default(&o);
    check(crt_render_batch(ctx.get(), jobs.data(), (uint32_t)jobs.size(), &o, ms.data(), stats));
    return ms;
}

}  // namespace crt
