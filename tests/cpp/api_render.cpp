// api_render.cpp - a program written against the C++ scene API
// (include/mpt/mpt.hpp) the way a program of the reference renderer is:
// it loads a .scene file, or builds a world from objects, and calls
// camera::render (camera.h:79-105). Used by tests/test_cpp_api.py.
//
//   api_render scene <file.scene> <width> <spp> <seed> <out.bin>
//       renders the file's world and lights with its camera at the given
//       width and spp; writes the radiance (h*w*3 doubles) to out.bin
//   api_render objects <out.bin>
//       renders a small world built from objects (spheres, a box() behind a
//       rotate/translate instance, a quad light in a light_list, a checker
//       floor) with save_every 2; prints the callback counts and writes the
//       radiance of the returned frame followed by the last callback frame
//   api_render lights <file.scene>
//       loads the file (its loader calls lights.build(world.bounding_box()),
//       scene.h:1248), builds the list again from world.bounding_box() as a
//       program would, and prints the world box, each light's
//       probability(k) and power(), the picks of a few xi, and the core's
//       probabilities for the uploaded scene (mpt_scene_light_probabilities)
//   api_render atmosphere <out.bin>
//       bakes mpt::atmosphere with the settings of scenes/sky_atmosphere.scene
//       (mie 1.5, intensity 3; sun at elevation 18, azimuth 35, intensity 4)
//       at 256 x 128 as its loader does; writes the pixels (h*w*3 doubles)
//       and then the sun's irradiance dimmed by sun_transmittance (3 doubles)
//   api_render denoise <file.scene> <out.bin>
//       renders the file at width 40, 4 spp with collect_features and runs
//       denoiser::run on the frame; writes radiance, albedo, normal, the
//       denoised radiance (h*w*3 doubles each) and the variance (h*w)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "mpt/mpt.hpp"

using namespace mpt;

static bool write_frame(const char* path, const std::vector<color>& a, const std::vector<color>* b) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    for (const auto& c : a) std::fwrite(c.v, sizeof(double), 3, f);
    if (b)
        for (const auto& c : *b) std::fwrite(c.v, sizeof(double), 3, f);
    return std::fclose(f) == 0;
}

int main(int argc, char** argv) {
    if (argc >= 7 && std::string(argv[1]) == "scene") {
        scene sc;
        std::string err;
        if (!load_scene(argv[2], sc, &err)) {
            std::fprintf(stderr, "%s\n", err.c_str());
            return 1;
        }
        sc.cam.image_width = std::atoi(argv[3]);
        sc.cam.samples_per_pixel = std::atoi(argv[4]);
        sc.cam.seed = std::strtoull(argv[5], nullptr, 10);
        try {
            frame f = sc.cam.render(sc.world, sc.lights);
            std::printf("height %d\n", sc.cam.height());
            return write_frame(argv[6], f.radiance, nullptr) ? 0 : 1;
        } catch (const std::exception& e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
    }
    if (argc >= 3 && std::string(argv[1]) == "atmosphere") {
        atmosphere_settings st;
        st.mie = 1.5;
        st.intensity = 3;
        atmosphere air(st);
        const double el = 18 * 3.1415926535897932385 / 180.0, az = 35 * 3.1415926535897932385 / 180.0;
        const vec3 sun_dir(std::cos(el) * std::sin(az), std::sin(el), -std::cos(el) * std::cos(az));
        const color e = 4.0 * color(1, 1, 1);
        std::vector<color> px = air.bake(256, 128, sun_dir, e);
        const color tr = air.sun_transmittance(sun_dir);
        px.push_back(color(e[0] * tr[0], e[1] * tr[1], e[2] * tr[2]));
        return write_frame(argv[2], px, nullptr) ? 0 : 1;
    }
    if (argc >= 3 && std::string(argv[1]) == "lights") {
        scene sc;
        std::string err;
        if (!load_scene(argv[2], sc, &err)) {
            std::fprintf(stderr, "%s\n", err.c_str());
            return 1;
        }
        try {
            const aabb box = sc.world.bounding_box();
            std::printf("box %.17g %.17g %.17g %.17g %.17g %.17g\n", box.x.min, box.x.max, box.y.min, box.y.max,
                        box.z.min, box.z.max);
            std::vector<double> loaded;
            for (size_t k = 0; k < sc.lights.objects.size(); k++) loaded.push_back(sc.lights.probability(k));
            sc.lights.build(box);
            for (size_t k = 0; k < sc.lights.objects.size(); k++)
                std::printf("light %zu %.17g %.17g %.17g\n", k, sc.lights.probability(k), loaded[k],
                            sc.lights.objects[k]->power());
            for (double xi : {0.0, 0.05, 0.1, 0.15, 0.5, 0.999999})
                std::printf("pick %.17g %zu\n", xi, sc.lights.objects.empty() ? size_t(0) : sc.lights.pick(xi));
            gpu_scene gs(sc.world, sc.lights, sc.cam.environment);
            std::vector<double> core(sc.lights.objects.size());
            if (mpt_scene_light_probabilities(gs.get(), core.data()) != MPT_OK) {
                std::fprintf(stderr, "%s\n", mpt_last_error());
                return 1;
            }
            for (size_t k = 0; k < core.size(); k++) std::printf("core %zu %.17g\n", k, core[k]);
            return 0;
        } catch (const std::exception& e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
    }
    if (argc >= 3 && std::string(argv[1]) == "objects") {
        hittable_list world;
        auto white = std::make_shared<lambertian>(color(0.73, 0.73, 0.73));
        auto checker = std::make_shared<checker_texture>(0.5, std::make_shared<solid_color>(color(0.2, 0.3, 0.1)),
                                                         std::make_shared<solid_color>(color(0.9, 0.9, 0.9)));
        world.add(std::make_shared<quad>(point3(-5, 0, -5), vec3(10, 0, 0), vec3(0, 0, 10),
                                         std::make_shared<lambertian>(checker)));
        world.add(std::make_shared<sphere>(point3(0, 1, 0), 1.0, std::make_shared<dielectric>(1.5)));
        world.add(std::make_shared<sphere>(point3(-2.2, 1, 0), 1.0, std::make_shared<metal>(color(0.8, 0.6, 0.4), 0.2)));
        auto b = box(point3(0, 0, 0), point3(1, 1.5, 1), white);
        transform x = transform::translate(vec3(1.6, 0, -0.5)).then_after(transform::rotate(25, vec3(0, 1, 0)));
        world.add(std::make_shared<instance>(b, x));
        auto light = std::make_shared<quad>(point3(-1, 4, -1), vec3(2, 0, 0), vec3(0, 0, 2),
                                            std::make_shared<diffuse_light>(color(6, 6, 6)));
        world.add(light);
        light_list lights;
        lights.add(light);
        lights.build(world.bounding_box());

        camera cam;
        cam.aspect_ratio = 4.0 / 3.0;
        cam.image_width = 48;
        cam.samples_per_pixel = 6;
        cam.max_depth = 8;
        cam.vfov = 40;
        cam.lookfrom = point3(0, 2.5, 7);
        cam.lookat = point3(0, 0.8, 0);
        cam.use_sky = true;
        cam.seed = 7;
        int calls = 0, last_done = 0;
        std::vector<color> last;
        try {
            frame f = cam.render(bvh(world), lights, 2, [&](int done, const frame& fr) {
                calls++;
                last_done = done;
                last = fr.radiance;
            });
            std::printf("height %d calls %d last %d\n", cam.height(), calls, last_done);
            return write_frame(argv[2], f.radiance, &last) ? 0 : 1;
        } catch (const std::exception& e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
    }
    if (argc >= 4 && std::string(argv[1]) == "denoise") {
        scene sc;
        std::string err;
        if (!load_scene(argv[2], sc, &err)) {
            std::fprintf(stderr, "%s\n", err.c_str());
            return 1;
        }
        sc.cam.image_width = 40;
        sc.cam.samples_per_pixel = 4;
        sc.cam.collect_features = true;
        sc.cam.seed = 3;
        try {
            frame f = sc.cam.render(sc.world, sc.lights);
            denoiser filter;
            std::vector<color> out = filter.run(f, sc.cam.image_width, sc.cam.height());
            std::printf("height %d\n", sc.cam.height());
            FILE* fp = std::fopen(argv[3], "wb");
            if (!fp) return 1;
            for (const auto* v : {&f.radiance, &f.albedo, &f.normal, &out})
                for (const auto& c : *v) std::fwrite(c.v, sizeof(double), 3, fp);
            std::fwrite(f.variance.data(), sizeof(double), f.variance.size(), fp);
            return std::fclose(fp) == 0 ? 0 : 1;
        } catch (const std::exception& e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
    }
    std::fprintf(stderr, "usage: api_render scene <file> <width> <spp> <seed> <out> | objects <out> | denoise <file> <out>\n");
    return 2;
}
