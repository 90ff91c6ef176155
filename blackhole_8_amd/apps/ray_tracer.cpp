// ray_tracer — headless counterpart of the reference's ray_tracer_test program
// (include/blackhole/ray_tracer_test.cc): the flat-space scene (800x450
// camera, fov pi/2, at (-400, 120, 100) turned by RotateY(-pi/24) then
// RotateZ(pi/12); three textured rectangles; the ground plane z = 0 with
// ChessPattern2D(10); RayTracer::Prograde with 10 chords per pixel, :50-95,
// :145) rendered on the GPU through the C ABI, with the reference's per-frame
// "movement test" (:237-261): the second rectangle slides along x by 3 and
// turns back outside 0..100, the first turns by pi/180 about an axis shifted
// by (60, 50) along its y/z axes, the third by pi/120 about an axis shifted by
// (-100, 50) along its x/z axes.
//
// Frames go through a frame ring to PPM/PNG files in place of the window and
// MJPG video.  A key script (--keys) replays the reference's bindings, one key
// after each frame (:196-235): moves by 0.9, turns by pi/180, upper case 10
// times as far, -/+ change the field of view by pi/54.
#include <cstdio>
#include <memory>
#include <sstream>
#include <string>

#include "app_common.hpp"

namespace {

const char kUsage[] =
    "usage: ray_tracer [--frames N] [--out DIR] [--format ppm|png|jpg|avi]\n"
    "                  [--width W --height H] [--textures A,B,C] [--keys STRING]\n"
    "                  [--depth D] [--device I] [--quiet] [--dry-run]\n"
    "  --frames N        frames to render (default 3)\n"
    "  --out DIR         write DIR/frame_NNNNN.<format>, or for avi\n"
    "                    DIR/video.avi (MJPG, 29 fps); default: no files\n"
    "  --width/--height  frame size (default 800x450, the reference's)\n"
    "  --textures A,B,C  PNG, JPEG or binary PPM images for the three rectangles,\n"
    "                    e.g. the reference's resource/mooni.jpeg,karina.jpeg,\n"
    "                    winter.jpg (default: seeded noise of 480x360, 300x400,\n"
    "                    320x320)\n"
    "  --keys STRING     key script, one key after each frame\n"
    "  --depth D         frame ring slots, 2..8 (default 2: double buffering)\n"
    "  --dry-run         print each frame's scene as a JSON line; no GPU\n";

constexpr int kSlot0 = 40;
constexpr uint64_t kNoiseSeed = 0xB1AC0200ull;

}  // namespace

int main(int argc, char** argv) {
  app::Args a(argc, argv);
  if (a.flag("--help") || a.flag("-h")) {
    std::fputs(kUsage, stdout);
    return 0;
  }
  try {
    const int frames = (int)a.num("--frames", 3);
    const int W = (int)a.num("--width", 800), H = (int)a.num("--height", 450);
    const int depth = (int)a.num("--depth", 2);
    const std::string out = a.get("--out", ""), fmt = a.get("--format", "ppm");
    const std::string keys = a.get("--keys", "");
    if (frames < 1 || W < 1 || H < 1) throw std::invalid_argument("frames/width/height must be positive");
    if (!app::valid_format(fmt)) throw std::invalid_argument("--format is ppm, png, jpg or avi");

    const bool dry = a.flag("--dry-run");
    std::unique_ptr<bh::Context> ctx;
    if (!dry) {
      ctx = std::make_unique<bh::Context>((int)a.num("--device", 0));
      if (!a.flag("--quiet")) std::printf("device: %s\n", ctx->device_name().c_str());
    }

    // Textures: the reference's photos (mooni.jpeg, karina.jpeg, winter.jpg,
    // decoded as imread decodes them) or, by default, seeded noise of
    // comparable sizes.
    const int sizes[3][2] = {{360, 480}, {400, 300}, {320, 320}};
    std::vector<std::string> paths;
    {
      std::stringstream ss(a.get("--textures", ""));
      std::string p;
      while (std::getline(ss, p, ',')) paths.push_back(p);
      if (!paths.empty() && paths.size() != 3) throw std::invalid_argument("--textures takes three files");
    }
    for (int i = 0; i < 3 && !dry; ++i) {
      bh::Image img;
      if (!paths.empty()) {
        img = app::load_image(paths[(size_t)i]);
      } else {
        img = bh::Image(sizes[i][0], sizes[i][1], 3);
        bh_gen_noise(img.data.data(), img.data.size(), kNoiseSeed + (uint64_t)i);
      }
      ctx->texture(kSlot0 + i, img);
    }

    bh::Camera camera(W, H, bh::kPi / 2);
    camera.move_to({-400, 120, 100});
    camera.rotate_y(-bh::kPi / 24);
    camera.rotate_z(bh::kPi / 12);

    bh::Rectangle mooni({100, 120, 100}, {100, 0, 100}, {100, 0, 0}, {100, 120, 0}, kSlot0);
    bh::Rectangle karina({100, 0, 150}, {100, -100, 150}, {100, -100, 0}, {100, 0, 0}, kSlot0 + 1);
    bh::Rectangle winter({-100, 50, 100}, {-100, -50, 100}, {-100, -50, 0}, {-100, 50, 0},
                         kSlot0 + 2);
    const bh::InfinitePlane ground({0, 0, 0}, 10);
    int karina_dir = 1;

    std::unique_ptr<app::Ring> ring;
    std::unique_ptr<app::Writer> writer;
    if (!dry) {
      ring = std::make_unique<app::Ring>(*ctx, W, H, depth);
      writer = std::make_unique<app::Writer>(ring->get(), W, H, out, fmt, a.flag("--quiet"));
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int f = 0; f < frames; ++f) {
      bh_flat_scene s{};
      s.camera = camera.get();
      s.n_rects = 3;
      s.rects[0] = mooni.get();
      s.rects[1] = karina.get();
      s.rects[2] = winter.get();
      s.n_planes = 1;
      s.planes[0] = ground.get();
      s.steps = 10;
      if (dry) {
        std::printf("{\"frame\": %d, \"camera\": %s, \"rects\": [%s, %s, %s]}\n", f,
                    app::camera_json(s.camera).c_str(), app::vertices_json(s.rects[0].vertex, 5).c_str(),
                    app::vertices_json(s.rects[1].vertex, 5).c_str(),
                    app::vertices_json(s.rects[2].vertex, 5).c_str());
      } else {
        int slot = -1;
        bh::check(bh_ring_render_flat(ring->get(), &s, &slot), ctx->get());
        writer->push(slot, f);
      }

      if (f < (int)keys.size()) app::camera_key(camera, keys[(size_t)f], 0.9, bh::kPi / 180.0, 10);

      // the reference's movement test
      karina.move_x(3 * karina_dir);
      if (const double x = karina.position()[0]; x > 100)
        karina_dir = -1;
      else if (x < 0)
        karina_dir = 1;
      {
        const double y = -120.0, z = -100.0;
        mooni.move_y(-y / 2);
        mooni.move_z(-z / 2);
        mooni.rotate_x(bh::kPi / 180);
        mooni.move_y(y / 2);
        mooni.move_z(z / 2);
      }
      {
        const double x = 100.0, z = -100.0;
        winter.move_x(-x);
        winter.move_z(-z / 2);
        winter.rotate_y(bh::kPi / 120);
        winter.move_x(x);
        winter.move_z(z / 2);
      }
    }
    if (dry) return 0;
    const std::string err = writer->finish();
    const double wall =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (!err.empty()) throw std::runtime_error(err);
    std::printf("frames %d  wall %.1f ms  %.1f fps  gpu %.3f ms/frame  %.1f Mrays/s\n",
                writer->frames(), wall, writer->frames() * 1e3 / wall,
                writer->gpu_ms_total() / writer->frames(),
                (double)W * H * writer->frames() / (wall * 1e3));
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
}
