// blackhole_solution — headless counterpart of the reference's
// blackhole_solution_test program (include/blackhole/blackhole_solution_test.cc):
// the same scene (960x540 camera, fov pi/2, at (-2000, 0, 400); hole of mass
// 10 at the origin; disc from radius 100 to 1000 in z = 0 spun by pi/180 about
// its axis after every frame, :54-62,103-112,148,407), rendered through the
// C ABI on the GPU with a frame ring in place of the two screens and their
// clean thread, and PPM/PNG files in place of the MJPG video.
//
// The reference's window, HUD text and live keyboard are out of scope; a key
// script (--keys) replays its bindings, one key after each frame (:351-406):
// w/s a/d z/x move the camera along its x/y/z axes by 10, q/e i/k j/l turn it
// about them by pi/1800, upper case 10 times as far, -/+ change the field of
// view by pi/54, and ^ v < > , . move the hole by 5 along z, -z, y, -y, x, -x
// (the arrow keys and , . of the reference).
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "app_common.hpp"

namespace {

const char kUsage[] =
    "usage: blackhole_solution [--frames N] [--out DIR] [--format ppm|png|jpg|avi]\n"
    "                          [--width W --height H] [--fp32] [--escape R]\n"
    "                          [--disc-texture FILE] [--keys STRING] [--depth D]\n"
    "                          [--device I] [--quiet] [--dry-run]\n"
    "  --frames N          frames to render (default 3)\n"
    "  --out DIR           write DIR/frame_NNNNN.<format>, or for avi\n"
    "                      DIR/video.avi (MJPG, 29 fps); default: no files\n"
    "  --width/--height    frame size (default 960x540, the reference's)\n"
    "  --fp32              FP32 march with the FP64 photon-sphere band\n"
    "                      (default: FP64, the reference's operation order)\n"
    "  --disc-texture FILE PNG, JPEG or binary PPM disc texture, e.g. the reference's\n"
    "                      resource/acc_disc.png (default: a seeded 598x578\n"
    "                      banded texture)\n"
    "  --keys STRING       key script, one key after each frame\n"
    "  --depth D           frame ring slots, 2..8 (default 2: double buffering)\n"
    "  --dry-run           print each frame's scene as a JSON line; no GPU\n";

constexpr uint64_t kDiscSeed = 0xB1AC0003ull;
constexpr int kDiscSlot = 1;

}  // namespace

int main(int argc, char** argv) {
  app::Args a(argc, argv);
  if (a.flag("--help") || a.flag("-h")) {
    std::fputs(kUsage, stdout);
    return 0;
  }
  try {
    const int frames = (int)a.num("--frames", 3);
    const int W = (int)a.num("--width", 960), H = (int)a.num("--height", 540);
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
      bh::Image disc_tex;
      const std::string tex_path = a.get("--disc-texture", "");
      if (!tex_path.empty()) {
        disc_tex = app::load_image(tex_path);
      } else {
        disc_tex = bh::Image(598, 578, 4);
        bh_gen_disc_texture(disc_tex.data.data(), 598, 578, 16, kDiscSeed);
      }
      ctx->texture(kDiscSlot, disc_tex);
    }

    bh::Camera camera(W, H, bh::kPi / 2);
    camera.move_to({-2000, 0, 400});
    // Annulus((1000,1000,0), (-1000,1000,0), (-1000,-0100,0), (1000,-1000,0),
    // 1000, 100): -0100 is an octal literal, -64.
    bh::Annulus disc({1000, 1000, 0}, {-1000, 1000, 0}, {-1000, -64, 0}, {1000, -1000, 0}, 1000,
                     100, kDiscSlot);
    bh::Pose hole(bh::Vec3{0, 0, 0});

    bh_options opt{};
    opt.precision = a.flag("--fp32") ? BH_PREC_FP32 : BH_PREC_FP64;
    opt.nstep = 20;
    opt.step_cap = 10000;
    opt.band = 1e-3;
    const std::string esc = a.get("--escape", "");
    if (!esc.empty()) {
      opt.escape = 1;
      opt.r_escape = std::stod(esc);
    }

    std::unique_ptr<app::Ring> ring;
    std::unique_ptr<app::Writer> writer;
    if (!dry) {
      ring = std::make_unique<app::Ring>(*ctx, W, H, depth);
      writer = std::make_unique<app::Writer>(ring->get(), W, H, out, fmt, a.flag("--quiet"));
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (int f = 0; f < frames; ++f) {
      bh_scene s{};
      s.camera = camera.get();
      for (int k = 0; k < 3; ++k) s.hole.position[k] = hole.position()[k];
      s.hole.mass = 10;
      s.has_disc = 1;
      s.disc = disc.get();
      s.background = -1;
      if (dry) {
        std::printf("{\"frame\": %d, \"camera\": %s, \"hole\": %s, \"disc\": %s, \"normal\": %s}\n",
                    f, app::camera_json(s.camera).c_str(), app::vec_json(s.hole.position).c_str(),
                    app::vertices_json(s.disc.vertex, 5).c_str(),
                    app::vec_json(s.disc.normal).c_str());
      } else {
        int slot = -1;
        bh::check(bh_ring_render(ring->get(), &s, &opt, &slot), ctx->get());
        writer->push(slot, f);
      }

      if (f < (int)keys.size()) {
        const char k = keys[(size_t)f];
        if (!app::camera_key(camera, k, 10, bh::kPi / 1800.0, 10)) {
          switch (k) {
            case '^': hole.move_z(5); break;
            case 'v': hole.move_z(-5); break;
            case '<': hole.move_y(5); break;
            case '>': hole.move_y(-5); break;
            case ',': hole.move_x(5); break;
            case '.': hole.move_x(-5); break;
            default: break;
          }
        }
      }
      disc.rotate_z(bh::kPi / 180);
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
