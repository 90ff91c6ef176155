// bh/blackhole.hpp — C++17 facade over the C ABI (bh_api.h): RAII context and
// camera path, scene builders with the reference's object poses, and C++
// exceptions on the host side (the ABI itself returns status codes).
//
// The builders mirror the reference's scene API: Camera<T> (camera.h:21-71),
// Object::Move*/Rotate* (object/object.h:58-88), StaticBlackhole
// (blackhole_solution.h:16-93), Annulus (object/vector_object.h:306-356),
// Rectangle and InfinitePlane (vector_object.h:84-239) and ChessPattern2D
// (object/pattern.h:14-50).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../bh_api.h"

namespace bh {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg)
      : std::runtime_error("bh error " + std::to_string(code) + ": " + msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const bh_ctx* ctx) {
  if (rc != BH_OK) throw Error(rc, bh_last_error(ctx));
}

using Vec3 = std::array<double, 3>;
constexpr double kPi = 3.14159265358979323846;

// An 8-bit image, rows x cols x channels, row-major (cv::Mat CV_8UCn layout).
struct Image {
  int rows = 0, cols = 0, channels = 3;
  std::vector<uint8_t> data;

  Image() = default;
  Image(int r, int c, int ch) : rows(r), cols(c), channels(ch), data((size_t)r * c * ch) {}
  bh_image view() const { return bh_image{data.data(), rows, cols, channels}; }
};

// Object<T> pose: a vertex list whose element 0 is the position, and three
// axes; rotations turn the axes and the vertices about the position.
class Pose {
 public:
  Pose() = default;
  explicit Pose(const Vec3& p) : verts_{p} {}

  const Vec3& position() const { return verts_[0]; }
  const Vec3& axis_x() const { return ax_[0]; }
  const Vec3& axis_y() const { return ax_[1]; }
  const Vec3& axis_z() const { return ax_[2]; }
  const std::vector<Vec3>& vertices() const { return verts_; }

  void move_to(const Vec3& p) { verts_[0] = p; }
  void move_x(double d) { move_along(ax_[0], d); }
  void move_y(double d) { move_along(ax_[1], d); }
  void move_z(double d) { move_along(ax_[2], d); }
  void rotate_x(double rad) { rotate(0, rad); }
  void rotate_y(double rad) { rotate(1, rad); }
  void rotate_z(double rad) { rotate(2, rad); }

 protected:
  void set_axes(const Vec3& x, const Vec3& y, const Vec3& z) { ax_ = {x, y, z}; }
  std::vector<Vec3> verts_{Vec3{0, 0, 0}};
  std::array<Vec3, 3> ax_{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};

 private:
  void move_along(const Vec3& a, double d) {
    for (auto& v : verts_)
      for (int k = 0; k < 3; ++k) v[k] += a[k] * d;
  }
  void rotate(int which, double rad) {
    double axes[9];
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) axes[3 * i + k] = ax_[i][k];
    std::vector<double> vs;
    for (const auto& v : verts_) vs.insert(vs.end(), v.begin(), v.end());
    bh_object_rotate(axes, vs.data(), (int)verts_.size(), which, rad);
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) ax_[i][k] = axes[3 * i + k];
    for (size_t j = 0; j < verts_.size(); ++j)
      for (int k = 0; k < 3; ++k) verts_[j][k] = vs[3 * j + k];
  }
};

// Camera<T>(width, height, fov): default axes x = (1,0,0), y = (0,-1,0),
// z = (0,0,-1) (camera.h:30-35).
class Camera : public Pose {
 public:
  Camera(int width, int height, double fov_x) : Pose(Vec3{0, 0, 0}), w_(width), h_(height) {
    set_axes({1, 0, 0}, {0, -1, 0}, {0, 0, -1});
    fov(fov_x);
  }
  void look_at(const Vec3& target, const Vec3& up = {0, 0, 1}) {
    bh_camera c = get();
    bh_camera_look_at(&c, target.data(), up.data());
    set_axes({c.axis_x[0], c.axis_x[1], c.axis_x[2]}, {c.axis_y[0], c.axis_y[1], c.axis_y[2]},
             {c.axis_z[0], c.axis_z[1], c.axis_z[2]});
  }
  int width() const { return w_; }
  int height() const { return h_; }
  double fov() const { return fov_; }
  // Camera::fov(fov_x) clamps to [0, nextafter(pi, 0)] (camera.h:37-41).
  void fov(double f) { fov_ = std::min(std::max(0.0, f), std::nextafter(kPi, 0.0)); }
  bh_camera get() const {
    bh_camera c{};
    for (int k = 0; k < 3; ++k) {
      c.position[k] = position()[k];
      c.axis_x[k] = axis_x()[k];
      c.axis_y[k] = axis_y()[k];
      c.axis_z[k] = axis_z()[k];
    }
    c.width = w_;
    c.height = h_;
    c.fov = fov_;
    return c;
  }

 private:
  int w_, h_;
  double fov_ = kPi / 2;
};

// Annulus(p1, p2, p3, p4, r_outer, r_inner): a disc in the plane of the
// texture square p1..p4, centred on the position.  The normal is taken at
// construction and kept through later moves and rotations, as the
// reference's norm_ member is (vector_object.h:320-326).
class Annulus : public Pose {
 public:
  Annulus(const Vec3& p1, const Vec3& p2, const Vec3& p3, const Vec3& p4, double r_outer,
          double r_inner, int texture_slot) {
    verts_ = {Vec3{0, 0, 0}, p1, p2, p3, p4};
    bh_annulus_init(&a_, p1.data(), p2.data(), p3.data(), p4.data(), r_outer, r_inner,
                    texture_slot);
  }
  bh_annulus get() const {
    bh_annulus a = a_;
    for (int v = 0; v < 5; ++v)
      for (int k = 0; k < 3; ++k) a.vertex[v][k] = verts_[v][k];
    return a;
  }

 private:
  bh_annulus a_{};
};

// Rectangle(p1, p2, p3, p4) with a texture.
class Rectangle : public Pose {
 public:
  Rectangle(const Vec3& p1, const Vec3& p2, const Vec3& p3, const Vec3& p4, int texture_slot)
      : slot_(texture_slot) {
    verts_ = {Vec3{0, 0, 0}, p1, p2, p3, p4};
  }
  bh_rect get() const {
    bh_rect r{};
    for (int v = 0; v < 5; ++v)
      for (int k = 0; k < 3; ++k) r.vertex[v][k] = verts_[v][k];
    r.texture = slot_;
    return r;
  }

 private:
  int slot_;
};

// InfinitePlane(position) with ChessPattern2D(pattern_size) (0: black).
class InfinitePlane : public Pose {
 public:
  InfinitePlane(const Vec3& p, double pattern_size) : Pose(p), size_(pattern_size) {}
  bh_plane get() const {
    bh_plane pl{};
    for (int k = 0; k < 3; ++k) {
      pl.position[k] = position()[k];
      pl.axis_x[k] = axis_x()[k];
      pl.axis_y[k] = axis_y()[k];
      pl.axis_z[k] = axis_z()[k];
    }
    pl.pattern_size = size_;
    return pl;
  }

 private:
  double size_;
};

// A rendered frame: 3 bytes per pixel in texture channel order, and the hit
// class per pixel.
struct Frame {
  int width = 0, height = 0;
  std::vector<uint8_t> rgb, hit;
  void resize(int w, int h) {
    width = w;
    height = h;
    rgb.resize((size_t)w * h * 3);
    hit.resize((size_t)w * h);
  }
};

class Context {
 public:
  explicit Context(int device = 0) { check(bh_ctx_create(device, &ctx_), nullptr); }
  ~Context() {
    if (ctx_) bh_ctx_destroy(ctx_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  Context(Context&& o) noexcept : ctx_(std::exchange(o.ctx_, nullptr)) {}

  bh_ctx* get() const { return ctx_; }

  std::string device_name() const {
    char name[256] = {0};
    int cu = 0, clk = 0;
    check(bh_device_info(ctx_, name, sizeof name, &cu, &clk), ctx_);
    return name;
  }

  // Material::SetTexture (material.h:29-35)
  void texture(int slot, const Image& img) {
    const bh_image v = img.view();
    check(bh_texture_set(ctx_, slot, &v), ctx_);
  }

  // The reference march of one frame into host memory.
  void render(const bh_scene& s, const bh_options& o, Frame& f) {
    f.resize(s.camera.width, s.camera.height);
    bh_outputs out{f.rgb.data(), f.hit.data(), nullptr, nullptr, nullptr};
    check(bh_render(ctx_, &s, &o, &out), ctx_);
  }

  // RayTracer::Prograde over every pixel (ray_tracer.h:68-85).
  void render_flat(const bh_flat_scene& s, Frame& f) {
    f.resize(s.camera.width, s.camera.height);
    bh_outputs out{f.rgb.data(), f.hit.data(), nullptr, nullptr, nullptr};
    check(bh_render_flat(ctx_, &s, &out), ctx_);
  }

 private:
  bh_ctx* ctx_ = nullptr;
};

// A hole (StaticBlackhole) scene with an optional disc.
struct Scene {
  Camera camera{960, 540, kPi / 2};
  Vec3 hole{0, 0, 0};
  double mass = 0.5;
  const Annulus* disc = nullptr;
  int background = -1;

  bh_scene get() const {
    bh_scene s{};
    s.camera = camera.get();
    for (int k = 0; k < 3; ++k) s.hole.position[k] = hole[k];
    s.hole.mass = mass;
    if (disc) {
      s.has_disc = 1;
      s.disc = disc->get();
    }
    s.background = background;
    return s;
  }
};

// A camera path captured as one HIP graph (bh_path_*).
class Path {
 public:
  Path(Context& ctx, const bh_scene& s, const bh_options& o, const std::vector<bh_camera>& cams)
      : ctx_(ctx.get()) {
    check(bh_path_create(ctx_, &s, &o, cams.data(), (int)cams.size(), &p_), ctx_);
  }
  ~Path() {
    if (p_) bh_path_destroy(p_);
  }
  Path(const Path&) = delete;
  Path& operator=(const Path&) = delete;
  float launch() {
    float ms = 0;
    check(bh_path_launch(p_, &ms), ctx_);
    return ms;
  }
  void read(int frame, Frame& f, int w, int h) {
    f.resize(w, h);
    check(bh_path_read(p_, frame, f.rgb.data(), f.hit.data()), ctx_);
  }

 private:
  bh_ctx* ctx_;
  bh_path* p_ = nullptr;
};

}  // namespace bh
