// app_common.hpp — pieces shared by the two demo programs: argument parsing,
// the scripted keyboard of the reference's interactive loop, and the frame
// writer thread that drains the frame ring.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bh/blackhole.hpp"

namespace app {

struct Args {
  std::vector<std::string> v;
  Args(int argc, char** argv) : v(argv + 1, argv + argc) {}
  bool flag(const std::string& name) const {
    for (const auto& a : v)
      if (a == name) return true;
    return false;
  }
  std::string get(const std::string& name, const std::string& def) const {
    for (size_t i = 0; i + 1 < v.size(); ++i)
      if (v[i] == name) return v[i + 1];
    return def;
  }
  long num(const std::string& name, long def) const {
    const std::string s = get(name, "");
    if (s.empty()) return def;
    char* end = nullptr;
    const long x = std::strtol(s.c_str(), &end, 10);
    if (!end || *end) throw std::invalid_argument(name + " expects an integer");
    return x;
  }
};

// The key bindings of the reference's loop (blackhole_solution_test.cc:351-406,
// ray_tracer_test.cc:196-235), one key applied after each frame.  Upper case
// moves or turns `faster` times as far.  '-' and '+' change the field of view
// by pi/54.  Returns false for a key the camera does not handle.
inline bool camera_key(bh::Camera& cam, char key, double move_d, double rotate_d, double faster) {
  const double m = (key >= 'A' && key <= 'Z') ? move_d * faster : move_d;
  const double r = (key >= 'A' && key <= 'Z') ? rotate_d * faster : rotate_d;
  switch (key) {
    case 'w': case 'W': cam.move_x(m); return true;
    case 's': case 'S': cam.move_x(-m); return true;
    case 'a': case 'A': cam.move_y(-m); return true;
    case 'd': case 'D': cam.move_y(m); return true;
    case 'z': case 'Z': cam.move_z(-m); return true;
    case 'x': case 'X': cam.move_z(m); return true;
    case 'q': case 'Q': cam.rotate_x(-r); return true;
    case 'e': case 'E': cam.rotate_x(r); return true;
    case 'i': case 'I': cam.rotate_y(r); return true;
    case 'k': case 'K': cam.rotate_y(-r); return true;
    case 'j': case 'J': cam.rotate_z(-r); return true;
    case 'l': case 'L': cam.rotate_z(r); return true;
    case '-': cam.fov(cam.fov() - bh::kPi / 54); return true;
    case '+': cam.fov(cam.fov() + bh::kPi / 54); return true;
    default: return false;
  }
}

// Frame file formats of --format: one image file per frame, or "avi": every
// frame into DIR/video.avi, Motion-JPEG at 29 frames/s as the reference
// records (blackhole_solution_test.cc:71-72).
inline bool valid_format(const std::string& f) {
  return f == "ppm" || f == "png" || f == "jpg" || f == "avi";
}
constexpr double kVideoFps = 29;

// Drains ring slots on its own thread: waits for each frame's copy, writes
// it (when a directory is given), prints its timing and hands the slot back.
// Stands in for the reference's clean thread and cv::VideoWriter::write.
class Writer {
 public:
  Writer(bh_ring* ring, int w, int h, std::string dir, std::string ext, bool quiet)
      : ring_(ring), w_(w), h_(h), dir_(std::move(dir)), ext_(std::move(ext)), quiet_(quiet) {
    if (!dir_.empty() && ext_ == "avi" &&
        bh_video_open((dir_ + "/video.avi").c_str(), h_, w_, kVideoFps, 95, &video_) != BH_OK)
      throw std::runtime_error(bh_io_last_error());
    th_ = std::thread([this] { run(); });
  }
  ~Writer() { finish(); }

  void push(int slot, int frame) {
    {
      std::lock_guard<std::mutex> lk(m_);
      q_.push_back({slot, frame, clock::now()});
    }
    cv_.notify_one();
  }
  // Waits for every queued frame; returns the first error (empty if none).
  std::string finish() {
    {
      std::lock_guard<std::mutex> lk(m_);
      done_ = true;
    }
    cv_.notify_one();
    if (th_.joinable()) th_.join();
    if (video_) {
      if (bh_video_close(video_) != BH_OK && err_.empty()) err_ = bh_io_last_error();
      video_ = nullptr;
    }
    return err_;
  }
  double gpu_ms_total() const { return gpu_ms_; }
  int frames() const { return frames_; }

 private:
  using clock = std::chrono::steady_clock;
  struct Item {
    int slot, frame;
    clock::time_point t0;
  };
  void run() {
    for (;;) {
      Item it;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [this] { return done_ || !q_.empty(); });
        if (q_.empty()) return;
        it = q_.front();
        q_.pop_front();
      }
      const uint8_t *rgb = nullptr, *hit = nullptr;
      float ms = 0;
      if (bh_ring_wait(ring_, it.slot, &rgb, &hit, &ms) != BH_OK) {
        if (err_.empty()) err_ = bh_last_error(nullptr);
        (void)bh_ring_release(ring_, it.slot);
        continue;
      }
      if (video_) {  // frames arrive in the order they were queued
        if (bh_video_write(video_, rgb) != BH_OK && err_.empty()) err_ = bh_io_last_error();
      } else if (!dir_.empty()) {
        char name[64];
        std::snprintf(name, sizeof name, "/frame_%05d.%s", it.frame, ext_.c_str());
        if (bh_image_write((dir_ + name).c_str(), rgb, h_, w_) != BH_OK && err_.empty())
          err_ = bh_io_last_error();
      }
      (void)bh_ring_release(ring_, it.slot);
      const double dt =
          std::chrono::duration<double, std::milli>(clock::now() - it.t0).count();
      gpu_ms_ += ms;
      ++frames_;
      if (!quiet_) std::printf("Took %.2fms (gpu %.3fms) frame %d\n", dt, ms, it.frame);
    }
  }

  bh_ring* ring_;
  int w_, h_;
  std::string dir_, ext_;
  bool quiet_;
  std::mutex m_;
  std::condition_variable cv_;
  std::deque<Item> q_;
  bool done_ = false;
  std::string err_;
  double gpu_ms_ = 0;
  int frames_ = 0;
  bh_video* video_ = nullptr;
  std::thread th_;
};

// RAII frame ring on a context.
class Ring {
 public:
  Ring(bh::Context& ctx, int w, int h, int depth) {
    bh::check(bh_ring_create(ctx.get(), w, h, depth, &r_), ctx.get());
  }
  ~Ring() {
    if (r_) bh_ring_destroy(r_);
  }
  Ring(const Ring&) = delete;
  Ring& operator=(const Ring&) = delete;
  bh_ring* get() const { return r_; }

 private:
  bh_ring* r_ = nullptr;
};

// "[x, y, z]" with every double printed round-trip exact.
inline std::string vec_json(const double* v) {
  char b[96];
  std::snprintf(b, sizeof b, "[%.17g, %.17g, %.17g]", v[0], v[1], v[2]);
  return b;
}

inline std::string camera_json(const bh_camera& c) {
  char f[32];
  std::snprintf(f, sizeof f, "%.17g", c.fov);
  return "{\"position\": " + vec_json(c.position) + ", \"axis_x\": " + vec_json(c.axis_x) +
         ", \"axis_y\": " + vec_json(c.axis_y) + ", \"axis_z\": " + vec_json(c.axis_z) +
         ", \"width\": " + std::to_string(c.width) + ", \"height\": " + std::to_string(c.height) +
         ", \"fov\": " + f + "}";
}

inline std::string vertices_json(const double (*v)[3], int n) {
  std::string s = "[";
  for (int i = 0; i < n; ++i) s += (i ? ", " : "") + vec_json(v[i]);
  return s + "]";
}

inline bh::Image load_image(const std::string& path) {
  int rows = 0, cols = 0;
  if (bh_image_load(path.c_str(), nullptr, 0, &rows, &cols) != BH_OK)
    throw std::runtime_error(path + ": " + bh_io_last_error());
  bh::Image img(rows, cols, 3);
  if (bh_image_load(path.c_str(), img.data.data(), img.data.size(), &rows, &cols) != BH_OK)
    throw std::runtime_error(path + ": " + bh_io_last_error());
  return img;
}

}  // namespace app
