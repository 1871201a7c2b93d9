// spatial_cell_sim: the command-line run of the reference (main.cu:142-722)
// over the C ABI.  Reads the config document (./config.json by default, the
// reference's 17 keys plus the extensions of sc_create) and a setup document
// (types, interactions and populations; ./setup.json by default), runs the
// init sequence, writes frame 0, steps, writes a frame every
// persistEveryNthFrame steps into ./results/frames.dat, and prints a timing
// line every reportEveryNthFrame steps.
//
// Frames are written as the reference's writer does it (main.cu:95-126,
// 666-680): the device reduces the state into its frame buffer at the end of
// a step, the copy to one of two pinned host buffers runs beside the next
// steps, and a writer thread appends the previous frame to the file while
// the device carries on.
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>

#include "../../include/spatial_cell.h"
#include "sc_json.h"

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

bool slurp(const std::string& path, std::string& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

// The setup document with "init": false in front of its keys, so that the
// driver runs the init sequence itself and writes frame 0 where the
// reference does (after the overlap removal, before the initial
// complexification: main.cu:283-338).
std::string without_init(const std::string& text, const scj::Value& doc) {
  const size_t brace = text.find('{');
  if (brace == std::string::npos) return "{\"init\": false}";
  return text.substr(0, brace + 1) + (doc.size() ? "\"init\": false, " : "\"init\": false") + text.substr(brace + 1);
}

// Appends frames to the file from its own thread; buffers go back to the
// stepping thread once written.
struct FrameWriter {
  std::ofstream out;
  std::mutex m;
  std::condition_variable cv;
  std::deque<std::pair<const char*, size_t>> queue;
  int in_flight = 0;
  bool done = false;
  double last_ms = 0.0;
  std::thread th;

  explicit FrameWriter(const std::string& path) : out(path, std::ios::binary | std::ios::trunc) {
    th = std::thread([this] { run(); });
  }
  bool ok() const { return (bool)out; }
  void run() {
    for (;;) {
      std::pair<const char*, size_t> job;
      {
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return done || !queue.empty(); });
        if (queue.empty()) return;
        job = queue.front();
        queue.pop_front();
      }
      Clock::time_point t = Clock::now();
      out.write(job.first, (std::streamsize)job.second);
      out.flush();
      std::lock_guard<std::mutex> lk(m);
      last_ms = ms_since(t);
      in_flight--;
      cv.notify_all();
    }
  }
  void push(const char* p, size_t n) {
    std::lock_guard<std::mutex> lk(m);
    queue.emplace_back(p, n);
    in_flight++;
    cv.notify_all();
  }
  // waits until at most `n` buffers are queued or being written
  void wait_below(int n) {
    std::unique_lock<std::mutex> lk(m);
    cv.wait(lk, [&] { return in_flight <= n; });
  }
  double last() {
    std::lock_guard<std::mutex> lk(m);
    return last_ms;
  }
  void finish() {
    {
      std::lock_guard<std::mutex> lk(m);
      done = true;
      cv.notify_all();
    }
    th.join();
  }
};

int usage(const char* argv0) {
  std::fprintf(stderr,
               "usage: %s [--config PATH] [--setup PATH] [--out PATH] [--steps N] [--device D] [--no-frames]\n"
               "  --config     config document (default ./config.json)\n"
               "  --setup      setup document: types, interactions, populations (default ./setup.json)\n"
               "  --out        frames file (default ./results/frames.dat)\n"
               "  --steps      steps to run (default: the config's \"steps\")\n"
               "  --device     HIP device ordinal (default 0)\n"
               "  --no-frames  write no frames file\n",
               argv0);
  return 2;
}

#define CHECK(call)                                                                       \
  do {                                                                                    \
    sc_status st_ = (call);                                                               \
    if (st_ != SC_OK) {                                                                   \
      std::fprintf(stderr, "%s failed (%d): %s\n", #call, (int)st_, sc_last_error(sim));  \
      return 1;                                                                           \
    }                                                                                     \
  } while (0)

}  // namespace

int main(int argc, char** argv) {
  std::string config_path = "./config.json", setup_path = "./setup.json", out_path = "./results/frames.dat";
  int steps = -1, device = 0;
  bool frames = true;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&](std::string& v) { if (i + 1 >= argc) return false; v = argv[++i]; return true; };
    std::string v;
    if (a == "--config" && next(v)) config_path = v;
    else if (a == "--setup" && next(v)) setup_path = v;
    else if (a == "--out" && next(v)) out_path = v;
    else if (a == "--steps" && next(v)) steps = std::atoi(v.c_str());
    else if (a == "--device" && next(v)) device = std::atoi(v.c_str());
    else if (a == "--no-frames") frames = false;
    else return usage(argv[0]);
  }
  std::string config, setup_text;
  if (!slurp(config_path, config)) { std::fprintf(stderr, "cannot read %s\n", config_path.c_str()); return 1; }
  if (!slurp(setup_path, setup_text)) { std::fprintf(stderr, "cannot read %s\n", setup_path.c_str()); return 1; }
  scj::Value setup_doc;
  std::string err;
  if (!scj::parse(setup_text, setup_doc, err)) { std::fprintf(stderr, "%s: %s\n", setup_path.c_str(), err.c_str()); return 1; }
  const bool run_init = setup_doc.get_bool("init", true);

  sc_sim* sim = nullptr;
  if (sc_create(config.c_str(), without_init(setup_text, setup_doc).c_str(), device, &sim) != SC_OK) {
    std::fprintf(stderr, "sc_create: %s\n", sc_last_error(nullptr));
    return 1;
  }
  sc_params P;
  CHECK(sc_get_params(sim, &P));
  if (steps < 0) steps = P.steps;
  // config.cuh:76-98 and main.cu:153-158
  std::printf("[Config]\n");
  std::printf("numParticles %d\nnumMetabolicParticles %d\nsteps %d\n", P.num_particles, P.num_metabolic_particles, steps);
  std::printf("simSize %f\nmaxInteractionDistance %f\nmaxDiffusionDistance %f\ngridCellSize %f\n", P.sim_size,
              P.max_interaction_distance, P.max_diffusion_distance, P.grid_cell_size);
  std::printf("movementNoiseScale %f\nrotationNoiseScale %f\nmetaboliteMovementNoiseScale %f\n", P.movement_noise_scale,
              P.rotation_noise_scale, P.metabolite_movement_noise_scale);
  std::printf("velocityDecay %f\nangularVelocityDecay %f\n", P.velocity_decay, P.angular_velocity_decay);
  std::printf("noiseCoordinationSteps %d\nrelaxationSteps %d\nmetaboliteDiffusionSteps %d\n",
              P.noise_coordination_steps, P.relaxation_steps, P.metabolite_diffusion_steps);
  std::printf("persistEveryNthFrame %d\nreportEveryNthFrame %d\n", P.persist_every_nth_frame, P.report_every_nth_frame);
  std::printf("boundary %s\n\n", P.boundary == SC_BOUNDARY_PERIODIC ? "periodic" : "clamp");
  if (P.grid_cell_size < P.max_interaction_distance)
    std::printf("WARNING! The maxInteractionDistance (%f) is less than gridCellSize (%f).\n", P.max_interaction_distance,
                P.grid_cell_size);
  if (P.grid_cell_size < P.max_diffusion_distance)
    std::printf("WARNING! The maxDiffusionDistance (%f) is less than gridCellSize (%f).\n", P.max_diffusion_distance,
                P.grid_cell_size);

  size_t frame_bytes = 0;
  CHECK(sc_frame_size(sim, &frame_bytes));
  FrameWriter* writer = nullptr;
  char* buf[2] = {nullptr, nullptr};
  if (frames) {
    writer = new FrameWriter(out_path);
    if (!writer->ok()) { std::fprintf(stderr, "cannot open %s\n", out_path.c_str()); return 1; }
    char header[SC_FRAME_HEADER_BYTES];
    CHECK(sc_frame_header(sim, header));
    writer->out.write(header, sizeof header);   // storage.cuh:22-27, before any frame is queued
    for (int k = 0; k < 2; k++) CHECK(sc_host_alloc(frame_bytes, (void**)&buf[k]));
  }
  int next_buf = 0;
  bool pending = false;   // a frame whose copy has not been waited for
  size_t written = 0;
  // hands the previous frame to the writer and queues this one's reduce and copy
  auto persist = [&]() -> sc_status {
    if (pending) {
      sc_status s = sc_frame_wait(sim);
      if (s != SC_OK) return s;
      writer->push(buf[next_buf ^ 1], frame_bytes);
    }
    writer->wait_below(1);   // the buffer about to be filled is no longer being written
    sc_status s = sc_emit_frame_async(sim, buf[next_buf], frame_bytes, &written);
    if (s != SC_OK) return s;
    pending = true;
    next_buf ^= 1;
    return SC_OK;
  };

  // init (main.cu:283-338): sort, overlap removal, frame 0, complexify at step -1
  if (run_init) {
    CHECK(sc_run_stage(sim, SC_STAGE_SORT, 0));
    CHECK(sc_run_stage(sim, SC_STAGE_REMOVE_INTERFERING, 0));
  }
  if (frames) CHECK(persist());
  if (run_init) CHECK(sc_run_stage(sim, SC_STAGE_COMPLEXIFY, -1));

  std::printf("\n[Simulating...]\n");
  const int report = P.report_every_nth_frame > 0 ? P.report_every_nth_frame : 10;
  const int every = P.persist_every_nth_frame > 0 ? P.persist_every_nth_frame : 1;
  Clock::time_point t0 = Clock::now(), tr = t0;
  int since = 0;
  for (int step = 0; step < steps; step++) {
    CHECK(sc_step(sim, 1));
    since++;
    if (frames && step % every == 0) CHECK(persist());   // main.cu:666-680
    if (step % report == 0) {
      CHECK(sc_synchronize(sim));
      sc_counts c;
      CHECK(sc_counts_get(sim, &c));
      const double ms = ms_since(tr) / since;
      std::printf("step %d, nActiveParticles %d, nActiveMetabolicParticles %d, full step time %f", step, c.n_active,
                  c.n_active_metabolic, ms);
      if (frames) std::printf(", persistFrame (previous) %f", writer->last());
      std::printf("\n");
      std::fflush(stdout);
      tr = Clock::now();
      since = 0;
    }
  }
  CHECK(sc_synchronize(sim));
  if (frames) {
    if (pending) writer->push(buf[next_buf ^ 1], frame_bytes);
    writer->finish();
    delete writer;
    for (int k = 0; k < 2; k++) sc_host_free(buf[k]);
  }
  std::printf("time %f\n", ms_since(t0) / 1000.0);
  sc_destroy(sim);
  std::printf("Done\n");
  return 0;
}
