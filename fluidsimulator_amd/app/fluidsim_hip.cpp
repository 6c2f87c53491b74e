// fluidsim_hip — the reference app's command line (app/src/main.cpp:19-279)
// with the hip backend behind it and the state resident on the GPU.
//
// The reference app steps a host State with fluid::step or fluid::cuda_step
// and writes a VTK frame from host memory every output interval.  Here the
// state is uploaded once; the GPU steps it between output frames (one
// synchronisation per chunk, fluid_hip_step); at an output frame the state
// is downloaded asynchronously into one of two page-locked buffers, and a
// writer thread formats the frame while the GPU steps the next chunk.
//
// Options of the reference (same names, values and checks): --backend,
// --no-output, --debug-print, --steps, --steps-per-sec, --enable-scorr,
// --enable-xsph, --enable-vorticity, --plane-restitution, --plane-friction,
// --threads, --no-omp, --fps, --duration, --scene, --output-dir.
// Extensions: --solver-iterations, --scorr-dq, --vort-epsilon, --seed and
// --jitter (seeded lattice jitter), --per-plane-collisions (per-plane
// restitution/friction from the scene file), --dump-dir / --dump-rho
// (binary state frames for the parity harness), --graph (hipGraph replay of
// a substep), --block-size, --autotune, --device.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "fluid_hip.h"
#include "frames.hpp"
#include "scene.hpp"

namespace {

enum class Kind { Flag, Value };
struct Opt {
  const char* name;
  Kind kind;
  const char* help;
};

const std::vector<Opt> kOpts = {
    {"backend", Kind::Value, "Backend to use: hip (cpu and cuda are the reference's own app)."},
    {"no-output", Kind::Flag, "Disable output for benchmarking."},
    {"debug-print", Kind::Flag, "Print each step's wall time (synchronises every step)."},
    {"steps", Kind::Value, "Number of simulation steps to run."},
    {"steps-per-sec", Kind::Value, "Simulation steps per second (sets dt = 1 / value)."},
    {"enable-scorr", Kind::Flag, "Enable s_corr constraint."},
    {"enable-xsph", Kind::Flag, "Enable XSPH viscosity."},
    {"enable-vorticity", Kind::Flag, "Enable vorticity confinement."},
    {"plane-restitution", Kind::Value, "Restitution for plane collisions (0 = no bounce)."},
    {"plane-friction", Kind::Value, "Tangential damping for plane collisions (0..1)."},
    {"threads", Kind::Value, "Host threads (accepted for compatibility; the step runs on the GPU)."},
    {"no-omp", Kind::Flag, "Accepted for compatibility; the step runs on the GPU."},
    {"fps", Kind::Value, "Output frames per second (controls output stride)."},
    {"duration", Kind::Value, "Simulation duration in seconds (overrides --steps)."},
    {"scene", Kind::Value, "Scene JSON to load."},
    {"output-dir", Kind::Value, "Directory for VTK output (for ParaView)."},
    // extensions
    {"solver-iterations", Kind::Value, "Constraint iterations per step (default 4)."},
    {"scorr-dq", Kind::Value, "s_corr reference distance as a fraction of h (default 0.3)."},
    {"vort-epsilon", Kind::Value, "Vorticity confinement strength (default 0.5)."},
    {"seed", Kind::Value, "Seed of the lattice jitter (with --jitter)."},
    {"jitter", Kind::Value, "Jitter every particle by this fraction of the spacing per axis."},
    {"per-plane-collisions", Kind::Flag,
     "Use the restitution/friction of each collision entry of the scene file."},
    {"dump-dir", Kind::Value, "Write binary state frames (.fsf) here at every output frame."},
    {"dump-rho", Kind::Flag, "Include each particle's density in the state frames."},
    {"scene-only", Kind::Flag,
     "Write the loaded scene (initial.fsf, params.txt) to --dump-dir and exit, no device."},
    {"graph", Kind::Flag, "Replay a captured hipGraph of one substep."},
    {"block-size", Kind::Value, "Threads per workgroup of the particle kernels (64..1024)."},
    {"autotune", Kind::Value, "Tune the launch configuration with this many trial steps first."},
    {"device", Kind::Value, "HIP device ordinal."},
};

struct Args {
  bool ok = true;
  std::string error;
  std::map<std::string, std::string> values;
  bool has(const std::string& k) const { return values.count(k) != 0; }
  std::string get(const std::string& k, const std::string& def = "") const {
    auto it = values.find(k);
    return it == values.end() ? def : it->second;
  }
};

Args parse(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    const std::string arg = argv[i];
    if (arg == "--") break;
    if (arg == "-h" || arg == "--help") {
      a.values["help"] = "1";
      continue;
    }
    if (arg.rfind("--", 0) != 0) continue;  // positionals are not used
    std::string name = arg.substr(2), value;
    if (name.empty()) {
      a.ok = false;
      a.error = "Invalid option: --";
      return a;
    }
    const size_t eq = name.find('=');
    const bool has_eq = eq != std::string::npos;
    if (has_eq) {
      value = name.substr(eq + 1);
      name = name.substr(0, eq);
    }
    const Opt* spec = nullptr;
    for (const Opt& o : kOpts)
      if (name == o.name) spec = &o;
    if (!spec) {
      a.ok = false;
      a.error = "Unknown option: --" + name;
      return a;
    }
    if (spec->kind == Kind::Flag) {
      if (has_eq && !value.empty()) {
        a.ok = false;
        a.error = "Option does not take a value: --" + name;
        return a;
      }
      a.values[name] = "1";
      continue;
    }
    if (!has_eq) {
      if (i + 1 >= argc) {
        a.ok = false;
        a.error = "Missing value for option: --" + name;
        return a;
      }
      value = argv[++i];
    }
    a.values[name] = value;
  }
  return a;
}

std::string usage(const char* argv0) {
  std::string u = std::string("Usage: ") + argv0 + " [options]\n\nOptions:\n  -h, --help\n";
  for (const Opt& o : kOpts) {
    u += std::string("  --") + o.name + (o.kind == Kind::Value ? " <value>" : "");
    u += std::string("\n      ") + o.help + "\n";
  }
  return u;
}

bool to_double(const std::string& s, double& out) {
  try {
    size_t k = 0;
    out = std::stod(s, &k);
    return true;
  } catch (const std::exception&) {
    return false;
  }
}

bool to_int(const std::string& s, int& out) {
  try {
    out = std::stoi(s);
    return true;
  } catch (const std::exception&) {
    return false;
  }
}

int check(int rc, const char* what) {
  if (rc != FLUID_HIP_OK) std::cerr << what << " failed: " << fluid_hip_last_error() << std::endl;
  return rc;
}

// A page-locked pair of SoA blocks for one output frame.
struct HostFrame {
  float* pos = nullptr;
  float* vel = nullptr;
  float time = 0.0f;
  int64_t step = 0;
  size_t index = 0;
};

}  // namespace

int main(int argc, char** argv) {
  const Args args = parse(argc, argv);
  if (!args.ok) {
    std::cerr << args.error << std::endl << usage(argv[0]);
    return 1;
  }
  if (args.has("help")) {
    std::cout << usage(argv[0]);
    return 0;
  }
  const std::string backend = args.get("backend", "hip");
  const bool output_enabled = !args.has("no-output");
  const bool debug_print = args.has("debug-print");
  const std::string output_dir = args.get("output-dir", "output");
  const std::string scene_path = args.get("scene");
  const std::string dump_dir = args.get("dump-dir");
  double steps_per_sec = std::numeric_limits<double>::quiet_NaN();
  double fps = -1.0, duration = -1.0;
  double restitution = std::numeric_limits<double>::quiet_NaN();
  double friction = std::numeric_limits<double>::quiet_NaN();
  int steps = 1, threads = 0;
  if (!to_int(args.get("steps", "1"), steps)) {
    std::cerr << "Invalid steps value." << std::endl << usage(argv[0]);
    return 1;
  }
  if (steps < 1) {
    std::cerr << "Steps must be >= 1." << std::endl;
    return 1;
  }
  auto positive = [&](const char* key, double& out) -> bool {
    if (!args.has(key)) return true;
    if (!to_double(args.get(key), out)) {
      std::cerr << "Invalid " << key << " value." << std::endl;
      return false;
    }
    if (!(out > 0.0)) {
      std::cerr << key << " must be > 0." << std::endl;
      return false;
    }
    return true;
  };
  if (!positive("steps-per-sec", steps_per_sec) || !positive("fps", fps) ||
      !positive("duration", duration))
    return 1;
  if (args.has("plane-restitution")) {
    if (!to_double(args.get("plane-restitution"), restitution)) {
      std::cerr << "Invalid plane-restitution value." << std::endl;
      return 1;
    }
    if (restitution < 0.0) {
      std::cerr << "plane-restitution must be >= 0." << std::endl;
      return 1;
    }
  }
  if (args.has("plane-friction")) {
    if (!to_double(args.get("plane-friction"), friction)) {
      std::cerr << "Invalid plane-friction value." << std::endl;
      return 1;
    }
    if (friction < 0.0 || friction > 1.0) {
      std::cerr << "plane-friction must be in [0, 1]." << std::endl;
      return 1;
    }
  }
  if (args.has("threads")) {
    if (!to_int(args.get("threads"), threads)) {
      std::cerr << "Invalid threads value." << std::endl;
      return 1;
    }
    if (threads < 1) {
      std::cerr << "threads must be >= 1." << std::endl;
      return 1;
    }
  }
  if (backend == "cpu" || backend == "cuda") {
    std::cerr << "This program runs the hip backend; --backend=" << backend
              << " is the reference's own app." << std::endl;
    return 1;
  }
  if (backend != "hip") {
    std::cerr << "Unsupported backend: " << backend << std::endl << usage(argv[0]);
    return 1;
  }

  const bool scene_only = args.has("scene-only");
  if (!scene_only) {
    std::cout << "FluidSimulator hip backend" << std::endl;
    std::cout << "hip_abi_version=" << fluid_hip_version() << std::endl;
    int device_count = 0;
    const char* dev_error = nullptr;
    if (!fluid_hip_device_available(&device_count, &dev_error)) {
      std::cerr << "HIP backend unavailable: "
                << (dev_error ? dev_error : "no HIP devices detected.") << std::endl;
      return 1;
    }
    std::cout << "backend=hip" << std::endl;
    std::cout << "hip_devices=" << device_count << std::endl;
  }

  fluid_hip_params params = fluidapp::default_params();
  fluidapp::Particles parts;
  if (!scene_path.empty()) {
    std::string error;
    fluidapp::SceneOptions so;
    so.per_plane_collisions = args.has("per-plane-collisions");
    if (!fluidapp::load_scene_json(scene_path, params, parts, &error, so)) {
      std::cerr << "Failed to load scene: " << error << std::endl;
      return 1;
    }
  } else {
    fluidapp::init_test_scene(params, parts);
  }
  if (steps_per_sec == steps_per_sec) params.dt = static_cast<float>(1.0 / steps_per_sec);
  if (duration > 0.0) {
    steps = static_cast<int>(std::ceil(duration / params.dt));
    if (steps < 1) steps = 1;
  }
  if (args.has("enable-scorr")) params.enable_scorr = 1;
  if (args.has("enable-xsph")) params.enable_xsph = 1;
  if (args.has("enable-vorticity")) params.enable_vorticity = 1;
  if (restitution == restitution) params.plane_restitution = static_cast<float>(restitution);
  if (friction == friction) params.plane_friction = static_cast<float>(friction);
  int ival = 0;
  double dval = 0.0;
  if (args.has("solver-iterations")) {
    if (!to_int(args.get("solver-iterations"), ival) || ival < 0) {
      std::cerr << "solver-iterations must be an integer >= 0." << std::endl;
      return 1;
    }
    params.solver_iterations = ival;
  }
  if (args.has("scorr-dq")) {
    if (!to_double(args.get("scorr-dq"), dval)) {
      std::cerr << "Invalid scorr-dq value." << std::endl;
      return 1;
    }
    params.scorr_dq_coeff = static_cast<float>(dval);
  }
  if (args.has("vort-epsilon")) {
    if (!to_double(args.get("vort-epsilon"), dval)) {
      std::cerr << "Invalid vort-epsilon value." << std::endl;
      return 1;
    }
    params.vort_epsilon = static_cast<float>(dval);
  }
  fluidapp::finish_plane_props(params);
  const size_t n = parts.n();
  if (args.has("jitter") || args.has("seed")) {
    double amp = 0.1;
    int seed = 0;
    if (args.has("jitter") && !to_double(args.get("jitter"), amp)) {
      std::cerr << "Invalid jitter value." << std::endl;
      return 1;
    }
    if (args.has("seed") && !to_int(args.get("seed"), seed)) {
      std::cerr << "Invalid seed value." << std::endl;
      return 1;
    }
    // the lattice spacing the scene was built with: particle_radius * 2
    const double d = static_cast<double>(params.particle_radius) * 2.0;
    const std::vector<double> u = fluidapp::uniform01(seed, 3 * n, 0);
    for (size_t k = 0; k < 3 * n; ++k)
      parts.pos[k] = static_cast<float>(static_cast<double>(parts.pos[k]) +
                                        d * amp * (2.0 * u[k] - 1.0));
  }

  if (scene_only) {
    // the scene as the solver would receive it, without touching a device
    if (dump_dir.empty() || !fluidapp::make_dirs(dump_dir)) {
      std::cerr << "--scene-only needs a writable --dump-dir." << std::endl;
      return 1;
    }
    fluidapp::Frame fr;
    fr.n = n;
    fr.rho0 = params.density;
    fr.pos = parts.pos;
    fr.vel = parts.vel;
    if (!fluidapp::write_frame(fluidapp::join_path(dump_dir, "initial.fsf"), fr) ||
        !fluidapp::write_params_text(fluidapp::join_path(dump_dir, "params.txt"), params)) {
      std::cerr << "Failed to write the scene." << std::endl;
      return 1;
    }
    std::cout << "particle_count=" << n << std::endl;
    return 0;
  }

  fluid_hip_config cfg;
  std::memset(&cfg, 0, sizeof(cfg));
  if (args.has("device") && !to_int(args.get("device"), cfg.device)) {
    std::cerr << "Invalid device value." << std::endl;
    return 1;
  }
  if (args.has("block-size") && !to_int(args.get("block-size"), cfg.block_size)) {
    std::cerr << "Invalid block-size value." << std::endl;
    return 1;
  }
  const int64_t count = static_cast<int64_t>(n);
  fluid_hip_ctx* ctx = nullptr;
  if (check(fluid_hip_create(&ctx, 1, &count, &params, &cfg), "fluid_hip_create")) return 1;
  if (check(fluid_hip_upload(ctx, parts.pos.data(), parts.vel.data(), 0.0f), "fluid_hip_upload")) {
    fluid_hip_destroy(ctx);
    return 1;
  }
  if (args.has("autotune")) {
    int trial = 0;
    if (!to_int(args.get("autotune"), trial) || trial < 1) {
      std::cerr << "autotune needs trial steps >= 1." << std::endl;
      fluid_hip_destroy(ctx);
      return 1;
    }
    fluid_hip_config best;
    double rep[4];
    if (check(fluid_hip_autotune(ctx, trial, 0, &best, rep), "fluid_hip_autotune")) {
      fluid_hip_destroy(ctx);
      return 1;
    }
    std::cout << "autotune block=" << best.block_size << " brick_edge=" << best.brick_edge
              << " xcd_remap=" << best.xcd_remap << " ms_per_step=" << rep[1] << std::endl;
  }

  const bool dumping = !dump_dir.empty();
  const bool writing = output_enabled || dumping;
  // the reference's output stride (app/src/main.cpp:239-244); state frames
  // follow the same stride as VTK frames
  int output_interval = 1;
  if (writing && fps > 0.0) {
    const double steps_per_frame = 1.0 / (params.dt * fps);
    output_interval = std::max(1, static_cast<int>(std::lround(steps_per_frame)));
  }
  if (output_enabled && !fluidapp::make_dirs(output_dir)) {
    std::cerr << "Failed to create output directory." << std::endl;
    fluid_hip_destroy(ctx);
    return 1;
  }
  if (dumping && !fluidapp::make_dirs(dump_dir)) {
    std::cerr << "Failed to create dump directory." << std::endl;
    fluid_hip_destroy(ctx);
    return 1;
  }

  // two page-locked frame buffers: one filling, one being written
  HostFrame buf[2];
  if (writing) {
    for (HostFrame& b : buf) {
      void* p = nullptr;
      if (check(fluid_hip_host_alloc(&p, static_cast<int64_t>(sizeof(float) * 6 * n)),
                "fluid_hip_host_alloc")) {
        fluid_hip_destroy(ctx);
        return 1;
      }
      b.pos = static_cast<float*>(p);
      b.vel = b.pos + 3 * n;
    }
  }
  std::vector<fluidapp::PvdEntry> pvd;
  bool write_failed = false;
  std::thread writer;
  auto write_frame = [&](const HostFrame& f, std::vector<float> rho) {
    if (output_enabled) {
      const std::string name = fluidapp::frame_name("frame", f.index, "vtp");
      if (!fluidapp::write_vtp(fluidapp::join_path(output_dir, name), f.pos, f.pos + n,
                               f.pos + 2 * n, n))
        write_failed = true;
    }
    if (dumping) {
      fluidapp::Frame fr;
      fr.n = n;
      fr.step = f.step;
      fr.time = f.time;
      fr.rho0 = params.density;
      fr.pos.assign(f.pos, f.pos + 3 * n);
      fr.vel.assign(f.vel, f.vel + 3 * n);
      fr.rho = std::move(rho);
      if (!fluidapp::write_frame(
              fluidapp::join_path(dump_dir, fluidapp::frame_name("state", f.index, "fsf")), fr))
        write_failed = true;
    }
  };

  const bool graph = args.has("graph");
  size_t frame_index = 0;
  int done = 0;
  int rc = 0;
  while (done < steps && rc == 0) {
    // the reference writes after step s (0-based) when s % interval == 0
    int chunk;
    if (debug_print) {
      chunk = 1;
    } else if (writing) {
      const int next = ((done + output_interval - 1) / output_interval) * output_interval;
      chunk = std::min(steps, next + 1) - done;
    } else {
      chunk = steps - done;
    }
    const auto t0 = std::chrono::steady_clock::now();
    rc = check(fluid_hip_step(ctx, chunk, graph ? 1 : 0), "fluid_hip_step");
    const auto t1 = std::chrono::steady_clock::now();
    if (rc) break;
    done += chunk;
    if (debug_print)
      std::cout << "step_done=" << done
                << " step_ms=" << std::chrono::duration<double, std::milli>(t1 - t0).count()
                << std::endl;
    if (writing && (done - 1) % output_interval == 0) {
      if (writer.joinable()) writer.join();  // the other buffer is free again
      HostFrame& f = buf[frame_index & 1];
      f.index = frame_index;
      f.step = done;
      rc = check(fluid_hip_download_async(ctx, f.pos, f.vel), "fluid_hip_download_async");
      if (!rc) rc = check(fluid_hip_download_wait(ctx, &f.time), "fluid_hip_download_wait");
      if (rc) break;
      std::vector<float> rho;
      if (dumping && args.has("dump-rho") && params.solver_iterations > 0) {
        rho.resize(n);
        rc = check(fluid_hip_get_diag(ctx, nullptr, nullptr, nullptr, rho.data(), nullptr),
                   "fluid_hip_get_diag");
        if (rc) break;
      }
      if (output_enabled)
        pvd.push_back({static_cast<double>(f.time), fluidapp::frame_name("frame", frame_index,
                                                                          "vtp")});
      writer = std::thread(write_frame, std::cref(f), std::move(rho));
      ++frame_index;
    }
  }
  if (writer.joinable()) writer.join();
  float end_time = 0.0f;
  if (!rc) rc = check(fluid_hip_download(ctx, nullptr, nullptr, &end_time), "fluid_hip_download");
  for (HostFrame& b : buf)
    if (b.pos) fluid_hip_host_free(b.pos);
  fluid_hip_destroy(ctx);
  if (rc) return 1;
  if (write_failed) {
    std::cerr << "Failed to write VTK frame." << std::endl;
    return 1;
  }
  std::cout << "particle_count=" << n << std::endl;
  std::cout << "end_time=" << end_time << std::endl;
  std::cout << "output_enabled=" << (output_enabled ? "true" : "false") << std::endl;
  if (output_enabled && !fluidapp::write_pvd(fluidapp::join_path(output_dir, "series.pvd"), pvd)) {
    std::cerr << "Failed to write PVD index." << std::endl;
    return 1;
  }
  return 0;
}
