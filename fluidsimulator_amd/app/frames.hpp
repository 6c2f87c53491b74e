// Frame output of the hip backend's tools.
//
// * Binary state frames (".fsf"), the parity harness's format: a 64-byte
//   little-endian header, then positions and velocities as SoA float32 blocks
//   (3n each, original particle order) and, when flagged, the density of the
//   last constraint iteration (n float32).  Bit-exact, unlike ASCII output.
//     0  char[8]  "FLUIDFR1"
//     8  uint32   version (1)
//    12  uint32   flags (bit 0: rho block present)
//    16  uint64   n
//    24  int64    step (substeps completed)
//    32  float32  time
//    36  float32  rest density rho0 (for density statistics)
//    40  char[24] zero
// * ParaView output in the reference's format (io/src/vtk_writer.cpp:25-130):
//   ASCII PolyData .vtp frames named <base>_<6-digit index>.vtp with
//   positions and one vertex per particle, and a .pvd collection listing
//   (time, file) per frame.
#pragma once

#include <sys/stat.h>

#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fluid_hip.h"

namespace fluidapp {

struct Frame {
  uint64_t n = 0;
  int64_t step = 0;
  float time = 0.0f;
  float rho0 = 0.0f;
  std::vector<float> pos, vel, rho;  // 3n, 3n, n or empty
};

constexpr char kFrameMagic[8] = {'F', 'L', 'U', 'I', 'D', 'F', 'R', '1'};

inline bool make_dirs(const std::string& path) {
  if (path.empty()) return false;
  std::string cur;
  for (size_t i = 0; i <= path.size(); ++i) {
    if (i == path.size() || path[i] == '/') {
      if (!cur.empty() && cur != "/" && ::mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST)
        return false;
    }
    if (i < path.size()) cur += path[i];
  }
  return true;
}

inline std::string join_path(const std::string& dir, const std::string& file) {
  if (dir.empty()) return file;
  if (dir.back() == '/') return dir + file;
  return dir + "/" + file;
}

inline std::string frame_name(const std::string& base, size_t index, const char* ext) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "_%06zu.%s", index, ext);
  return base + buf;
}

inline bool write_frame(const std::string& path, const Frame& f) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) return false;
  unsigned char head[64];
  std::memset(head, 0, sizeof(head));
  std::memcpy(head, kFrameMagic, 8);
  const uint32_t version = 1, flags = f.rho.empty() ? 0u : 1u;
  std::memcpy(head + 8, &version, 4);
  std::memcpy(head + 12, &flags, 4);
  std::memcpy(head + 16, &f.n, 8);
  std::memcpy(head + 24, &f.step, 8);
  std::memcpy(head + 32, &f.time, 4);
  std::memcpy(head + 36, &f.rho0, 4);
  bool ok = std::fwrite(head, 1, 64, fp) == 64;
  const size_t m = static_cast<size_t>(3 * f.n);
  ok = ok && std::fwrite(f.pos.data(), sizeof(float), m, fp) == m;
  ok = ok && std::fwrite(f.vel.data(), sizeof(float), m, fp) == m;
  if (flags) ok = ok && std::fwrite(f.rho.data(), sizeof(float), f.n, fp) == f.n;
  return std::fclose(fp) == 0 && ok;
}

inline bool read_frame(const std::string& path, Frame& f, std::string* error) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) {
    if (error) *error = "cannot open " + path;
    return false;
  }
  unsigned char head[64];
  bool ok = std::fread(head, 1, 64, fp) == 64 && std::memcmp(head, kFrameMagic, 8) == 0;
  uint32_t version = 0, flags = 0;
  if (ok) {
    std::memcpy(&version, head + 8, 4);
    std::memcpy(&flags, head + 12, 4);
    std::memcpy(&f.n, head + 16, 8);
    std::memcpy(&f.step, head + 24, 8);
    std::memcpy(&f.time, head + 32, 4);
    std::memcpy(&f.rho0, head + 36, 4);
    ok = version == 1 && f.n < (1ull << 34);
  }
  if (ok) {
    const size_t m = static_cast<size_t>(3 * f.n);
    f.pos.resize(m);
    f.vel.resize(m);
    ok = std::fread(f.pos.data(), sizeof(float), m, fp) == m &&
         std::fread(f.vel.data(), sizeof(float), m, fp) == m;
    f.rho.clear();
    if (ok && (flags & 1u)) {
      f.rho.resize(f.n);
      ok = std::fread(f.rho.data(), sizeof(float), f.n, fp) == f.n;
    }
  }
  std::fclose(fp);
  if (!ok && error) *error = path + " is not a version-1 state frame";
  return ok;
}

// One ASCII PolyData frame, formatted as the reference's VtkWriter formats
// it (9 fixed decimals, one vertex per point).
inline bool write_vtp(const std::string& path, const float* x, const float* y, const float* z,
                      size_t n) {
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) return false;
  std::fprintf(fp, "<?xml version=\"1.0\"?>\n");
  std::fprintf(fp, "<VTKFile type=\"PolyData\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
  std::fprintf(fp, "  <PolyData>\n");
  std::fprintf(fp,
               "    <Piece NumberOfPoints=\"%zu\" NumberOfVerts=\"%zu\" NumberOfLines=\"0\" "
               "NumberOfStrips=\"0\" NumberOfPolys=\"0\">\n",
               n, n);
  std::fprintf(fp, "      <Points>\n");
  std::fprintf(fp, "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (size_t i = 0; i < n; ++i)
    std::fprintf(fp, "          %.9f %.9f %.9f\n", static_cast<double>(x[i]),
                 static_cast<double>(y[i]), static_cast<double>(z[i]));
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "      </Points>\n");
  std::fprintf(fp, "      <Verts>\n");
  std::fprintf(fp, "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n");
  for (size_t i = 0; i < n; ++i) std::fprintf(fp, "          %zu\n", i);
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
  for (size_t i = 0; i < n; ++i) std::fprintf(fp, "          %zu\n", i + 1);
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "      </Verts>\n");
  std::fprintf(fp, "    </Piece>\n");
  std::fprintf(fp, "  </PolyData>\n");
  std::fprintf(fp, "</VTKFile>\n");
  return std::fclose(fp) == 0;
}

// The solver parameters as text, one `name value` line per field, reals as
// hexadecimal floats (exact), so that two loaders can be compared bit for
// bit.  Needs fluid_hip.h.
inline bool write_params_text(const std::string& path, const fluid_hip_params& p) {
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) return false;
  auto f = [&](const char* name, float v) { std::fprintf(fp, "%s %a\n", name, static_cast<double>(v)); };
  auto i = [&](const char* name, int v) { std::fprintf(fp, "%s %d\n", name, v); };
  f("dt", p.dt);
  f("density", p.density);
  f("particle_mass", p.particle_mass);
  f("h", p.h);
  f("particle_radius", p.particle_radius);
  f("epsilon", p.epsilon);
  i("solver_iterations", p.solver_iterations);
  i("use_uniform_grid", p.use_uniform_grid);
  i("enable_scorr", p.enable_scorr);
  i("enable_xsph", p.enable_xsph);
  i("enable_vorticity", p.enable_vorticity);
  f("scorr_k", p.scorr_k);
  i("scorr_n", p.scorr_n);
  f("scorr_dq_coeff", p.scorr_dq_coeff);
  f("visc_c", p.visc_c);
  f("plane_restitution", p.plane_restitution);
  f("plane_friction", p.plane_friction);
  f("vort_epsilon", p.vort_epsilon);
  f("vort_norm_eps", p.vort_norm_eps);
  f("gravity_x", p.gravity[0]);
  f("gravity_y", p.gravity[1]);
  f("gravity_z", p.gravity[2]);
  i("n_planes", p.n_planes);
  for (int k = 0; k < p.n_planes; ++k)
    std::fprintf(fp, "plane %a %a %a %a\n", static_cast<double>(p.planes[k][0]),
                 static_cast<double>(p.planes[k][1]), static_cast<double>(p.planes[k][2]),
                 static_cast<double>(p.planes[k][3]));
  return std::fclose(fp) == 0;
}

struct PvdEntry {
  double time;
  std::string file;
};

inline bool write_pvd(const std::string& path, const std::vector<PvdEntry>& entries) {
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) return false;
  std::fprintf(fp, "<?xml version=\"1.0\"?>\n");
  std::fprintf(fp, "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
  std::fprintf(fp, "  <Collection>\n");
  for (const auto& e : entries)
    std::fprintf(fp, "    <DataSet timestep=\"%.9f\" group=\"\" part=\"0\" file=\"%s\"/>\n", e.time,
                 e.file.c_str());
  std::fprintf(fp, "  </Collection>\n");
  std::fprintf(fp, "</VTKFile>\n");
  return std::fclose(fp) == 0;
}

}  // namespace fluidapp
