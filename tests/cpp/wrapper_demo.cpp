// Drives clothmi::ClothBodyHip the way the reference's scene objects drive a
// physics body (Rendering/SceneObject.cpp:104-134): FixedUpdate per substep, Update
// per frame. Reads a scene from a folder and writes the final positions and
// normals there.
//   wrapper_demo DIR SUBSTEPS_PER_FRAME FRAMES
// DIR holds params.bin (cm_params), xyz.f32 (3N), tri.u32 (3T), fixed.u8 (N) and
// inv_mass.f32 (1 float, free vertices); out_xyz.f32 / out_nrm.f32 are written.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "../../include/clothmi.hpp"

template <class T>
static std::vector<T> load(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  std::vector<T> out(raw.size() / sizeof(T));
  std::copy(raw.begin(), raw.begin() + out.size() * sizeof(T), (char*)out.data());
  return out;
}

template <class T>
static void save(const std::string& path, const std::vector<T>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write((const char*)v.data(), (std::streamsize)(v.size() * sizeof(T)));
}

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s DIR SUBSTEPS_PER_FRAME FRAMES\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  const int sub = std::atoi(argv[2]), frames = std::atoi(argv[3]);
  try {
    const std::vector<cm_params> prm = load<cm_params>(dir + "/params.bin");
    const std::vector<float> xyz = load<float>(dir + "/xyz.f32");
    const std::vector<unsigned int> tri = load<unsigned int>(dir + "/tri.u32");
    const std::vector<uint8_t> fx = load<uint8_t>(dir + "/fixed.u8");
    const std::vector<float> w = load<float>(dir + "/inv_mass.f32");
    if (prm.size() != 1 || w.size() != 1) throw std::runtime_error("bad params / inv_mass");
    std::vector<bool> fixed(fx.begin(), fx.end());
    clothmi::ClothBodyHip body(xyz, tri, prm[0], fixed, w[0]);
    clothmi::PhysicsBody& pb = body;
    for (int f = 0; f < frames; ++f) {
      for (int s = 0; s < sub; ++s) pb.FixedUpdate();
      pb.Update();
    }
    save(dir + "/out_xyz.f32", body.Positions());
    save(dir + "/out_nrm.f32", body.Normals());
    std::printf("wrapper_demo: %zu vertices, %zu triangles, %d frames\n", xyz.size() / 3,
                pb.GetTriangles().size() / 3, frames);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "wrapper_demo: %s\n", e.what());
    return 1;
  }
  return 0;
}
