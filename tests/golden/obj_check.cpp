// OBJ writers: dmc::dump_obj (the shim, reference form over triangles) and
// mdmc_write_obj (C ABI, over quads) on the same mesh.  Usage:
// obj_check <vertices.f32> <quads.i32> <out_shim.obj> <out_abi.obj>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <vector>

#include "dmc.hpp"

template <typename T>
static std::vector<T> slurp(const char *path) {
  std::ifstream f(path, std::ios::binary);
  std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  std::vector<T> out(b.size() / sizeof(T));
  std::copy(b.begin(), b.begin() + out.size() * sizeof(T), reinterpret_cast<char *>(out.data()));
  return out;
}

int main(int argc, char **argv) {
  if (argc != 5) return 2;
  const std::vector<float> v = slurp<float>(argv[1]);
  const std::vector<int32_t> q = slurp<int32_t>(argv[2]);
  std::vector<float3> verts(v.size() / 3);
  for (size_t i = 0; i < verts.size(); ++i) verts[i] = float3{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
  std::vector<uint3> tris;
  for (size_t i = 0; i + 3 < q.size(); i += 4) {
    tris.push_back(uint3{(unsigned)q[i], (unsigned)q[i + 1], (unsigned)q[i + 2]});
    tris.push_back(uint3{(unsigned)q[i + 2], (unsigned)q[i + 3], (unsigned)q[i]});
  }
  dmc::dump_obj(argv[3], verts, tris);
  return mdmc_write_obj(argv[4], v.data(), verts.size(), q.data(), q.size() / 4) == MDMC_OK ? 0 : 1;
}
