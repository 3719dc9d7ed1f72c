// Compile check for include/dmc.hpp: the reference call sequence
// (DMC_CUDA/DMC_CUDA/main.cpp:5-28) written against the drop-in header with
// the same unqualified names.
#include <string>
#include <vector>

#include "dmc.hpp"

int shim_compile_check() {
  float3 xyz_min = make_float3(0.f, 0.f, 0.f), xyz_max = make_float3(1.f, 1.f, 1.f);
  uint3 size3D = make_uint3(8, 8, 8);
  utils::Array3D<float> scalar_grid(size3D.x, size3D.y, size3D.z);
  float iso_value = 0.5f;
  float max_iter = 0;
  std::vector<float3> vertices;
  std::vector<uint3> triangles;
  dmc::setup_device_luts();
  dmc::run_dmc(vertices, triangles, scalar_grid, xyz_min, xyz_max, iso_value, max_iter);
  dmc::cleanup_device_luts();
  return (int)vertices.size() + (int)triangles.size();
}
