// Compile check for include/dmc.hpp: the reference call sequence
// (DMC_CUDA/DMC_CUDA/main.cpp:21-25) written against the drop-in header.
#include <vector>

#include "dmc.hpp"

int shim_compile_check() {
  dmc::float3 xyz_min = dmc::make_float3(0.f, 0.f, 0.f), xyz_max = dmc::make_float3(1.f, 1.f, 1.f);
  utils::Array3D<float> scalar_grid(8, 8, 8, 1.0f);
  std::vector<dmc::float3> vertices;
  std::vector<dmc::uint3> triangles;
  dmc::setup_device_luts();
  dmc::run_dmc(vertices, triangles, scalar_grid, xyz_min, xyz_max, 0.5f, 0);
  dmc::cleanup_device_luts();
  return (int)vertices.size() + (int)triangles.size();
}
