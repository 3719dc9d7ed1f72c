// C++ drop-in for the reference host API (DMC_CUDA/DMC_CUDA/include/dmc.hpp:40-57
// and include/utils.hpp:14-66), implemented over the mdmc C ABI.
//
// Differences a caller can observe:
//   * errors throw dmc::error (a std::runtime_error carrying the mdmc status)
//     instead of printing and calling exit(1) (include/cuda_includes.h:13-22);
//   * an empty surface returns empty vectors (the reference exits,
//     dmc.cu:2113-2114);
//   * num_smooth > 0 smooths as the reference does, except that a convex
//     ring's split corner is chosen by comparing cosines (IEEE sqrt and
//     division) instead of acosf/rsqrtf angles;
//   * float3/uint3 are this header's own 12-byte PODs (the reference uses the
//     CUDA vector types of the same layout); they are also declared in the
//     global namespace, as the reference's callers use them unqualified
//     (main.cpp:8-18), unless MDMC_DMC_NO_GLOBAL_TYPES is defined.
// Link with framework_amd/lib/libmdmc.so.
#ifndef MDMC_DMC_HPP
#define MDMC_DMC_HPP

#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mdmc.h"

namespace dmc {

struct float3 {
  float x, y, z;
};
struct uint3 {
  unsigned x, y, z;
};
static_assert(sizeof(float3) == 12 && sizeof(uint3) == 12, "reference vertex/triangle layout");

inline float3 make_float3(float x, float y, float z) { return float3{x, y, z}; }
inline uint3 make_uint3(unsigned x, unsigned y, unsigned z) { return uint3{x, y, z}; }

class error : public std::runtime_error {
 public:
  error(int status, const std::string &where)
      : std::runtime_error(where + ": " + mdmc_status_string(status) + " (" + mdmc_last_error() + ")"),
        status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

}  // namespace dmc

#ifndef MDMC_DMC_NO_GLOBAL_TYPES
using dmc::float3;
using dmc::make_float3;
using dmc::make_uint3;
using dmc::uint3;
#endif

namespace utils {

// Same interface as the reference's Array3D (include/utils.hpp:14-66): x
// fastest, std::vector storage.
template <typename T>
class Array3D {
 public:
  typedef T value_type;
  Array3D(const dmc::uint3 &size3D, const T &value = T())
      : dx_(size3D.x), dy_(size3D.y), dz_(size3D.z), data_((size_t)size3D.x * size3D.y * size3D.z, value) {}
  Array3D(unsigned dim_x, unsigned dim_y, unsigned dim_z, const T &value = T())
      : dx_(dim_x), dy_(dim_y), dz_(dim_z), data_((size_t)dim_x * dim_y * dim_z, value) {}
  const T &operator()(unsigned x, unsigned y, unsigned z) const {
    return data_[((size_t)z * dy_ + y) * dx_ + x];
  }
  T &operator()(unsigned x, unsigned y, unsigned z) { return data_[((size_t)z * dy_ + y) * dx_ + x]; }
  unsigned dim_x() const { return dx_; }
  unsigned dim_y() const { return dy_; }
  unsigned dim_z() const { return dz_; }
  T *data() { return data_.data(); }
  const T *data() const { return data_.data(); }
  size_t size() const { return data_.size(); }

 private:
  unsigned dx_, dy_, dz_;
  std::vector<T> data_;
};

// include/utils.hpp:118-122
inline float ijk_to_xyz(unsigned i, unsigned size, float f_range, float f_min) {
  return (float)i / size * f_range + f_min;
}

}  // namespace utils

namespace dmc {

typedef utils::Array3D<float> scalar_grid_type;

// dmc::setup_device_luts (include/dmc.hpp:40)
inline void setup_device_luts() {
  const int rc = mdmc_setup();
  if (rc != MDMC_OK) throw error(rc, "setup_device_luts");
}

// dmc::cleanup_device_luts (include/dmc.hpp:48)
inline void cleanup_device_luts() {
  const int rc = mdmc_cleanup();
  if (rc != MDMC_OK) throw error(rc, "cleanup_device_luts");
}

// dmc::run_dmc (include/dmc.hpp:56-57): vertices in the reference layout,
// two triangles (q0,q1,q2),(q2,q3,q0) per quad (dmc.cu:2039-2045).
inline void run_dmc(std::vector<float3> &vertices, std::vector<uint3> &triangles,
                    const scalar_grid_type &h_scalar_grid, const float3 &xyz_min,
                    const float3 &xyz_max, float iso_value, unsigned num_smooth = 0) {
  const uint32_t dims[3] = {h_scalar_grid.dim_x(), h_scalar_grid.dim_y(), h_scalar_grid.dim_z()};
  const float mn[3] = {xyz_min.x, xyz_min.y, xyz_min.z};
  const float mx[3] = {xyz_max.x, xyz_max.y, xyz_max.z};
  mdmc_mesh mesh;
  const int rc = mdmc_run(h_scalar_grid.data(), dims, mn, mx, iso_value, num_smooth, &mesh);
  if (rc != MDMC_OK) throw error(rc, "run_dmc");
  vertices.clear();
  triangles.clear();
  vertices.resize(mesh.num_vertices);
  for (uint32_t i = 0; i < mesh.num_vertices; ++i)
    vertices[i] = float3{mesh.vertices[3 * i], mesh.vertices[3 * i + 1], mesh.vertices[3 * i + 2]};
  triangles.resize(2 * (size_t)mesh.num_quads);
  for (uint32_t q = 0; q < mesh.num_quads; ++q) {
    const int32_t *p = mesh.quads + 4 * (size_t)q;
    triangles[2 * (size_t)q] = uint3{(unsigned)p[0], (unsigned)p[1], (unsigned)p[2]};
    triangles[2 * (size_t)q + 1] = uint3{(unsigned)p[2], (unsigned)p[3], (unsigned)p[0]};
  }
  mdmc_mesh_free(&mesh);
}

// dump_obj (dmc.cu:2238-2246): Wavefront OBJ through std::ostream at its
// default precision, 1-based triangle indices; byte-identical to
// mdmc_write_obj on the quads the triangles come from.
inline void dump_obj(const char *filename, const std::vector<float3> &vertices,
                     const std::vector<uint3> &triangles) {
  std::ofstream of(filename);
  for (const float3 &v : vertices) of << "v " << v.x << " " << v.y << " " << v.z << "\n";
  for (const uint3 &t : triangles) of << "f " << t.x + 1 << " " << t.y + 1 << " " << t.z + 1 << "\n";
  if (!of) throw error(MDMC_ERR_INVALID, "dump_obj");
}

}  // namespace dmc

#endif  // MDMC_DMC_HPP
