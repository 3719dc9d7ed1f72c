// C++ drop-in for the reference host API (DMC_CUDA/DMC_CUDA/include/dmc.hpp:10-59
// and include/utils.hpp:11-160), implemented over the mdmc C ABI.
//
// Differences a caller can observe:
//   * errors throw dmc::error (a std::runtime_error carrying the mdmc status)
//     instead of printing and calling exit(1) (include/cuda_includes.h:13-22);
//   * an empty surface returns empty vectors (the reference exits,
//     dmc.cu:2113-2114);
//   * num_smooth > 0 smooths as the reference does, except that a convex
//     ring's split corner is chosen by comparing cosines (IEEE sqrt and
//     division) instead of acosf/rsqrtf angles;
//   * the 11-pointer setup_device_luts hands out device copies owned by the
//     library (the 2B/3B list in ascending order); the matching cleanup
//     frees them and throws if given other pointers;
//   * float2/float3/uint3 are HIP's vector types when HIP's headers are in
//     use (__HIP_PLATFORM_AMD__ or hipcc), as the reference uses CUDA's;
//     otherwise this header's own PODs of the same layout (12-byte
//     float3/uint3, componentwise + - *), also declared in the global
//     namespace as the reference's callers use them unqualified
//     (main.cpp:8-18), unless MDMC_DMC_NO_GLOBAL_TYPES is defined.
// Link with framework_amd/lib/libmdmc.so.
#ifndef MDMC_DMC_HPP
#define MDMC_DMC_HPP

#include <stddef.h>
#include <stdint.h>

#include <fstream>
#include <functional>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mdmc.h"

#if defined(__HIP_PLATFORM_AMD__) || defined(__HIPCC__)
// With HIP's headers in play the shim uses HIP's vector types, as the
// reference uses CUDA's (12-byte float3/uint3 with componentwise operators).
#include <hip/hip_vector_types.h>
#define MDMC_DMC_HIP_VECTOR_TYPES 1
#endif

namespace dmc {

#ifdef MDMC_DMC_HIP_VECTOR_TYPES
using ::float2;
using ::float3;
using ::make_float2;
using ::make_float3;
using ::make_uint3;
using ::uint3;
#else
struct float2 {
  float x, y;
};
struct float3 {
  float x, y, z;
};
struct uint3 {
  unsigned x, y, z;
};

inline float2 make_float2(float x, float y) { return float2{x, y}; }
inline float3 make_float3(float x, float y, float z) { return float3{x, y, z}; }
inline uint3 make_uint3(unsigned x, unsigned y, unsigned z) { return uint3{x, y, z}; }

// componentwise arithmetic, as callers of the CUDA vector types expect
// (e.g. xyz_max - xyz_min, dmc.cu:2269)
inline float3 operator+(const float3 &a, const float3 &b) { return float3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline float3 operator-(const float3 &a, const float3 &b) { return float3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float3 operator*(const float3 &a, float s) { return float3{a.x * s, a.y * s, a.z * s}; }
inline float3 operator*(float s, const float3 &a) { return a * s; }
#endif
static_assert(sizeof(float3) == 12 && sizeof(uint3) == 12, "reference vertex/triangle layout");

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

#if !defined(MDMC_DMC_NO_GLOBAL_TYPES) && !defined(MDMC_DMC_HIP_VECTOR_TYPES)
using dmc::float2;
using dmc::float3;
using dmc::make_float2;
using dmc::make_float3;
using dmc::make_uint3;
using dmc::uint3;
#endif

// the stream output the reference declares for its vector types
// (include/utils.hpp:156-158)
inline std::ostream &operator<<(std::ostream &os, const dmc::float3 &f) {
  return os << "(" << f.x << ", " << f.y << ", " << f.z << ")";
}
inline std::ostream &operator<<(std::ostream &os, const dmc::uint3 &u) {
  return os << "(" << u.x << ", " << u.y << ", " << u.z << ")";
}

namespace utils {

// Same interface as the reference's Array3D (include/utils.hpp:14-66): x
// fastest, std::vector storage.
template <typename T>
class Array3D {
 public:
  typedef T value_type;
  typedef typename std::vector<value_type>::iterator iterator;
  typedef typename std::vector<value_type>::const_iterator const_iterator;

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
  iterator begin() { return data_.begin(); }
  iterator end() { return data_.end(); }
  const_iterator cbegin() const { return data_.cbegin(); }
  const_iterator cend() const { return data_.cend(); }
  T *data() { return data_.data(); }
  const T *data() const { return data_.data(); }
  size_t size() const { return data_.size(); }

 private:
  unsigned dx_, dy_, dz_;
  std::vector<T> data_;
};

// include/utils.hpp:68-96: x fastest, (k * nj + j) * ni + i
inline void index3D_to_1D(unsigned i, unsigned j, unsigned k, unsigned num_voxels_i, unsigned num_voxels_j,
                          unsigned &index1D) {
  index1D = (k * num_voxels_j + j) * num_voxels_i + i;
}
inline unsigned index3D_to_1D(unsigned i, unsigned j, unsigned k, unsigned num_voxels_i, unsigned num_voxels_j) {
  unsigned index1D;
  index3D_to_1D(i, j, k, num_voxels_i, num_voxels_j, index1D);
  return index1D;
}
inline void index3D_to_1D(const dmc::uint3 &index3D, const dmc::uint3 &num_voxels_dim, unsigned &index1D) {
  index3D_to_1D(index3D.x, index3D.y, index3D.z, num_voxels_dim.x, num_voxels_dim.y, index1D);
}
inline unsigned index3D_to_1D(const dmc::uint3 &index3D, const dmc::uint3 &num_voxels_dim) {
  unsigned index1D;
  index3D_to_1D(index3D, num_voxels_dim, index1D);
  return index1D;
}

// include/utils.hpp:98-116: the inverse
inline void index1D_to_3D(unsigned index1D, const dmc::uint3 &num_voxels_dim, dmc::uint3 &index3D) {
  const unsigned plane = num_voxels_dim.x * num_voxels_dim.y;
  index3D.z = index1D / plane;
  const unsigned rest = index1D % plane;
  index3D.y = rest / num_voxels_dim.x;
  index3D.x = rest % num_voxels_dim.x;
}
inline dmc::uint3 index1D_to_3D(unsigned index1D, const dmc::uint3 &num_voxels_dim) {
  dmc::uint3 index3D;
  index1D_to_3D(index1D, num_voxels_dim, index3D);
  return index3D;
}

// include/utils.hpp:118-122
inline float ijk_to_xyz(unsigned i, unsigned size, float f_range, float f_min) {
  return (float)i / size * f_range + f_min;
}

// include/utils.hpp:124-128: two components of a float3
inline dmc::float2 xy(const dmc::float3 &f) { return dmc::float2{f.x, f.y}; }
inline dmc::float2 yz(const dmc::float3 &f) { return dmc::float2{f.y, f.z}; }
inline dmc::float2 xz(const dmc::float3 &f) { return dmc::float2{f.x, f.z}; }

// include/utils.hpp:130-153: position of the first smallest (largest) of the
// arguments
template <typename Op, typename T>
size_t arg_best(const Op &, size_t best, size_t, const T &) {
  return best;
}
template <typename Op, typename T, typename... Ts>
size_t arg_best(const Op &better, size_t best, size_t at, const T &cur, const T &head, const Ts &...rest) {
  return better(head, cur) ? arg_best(better, at, at + 1, head, rest...) : arg_best(better, best, at + 1, cur, rest...);
}
template <typename T, typename... Ts>
size_t argmin(const T &head, const Ts &...rest) {
  return arg_best(std::less<T>(), 0, 1, head, rest...);
}
template <typename T, typename... Ts>
size_t argmax(const T &head, const Ts &...rest) {
  return arg_best(std::greater<T>(), 0, 1, head, rest...);
}

}  // namespace utils

namespace dmc {

// include/dmc.hpp:12-36
typedef uint8_t voxel_pt_index_type;
typedef uint8_t voxel_edge_index_type;
typedef uint8_t voxel_face_index_type;
typedef uint8_t voxel_config_type;
typedef unsigned voxel_index1D_type;
typedef uint8_t iso_vertex_m_type;
typedef unsigned vertex_index_type;
typedef unsigned flag_type;
typedef uint8_t check_dir_type;

const unsigned INVALID_UINT8 = 0xff;
const unsigned INVALID_UINT32 = 0xffffffff;
const unsigned VOXEL_NUM_PTS = 8;
const unsigned VOXEL_NUM_EDGES = 12;
const unsigned VOXEL_NUM_FACES = 6;
const unsigned NUM_CONFIGS = 256;
const voxel_index1D_type INVALID_INDEX_1D = INVALID_UINT32;
const voxel_config_type MAX_VOXEL_CONFIG_MASK = INVALID_UINT8;
const iso_vertex_m_type NO_VERTEX = INVALID_UINT8;

typedef utils::Array3D<float> scalar_grid_type;

// dmc::setup_device_luts (include/dmc.hpp:40)
inline void setup_device_luts() {
  const int rc = mdmc_setup();
  if (rc != MDMC_OK) throw error(rc, "setup_device_luts");
}

// dmc::setup_device_luts with the tables' device pointers handed back
// (include/dmc.hpp:42-46); the library owns them.
inline void setup_device_luts(iso_vertex_m_type **d_config_edge_lut1, iso_vertex_m_type **d_config_edge_lut2,
                              uint8_t **d_num_vertex_lut1, uint8_t **d_num_vertex_lut2,
                              voxel_config_type **d_config_2B_3B_lut,
                              voxel_face_index_type **d_config_2B_3B_ambiguous_face,
                              voxel_face_index_type **d_opposite_face_lut, check_dir_type **d_face_to_check_dir_lut,
                              uint8_t **d_edge_belonged_voxel_lut, voxel_edge_index_type **d_circular_edge_lut,
                              voxel_edge_index_type **d_voxel_local_edges) {
  mdmc_device_luts t;
  const int rc = mdmc_setup_device_luts(&t);
  if (rc != MDMC_OK) throw error(rc, "setup_device_luts");
  *d_config_edge_lut1 = t.config_edge_lut1;
  *d_config_edge_lut2 = t.config_edge_lut2;
  *d_num_vertex_lut1 = t.num_vertex_lut1;
  *d_num_vertex_lut2 = t.num_vertex_lut2;
  *d_config_2B_3B_lut = t.config_2B_3B_lut;
  *d_config_2B_3B_ambiguous_face = t.config_2B_3B_ambiguous_face;
  *d_opposite_face_lut = t.opposite_face_lut;
  *d_face_to_check_dir_lut = t.face_to_check_dir_lut;
  *d_edge_belonged_voxel_lut = t.edge_belonged_voxel_lut;
  *d_circular_edge_lut = t.circular_edge_lut;
  *d_voxel_local_edges = t.voxel_local_edges;
}

// dmc::cleanup_device_luts (include/dmc.hpp:48)
inline void cleanup_device_luts() {
  const int rc = mdmc_cleanup();
  if (rc != MDMC_OK) throw error(rc, "cleanup_device_luts");
}

// dmc::cleanup_device_luts with the pointers the 11-pointer setup handed out
// (include/dmc.hpp:50-54)
inline void cleanup_device_luts(iso_vertex_m_type *d_config_edge_lut1, iso_vertex_m_type *d_config_edge_lut2,
                                uint8_t *d_num_vertex_lut1, uint8_t *d_num_vertex_lut2,
                                voxel_config_type *d_config_2B_3B_lut,
                                voxel_face_index_type *d_config_2B_3B_ambiguous_face,
                                voxel_face_index_type *d_opposite_face_lut, check_dir_type *d_face_to_check_dir_lut,
                                uint8_t *d_edge_belonged_voxel_lut, voxel_edge_index_type *d_circular_edge_lut,
                                voxel_edge_index_type *d_voxel_local_edges) {
  const mdmc_device_luts t = {d_config_edge_lut1,       d_config_edge_lut2,        d_num_vertex_lut1,
                              d_num_vertex_lut2,        d_config_2B_3B_lut,        d_config_2B_3B_ambiguous_face,
                              d_opposite_face_lut,      d_face_to_check_dir_lut,   d_edge_belonged_voxel_lut,
                              d_circular_edge_lut,      d_voxel_local_edges};
  const int rc = mdmc_cleanup_device_luts(&t);
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
