// A program written against include/dmc.hpp the way the reference's own
// harness is written (test_dmc, DMC_CUDA/DMC_CUDA/src/dmc.cu:2248-2306): the
// 11-pointer table setup, a sphere field sqrt(x^2+y^2+z^2) on 21^3 points over
// [-5,5]^3 filled through Array3D and ijk_to_xyz, run_dmc at iso 4.1 with 15
// smoothing iterations, dump_obj, then the 11-pointer cleanup.
//
// Usage: dmc_harness OUT_DIR [RESOLUTION ISO NUM_SMOOTH]
// Writes into OUT_DIR: field.f32 (the grid, x fastest), vertices.f32,
// triangles.u32, sphere.obj, tables.bin (the 11 device tables read back, in
// argument order) and counts.txt ("V T").  Exit status 0 on success.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "dmc.hpp"

namespace {

template <typename T>
bool write_raw(const std::string &path, const T *data, size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(data), (std::streamsize)(n * sizeof(T)));
  return (bool)f;
}

bool read_table(std::vector<unsigned char> &out, const void *dptr, size_t n) {
  const size_t at = out.size();
  out.resize(at + n);
  return hipMemcpy(out.data() + at, dptr, n, hipMemcpyDeviceToHost) == hipSuccess;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s OUT_DIR [RESOLUTION ISO NUM_SMOOTH]\n", argv[0]);
    return 2;
  }
  const std::string out = argv[1];
  const unsigned resolution = argc > 2 ? (unsigned)std::atoi(argv[2]) : 20u;
  const float iso_value = argc > 3 ? (float)std::atof(argv[3]) : 4.1f;
  const unsigned num_smooth = argc > 4 ? (unsigned)std::atoi(argv[4]) : 15u;
  using namespace utils;
  using namespace dmc;
  try {
    iso_vertex_m_type *d_config_edge_lut1, *d_config_edge_lut2;
    uint8_t *d_num_vertex_lut1, *d_num_vertex_lut2;
    voxel_config_type *d_config_2B_3B_lut;
    voxel_face_index_type *d_config_2B_3B_ambiguous_face;
    voxel_face_index_type *d_opposite_face_lut;
    check_dir_type *d_face_to_check_dir_lut;
    uint8_t *d_edge_belonged_voxel_lut;
    voxel_edge_index_type *d_circular_edge_lut;
    voxel_edge_index_type *d_voxel_local_edges;
    setup_device_luts(&d_config_edge_lut1, &d_config_edge_lut2, &d_num_vertex_lut1, &d_num_vertex_lut2,
                      &d_config_2B_3B_lut, &d_config_2B_3B_ambiguous_face, &d_opposite_face_lut,
                      &d_face_to_check_dir_lut, &d_edge_belonged_voxel_lut, &d_circular_edge_lut,
                      &d_voxel_local_edges);

    std::vector<unsigned char> tables;
    bool ok = read_table(tables, d_config_edge_lut1, NUM_CONFIGS * VOXEL_NUM_EDGES) &&
              read_table(tables, d_config_edge_lut2, NUM_CONFIGS * VOXEL_NUM_EDGES) &&
              read_table(tables, d_num_vertex_lut1, NUM_CONFIGS) && read_table(tables, d_num_vertex_lut2, NUM_CONFIGS) &&
              read_table(tables, d_config_2B_3B_lut, 36) && read_table(tables, d_config_2B_3B_ambiguous_face, 36) &&
              read_table(tables, d_opposite_face_lut, VOXEL_NUM_FACES) &&
              read_table(tables, d_face_to_check_dir_lut, VOXEL_NUM_FACES) &&
              read_table(tables, d_edge_belonged_voxel_lut, VOXEL_NUM_EDGES) &&
              read_table(tables, d_circular_edge_lut, 12) && read_table(tables, d_voxel_local_edges, 3);
    if (!ok) {
      std::fprintf(stderr, "reading the device tables failed\n");
      return 1;
    }

    float3 xyz_min = make_float3(-5, -5, -5);
    float3 xyz_max = make_float3(5, 5, 5);
    float3 xyz_range = xyz_max - xyz_min;
    Array3D<float> scalar_grid(resolution + 1, resolution + 1, resolution + 1);
    for (unsigned k = 0; k < scalar_grid.dim_z(); ++k) {
      const float z = ijk_to_xyz(k, resolution, xyz_range.z, xyz_min.z);
      for (unsigned j = 0; j < scalar_grid.dim_y(); ++j) {
        const float y = ijk_to_xyz(j, resolution, xyz_range.y, xyz_min.y);
        for (unsigned i = 0; i < scalar_grid.dim_x(); ++i) {
          const float x = ijk_to_xyz(i, resolution, xyz_range.x, xyz_min.x);
          scalar_grid(i, j, k) = std::sqrt(x * x + y * y + z * z);
        }
      }
    }

    std::vector<float3> compact_vertices;
    std::vector<uint3> compact_triangles;
    run_dmc(compact_vertices, compact_triangles, scalar_grid, xyz_min, xyz_max, iso_value, num_smooth);
    dump_obj((out + "/sphere.obj").c_str(), compact_vertices, compact_triangles);

    cleanup_device_luts(d_config_edge_lut1, d_config_edge_lut2, d_num_vertex_lut1, d_num_vertex_lut2,
                        d_config_2B_3B_lut, d_config_2B_3B_ambiguous_face, d_opposite_face_lut,
                        d_face_to_check_dir_lut, d_edge_belonged_voxel_lut, d_circular_edge_lut,
                        d_voxel_local_edges);

    ok = write_raw(out + "/field.f32", scalar_grid.data(), scalar_grid.size()) &&
         write_raw(out + "/vertices.f32", &compact_vertices[0].x, 3 * compact_vertices.size()) &&
         write_raw(out + "/triangles.u32", &compact_triangles[0].x, 3 * compact_triangles.size()) &&
         write_raw(out + "/tables.bin", tables.data(), tables.size());
    std::ofstream(out + "/counts.txt") << compact_vertices.size() << " " << compact_triangles.size() << "\n";
    if (!ok) {
      std::fprintf(stderr, "writing the outputs failed\n");
      return 1;
    }
    std::printf("vertices %zu triangles %zu\n", compact_vertices.size(), compact_triangles.size());
  } catch (const dmc::error &e) {
    std::fprintf(stderr, "dmc::error: %s\n", e.what());
    return 1;
  }
  return 0;
}
