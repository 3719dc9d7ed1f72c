// Host check of include/dmc.hpp's utilities (no device call): the reference's
// public constants and typedefs (include/dmc.hpp:12-36), Array3D iterators
// and element order (include/utils.hpp:14-66), index3D_to_1D /
// index1D_to_3D against brute-force enumeration (utils.hpp:68-116),
// ijk_to_xyz, the float2 views and argmin / argmax (utils.hpp:118-153).
// Prints OK and exits 0 when every check holds.
#include <cstdio>
#include <numeric>
#include <type_traits>

#include "dmc.hpp"

static int failures = 0;
#define CHECK(cond)                                            \
  do {                                                         \
    if (!(cond)) {                                             \
      std::printf("FAIL line %d: %s\n", __LINE__, #cond);      \
      ++failures;                                              \
    }                                                          \
  } while (0)

int main() {
  using namespace dmc;
  static_assert(std::is_same<voxel_config_type, uint8_t>::value, "config type");
  static_assert(std::is_same<voxel_index1D_type, unsigned>::value, "index type");
  static_assert(std::is_same<vertex_index_type, unsigned>::value, "vertex index type");
  static_assert(std::is_same<flag_type, unsigned>::value, "flag type");
  CHECK(INVALID_UINT8 == 0xffu && INVALID_UINT32 == 0xffffffffu);
  CHECK(VOXEL_NUM_PTS == 8 && VOXEL_NUM_EDGES == 12 && VOXEL_NUM_FACES == 6 && NUM_CONFIGS == 256);
  CHECK(INVALID_INDEX_1D == 0xffffffffu && MAX_VOXEL_CONFIG_MASK == 0xff && NO_VERTEX == 0xff);

  // Array3D: x fastest, iterators walk the storage in that order
  utils::Array3D<float> a(make_uint3(5, 4, 3), 0.0f);
  CHECK(a.size() == 60 && a.dim_x() == 5 && a.dim_y() == 4 && a.dim_z() == 3);
  float v = 0.0f;
  for (auto it = a.begin(); it != a.end(); ++it) *it = v++;
  CHECK(a(0, 0, 0) == 0.0f && a(1, 0, 0) == 1.0f && a(0, 1, 0) == 5.0f && a(0, 0, 1) == 20.0f);
  CHECK(a(4, 3, 2) == 59.0f);
  const utils::Array3D<float> &ca = a;
  CHECK(std::accumulate(ca.cbegin(), ca.cend(), 0.0f) == 59.0f * 60.0f / 2.0f);
  CHECK(ca.data() == &ca(0, 0, 0));

  // index helpers against enumeration order
  const uint3 dims = make_uint3(7, 3, 4);
  unsigned n = 0;
  for (unsigned k = 0; k < dims.z; ++k)
    for (unsigned j = 0; j < dims.y; ++j)
      for (unsigned i = 0; i < dims.x; ++i, ++n) {
        unsigned i1 = 12345;
        utils::index3D_to_1D(i, j, k, dims.x, dims.y, i1);
        CHECK(i1 == n);
        CHECK(utils::index3D_to_1D(i, j, k, dims.x, dims.y) == n);
        CHECK(utils::index3D_to_1D(make_uint3(i, j, k), dims) == n);
        uint3 b;
        utils::index1D_to_3D(n, dims, b);
        CHECK(b.x == i && b.y == j && b.z == k);
        const uint3 c = utils::index1D_to_3D(n, dims);
        CHECK(c.x == i && c.y == j && c.z == k);
      }

  // coordinates: i / size * range + min in float
  CHECK(utils::ijk_to_xyz(0, 20, 10.0f, -5.0f) == -5.0f);
  CHECK(utils::ijk_to_xyz(20, 20, 10.0f, -5.0f) == 5.0f);
  CHECK(utils::ijk_to_xyz(3, 7, 2.0f, 1.0f) == (float)3 / 7 * 2.0f + 1.0f);

  const float3 p = make_float3(1.0f, 2.0f, 3.0f);
  CHECK(utils::xy(p).x == 1.0f && utils::xy(p).y == 2.0f);
  CHECK(utils::yz(p).x == 2.0f && utils::yz(p).y == 3.0f);
  CHECK(utils::xz(p).x == 1.0f && utils::xz(p).y == 3.0f);
  const float3 d = make_float3(5, 5, 5) - make_float3(-5, -5, -5);
  CHECK(d.x == 10.0f && d.y == 10.0f && d.z == 10.0f);

  // first smallest / largest wins
  CHECK(utils::argmin(3, 1, 2) == 1);
  CHECK(utils::argmin(1, 1, 0, 0) == 2);
  CHECK(utils::argmax(3, 7, 7, 1) == 1);
  CHECK(utils::argmax(5.0f) == 0);
  CHECK(utils::argmin(2.0f, 4.0f, -1.0f, 6.0f) == 2);

  if (failures) return 1;
  std::printf("OK\n");
  return 0;
}
