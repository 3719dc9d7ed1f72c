/* Prints sizeof / offsetof of the C ABI structs for the ctypes layout test. */
#include <stddef.h>
#include <stdio.h>

#include "mdmc.h"

#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m))
#define S(T) printf(#T " %zu\n", sizeof(T))

int main(void) {
  S(mdmc_counts); S(mdmc_mesh); S(mdmc_volume); S(mdmc_gen_spec); S(mdmc_server_config); S(mdmc_request);
  S(mdmc_result); S(mdmc_server_stats);
  F(mdmc_volume, field); F(mdmc_volume, iso); F(mdmc_volume, vertices); F(mdmc_volume, quad_capacity);
  F(mdmc_gen_spec, kind); F(mdmc_gen_spec, seed); F(mdmc_gen_spec, nblobs); F(mdmc_gen_spec, noise_amp);
  F(mdmc_request, field_kind); F(mdmc_request, field); F(mdmc_request, nblobs); F(mdmc_request, tag);
  F(mdmc_result, status); F(mdmc_result, counts); F(mdmc_result, vertices); F(mdmc_result, t_done);
  F(mdmc_server_stats, busy_seconds);
  return 0;
}
