/*
 * Sizes and field offsets of every structure in include/bpc_paco.h, printed as
 * JSON, so that the Python mirror's ctypes structures can be checked against the
 * compiler's layout (tests/test_abi_layout.py).
 */
#include <stddef.h>
#include <stdio.h>

#include "bpc_paco.h"

#define FIELD(T, f) printf("%s\"%s\": %zu", first++ ? ", " : "", #f, offsetof(T, f))
#define BEGIN(T)                                      \
  do {                                                \
    int first = 0;                                    \
    printf("%s\"%s\": {\"sizeof\": %zu, \"fields\": {", nstruct++ ? ", " : "", #T, sizeof(T));
#define END() \
  printf("}}"); \
  } while (0)

int main(void) {
  int nstruct = 0;
  printf("{");
  BEGIN(bpc_params);
  FIELD(bpc_params, cpass3); FIELD(bpc_params, dwt_levels); FIELD(bpc_params, cblock_length);
  FIELD(bpc_params, lut_n_bitplanes); FIELD(bpc_params, cb_stride_words); FIELD(bpc_params, layout);
  END();
  BEGIN(bpc_image);
  FIELD(bpc_image, src); FIELD(bpc_image, dst); FIELD(bpc_image, lut); FIELD(bpc_image, lut_len);
  FIELD(bpc_image, ncw); FIELD(bpc_image, size_x); FIELD(bpc_image, size_y); FIELD(bpc_image, params);
  FIELD(bpc_image, fill_unused); FIELD(bpc_image, int16_coeffs);
  END();
  BEGIN(bpc_server_config);
  FIELD(bpc_server_config, max_inflight); FIELD(bpc_server_config, max_batch_images);
  FIELD(bpc_server_config, max_batch_codeblocks); FIELD(bpc_server_config, hbm_budget_bytes);
  FIELD(bpc_server_config, hbm_reserve_bytes); FIELD(bpc_server_config, pool_limit_bytes);
  FIELD(bpc_server_config, waves_per_block); FIELD(bpc_server_config, vgpr_cap_waves);
  FIELD(bpc_server_config, mem_info); FIELD(bpc_server_config, mem_info_ctx);
  END();
  BEGIN(bpc_request);
  FIELD(bpc_request, direction); FIELD(bpc_request, src); FIELD(bpc_request, lut); FIELD(bpc_request, lut_len);
  FIELD(bpc_request, size_x); FIELD(bpc_request, size_y); FIELD(bpc_request, params);
  FIELD(bpc_request, int16_coeffs);
  END();
  BEGIN(bpc_result);
  FIELD(bpc_result, status); FIELD(bpc_result, output); FIELD(bpc_result, ncw); FIELD(bpc_result, queue_us);
  FIELD(bpc_result, latency_us);
  END();
  BEGIN(bpc_server_stats_t);
  FIELD(bpc_server_stats_t, batches); FIELD(bpc_server_stats_t, completed); FIELD(bpc_server_stats_t, pending);
  FIELD(bpc_server_stats_t, committed_bytes); FIELD(bpc_server_stats_t, max_batch_images);
  END();
  printf("}\n");
  return 0;
}
