/*
 * A C caller of the drop-in boundary: the reference's example main
 * (source/main_example.cu:452-548) against libbpc_paco.so through
 * include/bpc_paco.h only.  It loads the sample image, loads the probability
 * tables from the reference's text files, codes the image in place with
 * bpc_run(BPC_CODE, ...) as CUDA_BPC(CODE, ...) does, checks the codestream against
 * an expected one when given, decodes it in place and checks the reconstruction
 * against the original (check_Image / Print_Results, main_example.cu:182-217).
 *
 * usage: bpc_main <image.i32> <size> <lut folder> <cpass3> [expected.i32 [codestream_out.i32]]
 * image.i32: size*size little-endian int32 coefficients (raster, already transformed).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bpc_paco.h"

static int read_file(const char* path, void* buf, size_t bytes) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  const size_t got = fread(buf, 1, bytes, f);
  fclose(f);
  return got == bytes ? 0 : -1;
}

static int write_file(const char* path, const void* buf, size_t bytes) {
  FILE* f = fopen(path, "wb");
  if (!f) return -1;
  const size_t put = fwrite(buf, 1, bytes, f);
  return (fclose(f) == 0 && put == bytes) ? 0 : -1;
}

/* First mismatch and count of mismatching words, and the codeblock of the first one
 * (check_Image, main_example.cu:182-202). */
static long compare(const int32_t* a, const int32_t* b, int size, long* first) {
  long wrong = 0;
  *first = -1;
  for (long k = 0; k < (long)size * size; ++k)
    if (a[k] != b[k]) {
      if (*first < 0) *first = k;
      ++wrong;
    }
  return wrong;
}

int main(int argc, char** argv) {
  if (argc < 5) {
    fprintf(stderr, "usage: %s <image.i32> <size> <lut folder> <cpass3> [expected.i32 [codestream_out.i32]]\n",
            argv[0]);
    return 2;
  }
  const int size = atoi(argv[2]);
  const int cpass3 = atoi(argv[4]);
  if (size <= 0 || size % 64 || (cpass3 != 0 && cpass3 != 1)) {
    fprintf(stderr, "bad size or cpass3\n");
    return 2;
  }
  const size_t bytes = (size_t)size * size * sizeof(int32_t);
  int32_t* input = (int32_t*)malloc(bytes);
  int32_t* original = (int32_t*)malloc(bytes);
  int32_t* expected = (int32_t*)malloc(bytes);
  if (!input || !original || !expected) return 3;
  if (read_file(argv[1], input, bytes)) {
    fprintf(stderr, "cannot read %s\n", argv[1]);
    return 3;
  }
  memcpy(original, input, bytes);

  printf("%s\n", cpass3 ? "+++ 3 Coding passes encoder +++" : "+++ 2 Coding passes encoder +++");
  int32_t* lut = NULL;
  int lut_len = 0, nbp = 0;
  if (bpc_lut_load_text(argv[3], cpass3, 5, &lut, &lut_len, &nbp) != BPC_OK) {
    fprintf(stderr, "load_LUTs: %s\n", bpc_last_error());
    return 4;
  }
  bpc_params p;
  bpc_default_params(&p); /* DWT_LEVELS 5, CBLOCK_LENGTH 64 (common.h) */
  p.cpass3 = cpass3;
  p.lut_n_bitplanes = nbp;
  if (bpc_set_config(&p) != BPC_OK) {
    fprintf(stderr, "config: %s\n", bpc_last_error());
    return 4;
  }

  int rc = bpc_run(BPC_CODE, input, size, size, lut);
  if (rc != BPC_OK) {
    fprintf(stderr, "CODE: status %d: %s\n", rc, bpc_last_error());
    return 5;
  }
  int ok = 1;
  if (argc > 5) {
    if (read_file(argv[5], expected, bytes)) {
      fprintf(stderr, "cannot read %s\n", argv[5]);
      return 3;
    }
    long first;
    const long wrong = compare(input, expected, size, &first);
    if (wrong) {
      printf("TEST ENCODER ERROR: %ld words differ, first at %ld (codeblock %ld)\n", wrong, first,
             first / (64 * 64));
      ok = 0;
    } else {
      printf("TEST ENCODER OK\n");
    }
    if (argc > 6 && write_file(argv[6], input, bytes)) {
      fprintf(stderr, "cannot write %s\n", argv[6]);
      return 3;
    }
  }

  rc = bpc_run(BPC_DECODE, input, size, size, lut);
  if (rc != BPC_OK) {
    fprintf(stderr, "DECODE: status %d: %s\n", rc, bpc_last_error());
    return 5;
  }
  long first;
  const long wrong = compare(input, original, size, &first);
  if (wrong) {
    printf("TEST DECODER ERROR: %ld samples differ, first at (%ld, %ld)\n", wrong, first % size, first / size);
    ok = 0;
  } else {
    printf("TEST DECODER %dx%d OK\n", size, size);
  }
  bpc_lut_free(lut);
  free(input);
  free(original);
  free(expected);
  return ok ? 0 : 1;
}
