/*
 * bpc_paco.h -- C-ABI of the MI355X BPC-PaCo codec (libbpc_paco.so).
 *
 * Drop-in boundary for the reference's host entry point
 *     void CUDA_BPC(int Direction, DATATYPE* HData, int HDSize_X, int HDSize_Y, int* LUT)
 * (source/BPC_PaCo.cu:1305-1325, declared for callers at source/main_example.cu:8)
 * and for the LUT loader load_LUTs (source/main_example.cu:219-448).
 *
 * Plain pointers and sizes only.  Every function returns BPC_OK (0) or a negative
 * status; bpc_last_error() gives a message for the calling thread.
 */
#ifndef BPC_PACO_H
#define BPC_PACO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference prints and exit()s instead, common.cu:3-8) ---- */
#define BPC_OK 0
#define BPC_ERR_ARG (-1)       /* invalid argument or configuration */
#define BPC_ERR_HIP (-2)       /* HIP runtime error (allocation, copy, launch) */
#define BPC_ERR_OVERFLOW (-3)  /* a codeblock needs more codeword slots than its region holds */
#define BPC_ERR_RANGE (-4)     /* |coefficient| >= 2^28, LUT entry outside [0,127], or LUT row outside the array */
#define BPC_ERR_MALFORMED (-5) /* decoder: MSB word of a region not in [0,27] and not 32 */
#define BPC_ERR_WIDE (-6)      /* device-resident call: a codeblock's MSB is beyond the kernel's bitplane
                                  registers (only when MSB >= lut_n_bitplanes); bpc_run_ex retries itself */
#define BPC_ERR_NODEVICE (-7)  /* no HIP device */
#define BPC_ERR_TIMEOUT (-8)   /* bpc_server_wait: the request did not finish in time */

/* ---- direction (enum BPC, common.h:8-11) ---- */
#define BPC_CODE 0
#define BPC_DECODE 1

/* ---- coefficient layouts ---- */
#define BPC_LAYOUT_RASTER 0 /* row-major X*Y image (the reference's layout) */
#define BPC_LAYOUT_TILED 1  /* codeblock-contiguous: codeblock t (row-major tile order) at t*64*rows,
                               row-major inside the codeblock */

/* Runtime form of the reference's compile-time configuration (common.h:19-51). */
typedef struct bpc_params {
  int32_t cpass3;          /* CPASS_3 (common.h:24): 0 = SPP+MRP, 1 = SPP+MRP+CP. Reference default 0 */
  int32_t dwt_levels;      /* DWT_LEVELS (common.h:19). Default 5 */
  int32_t cblock_length;   /* CBLOCK_LENGTH (common.h:45): codeblock rows, 2..64. Default 64 */
  int32_t lut_n_bitplanes; /* LUT_N_BITPLANES (common.h:28): LUT rows per group. Default 15 */
  int32_t cb_stride_words; /* int32 words per codeblock region, 2..2^26; 0 = 64*cblock_length, the
                              reference's layout (BPC_PaCo.cu:1130), in which the codestream has the
                              image's size */
  int32_t layout;          /* BPC_LAYOUT_RASTER (default) or BPC_LAYOUT_TILED */
} bpc_params;

/* Reference defaults: CPASS_3=0, DWT_LEVELS=5, CBLOCK_LENGTH=64, LUT_N_BITPLANES=15. */
void bpc_default_params(bpc_params* p);
/* LUT length in ints: (3*L+1)*NBP*(1+9+4), plus (3*L+1)*NBP*(9+4) when cpass3
 * (refinement_N + significance_N + sign_N [+ significance_N + sign_N], main_example.cu:266-278). */
int bpc_lut_len(const bpc_params* p);
/* int32 words of one codestream buffer for an image: codeblocks * stride. */
int64_t bpc_stream_words(int size_x, int size_y, const bpc_params* p);
const char* bpc_last_error(void);
int bpc_device_count(void);

/* ---- drop-in entry point ------------------------------------------------------
 * Replaces CUDA_BPC (BPC_PaCo.cu:1305-1325).  hdata is a host buffer of
 * size_x*size_y int32, overwritten in place: CODE turns two's-complement
 * coefficients into per-codeblock codestream regions (word 0 = MSB, then the
 * codewords, unused words = -1); DECODE turns the regions back into coefficients.
 * d_lut is a device LUT (from bpc_lut_load_text / bpc_lut_upload).
 * Deviation: the reference frees d_lut (BPC_PaCo.cu:1300); this library leaves it
 * to the caller (bpc_lut_free).  Uses the configuration set by bpc_set_config
 * (reference defaults until then). */
int bpc_run(int direction, int32_t* hdata, int size_x, int size_y, const int32_t* d_lut);
int bpc_set_config(const bpc_params* p);
int bpc_get_config(bpc_params* p);
/* Same with explicit configuration and LUT length.  With a non-default
 * cb_stride_words the codestream is bpc_stream_words() words: hdata must hold
 * max(size_x*size_y, bpc_stream_words()) int32. */
int bpc_run_ex(int direction, int32_t* hdata, int size_x, int size_y, const int32_t* d_lut, int lut_len,
               const bpc_params* p);

/* ---- probability LUTs (main_example.cu:219-448) ---- */
/* Loads ref.txt, sig.txt, sign.txt (+ cp_sig.txt, cp_sign.txt when cpass3) from
 * folder with the reference loader's layout and padding rule and uploads them.
 * *nbp receives max_bitplane+1 (the LUT_N_BITPLANES the table was laid out for). */
int bpc_lut_load_text(const char* folder, int cpass3, int dwt_levels, int32_t** d_lut, int* lut_len, int* nbp);
/* Same, into a host array of capacity cap ints (no device needed). */
int bpc_lut_parse_text(const char* folder, int cpass3, int dwt_levels, int32_t* host_lut, int cap, int* lut_len,
                       int* nbp);
int bpc_lut_upload(const int32_t* host_lut, int lut_len, int32_t** d_lut);
int bpc_lut_free(int32_t* d_lut);
/* Builds a LUT by counting symbol-0 frequencies of every (pass, level, subband,
 * bitplane, context) over the codeblocks of a training image (raster, two's
 * complement), with the coder's own context modelling.  P = clamp(round(128*n0/n), 1, 127),
 * 64 where a row was never visited.  nthreads <= 0: all host threads. */
int bpc_lut_train(const int32_t* img, int size_x, int size_y, const bpc_params* p, int32_t* host_lut, int lut_len,
                  int nthreads);
/* The same trainer on the device, asynchronous on `stream`: d_img is a training
 * image of size_x*size_y int32 in p->layout (raster or codeblock-contiguous).
 * Writes bpc_lut_len(p) entries to each non-null output: d_lut (the table, ready
 * for bpc_encode_device), d_n0 / d_n (per LUT entry: symbols 0 / all symbols counted).
 * Codeblocks with a magnitude >= 2^28 are skipped, as by bpc_lut_train.  Uses the
 * library's per-device scratch (ordered with the codec launches that share it).
 * No counterpart in the reference, which ships trained tables only
 * (main_example.cu:219-448 reads them). */
int bpc_lut_train_device(const int32_t* d_img, int size_x, int size_y, const bpc_params* p, int32_t* d_lut,
                         uint32_t* d_n0, uint32_t* d_n, void* stream);

/* ---- device-resident codec (no host copies; asynchronous on `stream`) ----------
 * d_coeffs: size_x*size_y int32 in p->layout; d_stream: bpc_stream_words() int32.
 * d_ncw (nullable): codeword slots used per codeblock (encoder).  d_status: a device
 * uint32 that collects error bits (nullable: a library-owned word); read and clear
 * it with bpc_status_take(), which synchronizes `stream`.  The encoder writes
 * every word of d_stream when fill_unused != 0 (reference layout), otherwise only
 * word 0 and the used codeword slots of each region. */
int bpc_encode_device(const int32_t* d_coeffs, int32_t* d_stream, int size_x, int size_y, const int32_t* d_lut,
                      int lut_len, const bpc_params* p, int fill_unused, int32_t* d_ncw, uint32_t* d_status,
                      void* stream);
int bpc_decode_device(const int32_t* d_stream, int32_t* d_coeffs, int size_x, int size_y, const int32_t* d_lut,
                      int lut_len, const bpc_params* p, uint32_t* d_status, void* stream);
/* The same two calls with int16 coefficients (size_x*size_y int16 in p->layout):
 * half the coefficient bytes for data of at most 15 magnitude bits.  The
 * codestream is the same as from int32 coefficients of equal value.  The decoder
 * saturates values outside [-32768, 32767] and reports BPC_ERR_RANGE.  No
 * counterpart in the reference, whose 16/32-bit switch is unused
 * (DATATYPE_16BITS_or_32BITS, common.h:110-141: both branches are int). */
int bpc_encode_device_i16(const int16_t* d_coeffs, int32_t* d_stream, int size_x, int size_y, const int32_t* d_lut,
                          int lut_len, const bpc_params* p, int fill_unused, int32_t* d_ncw, uint32_t* d_status,
                          void* stream);
int bpc_decode_device_i16(const int32_t* d_stream, int16_t* d_coeffs, int size_x, int size_y, const int32_t* d_lut,
                          int lut_len, const bpc_params* p, uint32_t* d_status, void* stream);
/* No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  bpc_encode_device (direction BPC_CODE) or bpc_decode_device
 * (BPC_DECODE, where fill_unused and d_ncw are ignored) with the same results, but
 * no call to hipMalloc, hipFree or any allocating routine.  Images whose waves all
 * fit on the device at once keep their bitplanes in registers and use no scratch;
 * larger ones use a fixed array of BPC_NOALLOC_SCRATCH_BYTES per device (held until
 * the process ends) for the bitplane slots, and the launch uses at most as many
 * waves as it has slots for (16 KiB per wave for up to 16 LUT bitplanes, 20 KiB up
 * to 20, 28 KiB beyond; any image size is accepted, larger ones take more
 * codeblocks per wave).  d_status is required.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams, or from different threads) race and must be ordered by the
 * caller.  No part of the array needs clearing: each launch writes every slot word
 * it reads. */
#define BPC_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int bpc_codec_device_noalloc(int direction, const int32_t* d_src, int32_t* d_dst, int size_x, int size_y,
                             const int32_t* d_lut, int lut_len, const bpc_params* p, int fill_unused,
                             int32_t* d_ncw, uint32_t* d_status, void* stream);
/* Synchronizes stream, maps the status bits to a BPC_* code and clears them. */
int bpc_status_take(uint32_t* d_status, void* stream);

/* ---- compact codestream container (no counterpart in the reference) -------------
 * The reference layout gives every codeblock a fixed region of int32 words
 * [MSB, codewords..., -1...] (BPC_PaCo.cu:1130,1147-1157,1284).  The compact form
 * keeps per codeblock one MSB byte (255 marks a malformed MSB word) and its
 * codewords as uint16 at payload[offsets[cb] .. offsets[cb+1]), offsets[0] = 0.
 * All three calls are asynchronous on `stream`; errors go to d_status
 * (nullable: the library's status word), read with bpc_status_take.
 *
 * Index: d_msb[ncb] and d_offsets[ncb+1].  Codeword counts come from d_ncw (the
 * encoder's per-codeblock counts) or, when d_ncw is NULL, from the -1 fill: the
 * words before the first -1 after the MSB (regions must then be -1-filled). */
int bpc_compact_index_device(const int32_t* d_stream, int64_t ncb, int stride_words, const int32_t* d_ncw,
                             uint8_t* d_msb, uint64_t* d_offsets, uint32_t* d_status, void* stream);
/* Pack: codewords of every region into d_payload (payload_cap uint16 entries;
 * offsets[ncb] of them are written). */
int bpc_compact_pack_device(const int32_t* d_stream, int64_t ncb, int stride_words, const uint64_t* d_offsets,
                            uint16_t* d_payload, uint64_t payload_cap, uint32_t* d_status, void* stream);
/* Unpack: back to the reference layout (MSB word, codewords, -1 fill); a codeblock
 * whose codewords do not fit its region keeps only its MSB (BPC_ERR_OVERFLOW). */
int bpc_compact_unpack_device(const uint8_t* d_msb, const uint64_t* d_offsets, const uint16_t* d_payload,
                              uint64_t payload_len, int64_t ncb, int stride_words, int32_t* d_stream,
                              uint32_t* d_status, void* stream);
/* Decode straight from the compact form into size_x*size_y int32 coefficients in
 * p->layout (no regions in between): the result, including errors, is that of
 * bpc_compact_unpack_device with p's region size followed by bpc_decode_device.
 * d_msb and d_offsets cover the image's codeblocks; d_payload may be NULL when
 * payload_len is 0.  The counterpart of the reference decoder's region reads
 * (BPC_PaCo.cu:1216 MSB, :304-316 codewords), with the regions left out. */
int bpc_decode_compact_device(const uint8_t* d_msb, const uint64_t* d_offsets, const uint16_t* d_payload,
                              uint64_t payload_len, int32_t* d_coeffs, int size_x, int size_y, const int32_t* d_lut,
                              int lut_len, const bpc_params* p, uint32_t* d_status, void* stream);

/* ---- per-codeblock codestream files (main_example.cu:51-179) -------------------
 * File k.bin belongs to tile tile_of_file[k] in load_Data's order (LL band, then
 * per level from the coarsest HL, LH, HH, each row-major).  Returns the number of
 * files, or BPC_ERR_ARG when a subband is not a whole number of codeblocks. */
int bpc_cb_file_order(int size_x, int size_y, int dwt_levels, int cblock_length, int32_t* tile_of_file, int cap);
/* Writes folder/k.bin for every codeblock from a host copy of the compact form:
 * the MSB and the codewords as big-endian 16-bit values (2*(1+n) bytes). */
int bpc_bin_export(const char* folder, const uint8_t* msb, const uint64_t* offsets, const uint16_t* payload,
                   int size_x, int size_y, const bpc_params* p);
/* load_Data with CStreamDataLoad (main_example.cu:71-179) and load_CStream (:51-68):
 * fills every region of host_stream (bpc_stream_words int32) from folder/k.bin,
 * big-endian 16-bit words, -1 past the end of each file. */
int bpc_bin_import(const char* folder, int size_x, int size_y, const bpc_params* p, int32_t* host_stream);

/* ---- reversible 5/3 wavelet transform (no counterpart in the reference) ----------
 * The step before encoding and after decoding: the reference codes coefficients
 * that arrive transformed (its sample image, main_example.cu:71-179).  Integer
 * lifting with whole-sample symmetric extension (ISO/IEC 15444-1 Annex F): per
 * level, columns then rows of the current LL band, bands stored in the Mallat
 * layout Find_Subband (BPC_PaCo.cu:84-111) reads (HL right, LH below, HH diagonal).
 * Raster int32 size_x*size_y in and out (no overlap); size_x and size_y must be
 * divisible by 2^levels.  Asynchronous on `stream`; uses the library's scratch. */
int bpc_dwt53_forward_device(const int32_t* d_img, int32_t* d_coeffs, int size_x, int size_y, int levels,
                             void* stream);
int bpc_dwt53_inverse_device(const int32_t* d_coeffs, int32_t* d_img, int size_x, int size_y, int levels,
                             void* stream);

/* ---- batched codec (many images of uneven size in one persistent launch) --------
 * The codeblocks of all images go through one device work queue: sorted by MSB
 * inside each image and paired there (one wave64 codes two codeblocks of one
 * image), then ordered by pass mode and cost, costliest first, and drained by a
 * persistent kernel.  Each image's output is identical to bpc_encode_device /
 * bpc_decode_device on that image alone.  No counterpart in the reference, which
 * codes one image per call (CUDA_BPC, BPC_PaCo.cu:1305). */
typedef struct bpc_image {
  const int32_t* src; /* encoder: coefficients (params.layout) ; decoder: codestream regions */
  int32_t* dst;       /* encoder: codestream regions ; decoder: coefficients */
  const int32_t* lut; /* device LUT of this image */
  int32_t lut_len;
  int32_t* ncw;       /* encoder: codeword slots per codeblock (nullable) */
  int32_t size_x, size_y;
  bpc_params params;  /* per image; cblock_length must be the same for every image of a batch */
  int32_t fill_unused; /* encoder: -1 in unused region words */
  int32_t int16_coeffs; /* 1: the coefficients (encoder src, decoder dst) are int16_t, as for
                           bpc_encode_device_i16 / bpc_decode_device_i16; 0: int32_t */
} bpc_image;

typedef struct bpc_batch bpc_batch;
/* Device work-queue buffers for up to max_images (<= 4096) images / max_codeblocks
 * codeblocks; an image in a batch holds at most 2^20 - 1 codeblocks. */
int bpc_batch_create(bpc_batch** batch, int max_images, long long max_codeblocks);
int bpc_batch_destroy(bpc_batch* batch);
/* Launch shape of the persistent kernel: waves per workgroup (1, 2 or 4, i.e. 2, 4
 * or 8 codeblocks per workgroup) and resident waves per SIMD (0 = as many as fit). */
int bpc_batch_set_tuning(bpc_batch* batch, int waves_per_block, int waves_per_simd);
/* Register budget of the persistent kernel: 0 = the compiler's (4 waves per SIMD
 * for the coder), 5..8 = registers capped so that that many waves fit per SIMD
 * (separately compiled variants of the 16- and 20-plane kernels; the 28-plane
 * kernel keeps the compiler's budget).  The results are the same. */
int bpc_batch_set_vgpr_cap(bpc_batch* batch, int waves_per_simd);
/* Encodes (BPC_CODE) or decodes (BPC_DECODE) all images, asynchronously on stream.
 * Codeblocks whose MSB needs more bitplanes than the images' lut_n_bitplanes
 * provide for are coded by a second, 28-plane pass, so every legal codeblock is
 * coded.  use_graph != 0 replays a captured hipGraph of the launch sequence: the
 * kernels read the run's shape (image count, geometry, parameters, codeblock
 * counts, buffer pointers, LUTs, d_status) from words uploaded per run, so one
 * graph per (direction, bitplane count of the widest image, tuning) serves
 * batches of any composition; it is captured on first use, at most 32 are kept
 * per batch object and the least recently used is dropped.  Errors land in
 * d_status (bpc_status_take). */
int bpc_batch_run(bpc_batch* batch, int direction, const bpc_image* images, int n_images, int use_graph,
                  uint32_t* d_status, void* stream);
/* Number of captured graphs the batch object holds. */
int bpc_batch_graph_count(bpc_batch* batch);

/* ---- continuous-batching server ------------------------------------------------
 * Encode / decode requests on device-resident inputs are queued on arrival.  A
 * scheduler thread admits them in arrival order while the device's free memory
 * (less hbm_reserve_bytes) and the optional output budget allow their output
 * buffers, groups admitted requests of one direction into a batch (the batched
 * codec above) and launches it on one of max_inflight streams, so requests that
 * arrive while others are on the GPU start as soon as a slot and memory free up
 * and their work overlaps on the device.  Outputs are allocated by the server and
 * stay in HBM until bpc_server_release.  No counterpart in the reference (one
 * synchronous image per CUDA_BPC call, which allocates its buffers per call,
 * BPC_PaCo.cu:1277-1289). */
/* Device memory probe: fills free and total bytes, returns 0 on success. */
typedef int (*bpc_mem_info_fn)(void* ctx, size_t* free_bytes, size_t* total_bytes);
typedef struct bpc_server_config {
  int32_t max_inflight;         /* batches on the device at once (streams), 1..8 */
  int32_t max_batch_images;     /* images per batch, 1..4096 */
  long long max_batch_codeblocks; /* codeblocks per batch; a larger request is refused at submit */
  size_t hbm_budget_bytes;      /* cap on output bytes held by admitted, unreleased requests; 0 = none */
  size_t hbm_reserve_bytes;     /* device memory admission leaves free (batch workspaces, other users) */
  size_t pool_limit_bytes;      /* released output buffers kept for reuse, at most this many bytes */
  int32_t waves_per_block;      /* tuning of the batched kernel (1, 2 or 4) */
  int32_t vgpr_cap_waves;       /* its register budget: 0 or 5..8 (bpc_batch_set_vgpr_cap) */
  bpc_mem_info_fn mem_info;     /* NULL: hipMemGetInfo; a caller may inject its own figure */
  void* mem_info_ctx;
} bpc_server_config;

typedef struct bpc_request {
  int32_t direction;   /* BPC_CODE or BPC_DECODE */
  const int32_t* src;  /* device: coefficients (encode) or codestream regions (decode) */
  const int32_t* lut;
  int32_t lut_len;
  int32_t size_x, size_y;
  bpc_params params;
  int32_t int16_coeffs; /* 1: coefficients are int16_t (encode input, decode output), else int32_t */
} bpc_request;

typedef struct bpc_result {
  int32_t status;      /* BPC_OK or an error of the request's batch */
  void* output;        /* device: codestream regions (encode) or coefficients (decode) */
  int32_t* ncw;        /* device: codeword slots per codeblock (encode) */
  double queue_us;     /* submit -> batch launch */
  double latency_us;   /* submit -> completion seen by the scheduler */
} bpc_result;

typedef struct bpc_server_stats_t {
  unsigned long long batches, completed, pending;
  size_t committed_bytes, max_batch_images;
} bpc_server_stats_t;

typedef struct bpc_server bpc_server;
void bpc_server_default_config(bpc_server_config* cfg);
int bpc_server_create(bpc_server** server, const bpc_server_config* cfg);
/* Refuses (BPC_ERR_ARG) a request that could never run: bad parameters or geometry,
 * more codeblocks than max_batch_codeblocks, a short LUT, or an output larger than
 * the budget or the device less the reserve. */
int bpc_server_submit(bpc_server* server, const bpc_request* req, uint64_t* id);
/* Blocks until request id finished (timeout_ms < 0: no limit); fills res. */
int bpc_server_wait(bpc_server* server, uint64_t id, int timeout_ms, bpc_result* res);
/* Returns the request's output buffers to the server (after bpc_server_wait). */
int bpc_server_release(bpc_server* server, uint64_t id);
int bpc_server_stats(bpc_server* server, bpc_server_stats_t* stats);
int bpc_server_destroy(bpc_server* server);

/* ---- device memory / stream / event helpers (thin HIP wrappers) ---- */
int bpc_dev_malloc(void** ptr, size_t bytes);
int bpc_dev_free(void* ptr);
int bpc_dev_memset(void* ptr, int value, size_t bytes, void* stream);
int bpc_copy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int bpc_copy_d2h(void* dst, const void* src, size_t bytes, void* stream);
/* Asynchronous device-to-device copy by a plain copy kernel; 16-byte aligned
 * pointers, bytes a multiple of 16. */
int bpc_copy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int bpc_stream_create(void** stream);
int bpc_stream_destroy(void* stream);
int bpc_stream_sync(void* stream);
int bpc_device_sync(void);
int bpc_event_create(void** ev);
int bpc_event_destroy(void* ev);
int bpc_event_record(void* ev, void* stream);
int bpc_event_elapsed_ms(void* start, void* stop, float* ms);
int bpc_mem_info(size_t* free_bytes, size_t* total_bytes);

#ifdef __cplusplus
}
#endif
#endif
