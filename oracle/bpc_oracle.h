/*
 * bpc_oracle.h -- CPU oracle for the BPC-PaCo codeblock coder.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CUDA implementation (source/BPC_PaCo.cu) with the 32-lane warp emulated lane
 * by lane.  It is the checker for the HIP product path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped.
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef BPC_ORACLE_H
#define BPC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (shared numbering with include/bpc_paco.h). */
#define ORC_OK 0
#define ORC_ERR_ARG (-1)
#define ORC_ERR_OVERFLOW (-3)  /* codeword slots past the codeblock region */
#define ORC_ERR_RANGE (-4)     /* |coefficient| >= 2^28, or LUT row out of the array */
#define ORC_ERR_MALFORMED (-5) /* decoder: MSB word not in [0,27] and not 32 */

/* Runtime form of the reference's compile-time configuration (common.h:19-51). */
typedef struct orc_params {
  int32_t cpass3;          /* CPASS_3: 0 -> SPP,MRP ; 1 -> SPP,MRP,CP      common.h:24-26 */
  int32_t dwt_levels;      /* DWT_LEVELS                                   common.h:19-21 */
  int32_t cblock_length;   /* CBLOCK_LENGTH (rows per codeblock), 2..64    common.h:45-47 */
  int32_t lut_n_bitplanes; /* LUT_N_BITPLANES (15 in the reference)        common.h:28    */
  int32_t cb_stride_words; /* int32 words per codeblock region; 0 -> 64*cblock_length (BPC_PaCo.cu:1130) */
} orc_params;

/* LUT length in ints for a configuration: (3L+1)*NBP*(1+9+4 [+9+4]). */
int orc_lut_len(const orc_params* p);

/* Encode a raster int32 image (two's complement) into per-codeblock regions.
 * stream must hold nCB*stride words; it is filled with -1 first (BPC_PaCo.cu:1284).
 * ncw (nullable) receives the number of codeword slots used per codeblock.
 * nthreads <= 0 lets OpenMP choose.  Returns ORC_OK or a negative status. */
int orc_encode(const int32_t* img, int size_x, int size_y, const int32_t* lut, int lut_len,
               const orc_params* p, int32_t* stream, int32_t* ncw, int nthreads);

/* Decode per-codeblock regions back to a raster int32 image. */
int orc_decode(const int32_t* stream, int size_x, int size_y, const int32_t* lut, int lut_len,
               const orc_params* p, int32_t* img, int nthreads);

/* Single-codeblock forms (codeblock index cb in row-major tile order).
 * orc_encode_cb writes only region words it uses (caller pre-fills with -1). */
int orc_encode_cb(const int32_t* img, int size_x, int size_y, int cb, const int32_t* lut, int lut_len,
                  const orc_params* p, int32_t* region, int32_t* ncw);
int orc_decode_cb(const int32_t* region, int size_x, int size_y, int cb, const int32_t* lut, int lut_len,
                  const orc_params* p, int32_t* img);

/* Per-symbol trace of one codeblock: the reference's debugging printfs
 * (THREAD_DEBUG_ID, BPC_PaCo.cu:266,317,434,445) kept as one record per coded
 * symbol of every lane, in coding order (plane, pass, row, column parity, call
 * site, lane).  L and S are the interval before the symbol, after any slot
 * reservation that symbol made (opened = 1).  sym is the arithmetic-coded symbol
 * (for a sign, before the predicted-sign flip).  The region / image is written as
 * by orc_encode_cb / orc_decode_cb; *n receives the number of symbols (records
 * beyond cap are counted but not stored). */
typedef struct orc_trace_rec {
  int16_t plane;
  int8_t pass;   /* 0 significance propagation, 1 refinement, 2 cleanup */
  int8_t site;   /* 0 significance, 1 sign, 2 refinement */
  int16_t row;
  int8_t column; /* codeblock column 0..63 (2 * lane + parity) */
  int8_t lane;
  int32_t lut_index;
  int32_t prob;
  uint32_t L, S;
  uint32_t slot; /* codeword slot the symbol is coded into */
  uint32_t sym;
  uint32_t opened;
} orc_trace_rec;
int orc_encode_cb_trace(const int32_t* img, int size_x, int size_y, int cb, const int32_t* lut, int lut_len,
                        const orc_params* p, int32_t* region, int32_t* ncw, orc_trace_rec* trace, int cap, int* n);
int orc_decode_cb_trace(const int32_t* region, int size_x, int size_y, int cb, const int32_t* lut, int lut_len,
                        const orc_params* p, int32_t* img, orc_trace_rec* trace, int cap, int* n);

/* Same as orc_encode/orc_decode restricted to a list of codeblocks (OpenMP over the list). */
int orc_encode_list(const int32_t* img, int size_x, int size_y, const int32_t* cbs, int ncbs,
                    const int32_t* lut, int lut_len, const orc_params* p, int32_t* stream,
                    int32_t* ncw, int nthreads);
int orc_decode_list(const int32_t* stream, int size_x, int size_y, const int32_t* cbs, int ncbs,
                    const int32_t* lut, int lut_len, const orc_params* p, int32_t* img, int nthreads);

/* Symbol statistics of an image: the encoder's modelling (passes, scan order,
 * contexts) with every coded symbol counted at the LUT entry whose probability
 * would code it, in n[] (all symbols) and n0[] (symbols 0); no codewords are
 * formed.  Planes >= lut_n_bitplanes are modelled but not counted; codeblocks
 * with a magnitude >= 2^28 count nothing.  Arrays hold lut_len entries. */
int orc_symbol_counts(const int32_t* img, int size_x, int size_y, const orc_params* p, uint32_t* n0, uint32_t* n,
                      int lut_len, int nthreads);

/* Codeblock file order of load_Data (main_example.cu:71-179): tile_of_file[k] =
 * row-major tile index of file k.bin / k.txt.  Returns the file count, or
 * ORC_ERR_ARG when a band is not a whole number of 64 x cblock_length codeblocks. */
int orc_cb_file_order(int size_x, int size_y, int dwt_levels, int cblock_length, int32_t* tile_of_file, int cap);
/* load_CStream (main_example.cu:51-68): words big-endian 16-bit values from a file's
 * bytes into a region; values past the end of the file read as -1. */
void orc_load_cstream(const uint8_t* bytes, size_t nbytes, int32_t* region, int words);

/* ---- primitives exposed for unit tests ---------------------------------- */
/* Significance context: number of significant 8-neighbours (BPC_PaCo.cu:164-172). */
int orc_sig_context(const uint32_t nb[8]);
/* Sign context from (h, v) sums (BPC_PaCo.cu:178-211). */
int orc_sign_context_hv(int horizontal, int vertical);
/* Sign context from the U, L, R, D words (BPC_PaCo.cu:214-226). */
int orc_sign_context(uint32_t u, uint32_t l, uint32_t r, uint32_t d);
/* Decomposition level and subband of a lane coordinate (BPC_PaCo.cu:84-111). */
void orc_find_subband(int x, int y, int size_x, int size_y, int dwt_levels, int* level, int* subband);
/* MSB of a codeblock (BPC_PaCo.cu:114-161): floor(log2(OR |v|)), 32 when all zero. */
int orc_codeblock_msb(const int32_t* img, int size_x, int size_y, int cb, int cblock_length);

/* One lane's arithmetic coder in isolation (BPC_PaCo.cu:261-339), with a private
 * slot counter.  Encoder: codes n symbols with probabilities probs[] and returns the
 * number of codewords written to cw[] (including the final flush of BPC_PaCo.cu:1020).
 * Decoder: reads codewords from cw[] and returns the decoded symbols. */
int orc_ac_encode_seq(const uint8_t* syms, const int32_t* probs, int n, uint32_t* cw, int cw_cap);
int orc_ac_decode_seq(const uint32_t* cw, int n_cw, const int32_t* probs, int n, uint8_t* syms);

/* Restatement of main_example.cu:219-448 (load_LUTs): parses ref/sig/sign
 * (+cp_sig/cp_sign) text files from folder (which must end in '/').
 * Writes up to cap ints into out; *nbp receives max_bitplane+1 and *len the LUT length.
 * Entries the files never reach are set to 64. */
int orc_lut_load_text(const char* folder, int cpass3, int dwt_levels, int32_t* out, int cap,
                      int* nbp, int* len);

#ifdef __cplusplus
}
#endif
#endif
