/*
 * bpc_oracle.c -- CPU oracle for the BPC-PaCo codeblock coder.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product path).  A plain-C
 * restatement of source/BPC_PaCo.cu from the reference: one CUDA warp codes one
 * codeblock and each of its 32 lanes owns a 2-column stripe.  The warp is
 * emulated lane by lane and phase by phase:
 *   - every scan step first snapshots the neighbour words of all lanes (the
 *     warp shuffles of BPC_PaCo.cu:342-388 read the values at step start);
 *   - every arithmetic-coder call site is then run for the lanes present at
 *     that site in ascending lane order, so that a lane needing a new codeword
 *     takes slot 1 + counter and bumps the counter.  This is exactly the slot the
 *     warp ballot / popcount prefix of BPC_PaCo.cu:268-277 and :304-316 gives:
 *     rank among the lanes at the site with an exhausted interval, plus the
 *     per-warp counter, which the highest such lane then advances.
 *
 * Decisions on behaviour the reference leaves undefined (applied identically by
 * the HIP kernels):
 *   - all-zero codeblock: Find_MSB yields 32 (BPC_PaCo.cu:159); the MSB word 32 is
 *     written and nothing is coded; the decoder maps MSB word 32 to zeros;
 *   - |coefficient| >= 2^28 would collide with the R/C/S flags (bits 29..31):
 *     rejected with ORC_ERR_RANGE;
 *   - LUT rows are addressed exactly as BPC_PaCo.cu:228-258 (so MSB >= NBP aliases
 *     into the next group as in the reference), but a row outside the array is
 *     rejected with ORC_ERR_RANGE;
 *   - a codeword slot at or past the region end (BPC_PaCo.cu:1130) is not written
 *     and reported as ORC_ERR_OVERFLOW (the reference would spill into the next
 *     codeblock's region).
 */
#include "bpc_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define S_BIT 0x80000000u /* significant            (BPC_PaCo.cu:439)  */
#define C_BIT 0x40000000u /* visit in cleanup pass  (BPC_PaCo.cu:37)   */
#define R_BIT 0x20000000u /* refine in later planes (BPC_PaCo.cu:604)  */

#define NLANES 32     /* WARPSIZE, common.h:40 */
#define MAXROWS 64    /* largest CBLOCK_LENGTH handled here */
#define N_REF 1       /* N_CONTEXT_REFINEMENT, common.h:31 */
#define N_SIG 9       /* N_CONTEXT_SIGNIFICANCE, common.h:33 */
#define N_SIGN 4      /* N_CONTEXT_SIGN, common.h:32 */
#define N_SUBBANDS 3  /* LUT_N_SUBBANDS, common.h:29 */

/* ------------------------------------------------------------------------- */
/* configuration helpers                                                      */

static int stride_of(const orc_params* p) {
  return p->cb_stride_words > 0 ? p->cb_stride_words : 64 * p->cblock_length;
}

static int ngroups(const orc_params* p) { return N_SUBBANDS * p->dwt_levels + 1; }

int orc_lut_len(const orc_params* p) {
  int per = N_REF + N_SIG + N_SIGN + (p->cpass3 ? N_SIG + N_SIGN : 0);
  return ngroups(p) * p->lut_n_bitplanes * per;
}

static int check_params(int size_x, int size_y, const orc_params* p, int lut_len) {
  if (!p) return ORC_ERR_ARG;
  if (p->cpass3 != 0 && p->cpass3 != 1) return ORC_ERR_ARG;
  if (p->dwt_levels < 0 || p->dwt_levels > 16) return ORC_ERR_ARG;
  if (p->cblock_length < 2 || p->cblock_length > MAXROWS) return ORC_ERR_ARG;
  if (p->lut_n_bitplanes < 1 || p->lut_n_bitplanes > 32) return ORC_ERR_ARG;
  if (stride_of(p) < 2) return ORC_ERR_ARG;
  if (size_x <= 0 || size_y <= 0 || size_x % 64 || size_y % p->cblock_length) return ORC_ERR_ARG;
  if (lut_len < orc_lut_len(p)) return ORC_ERR_ARG;
  return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* primitives                                                                 */

/* BPC_PaCo.cu:164-172 */
int orc_sig_context(const uint32_t nb[8]) {
  int c = 0;
  for (int k = 0; k < 8; ++k) c += (int)(nb[k] >> 31);
  return c;
}

/* BPC_PaCo.cu:178-211 */
int orc_sign_context_hv(int horizontal, int vertical) {
  if (horizontal == 0) return vertical == 0 ? 0 : (vertical > 0 ? 2 : 3);
  if (horizontal > 0) return vertical == 0 ? 4 : (vertical > 0 ? 6 : 0);
  return vertical == 0 ? 5 : (vertical > 0 ? 1 : 7);
}

/* chi(n): 0 if not significant, -1 if negative, +1 otherwise (BPC_PaCo.cu:220-223) */
static int chi(uint32_t w) { return (w >> 31) == 0 ? 0 : ((w & 1) ? -1 : 1); }

/* BPC_PaCo.cu:214-226 */
int orc_sign_context(uint32_t u, uint32_t l, uint32_t r, uint32_t d) {
  return orc_sign_context_hv(chi(l) + chi(r), chi(u) + chi(d));
}

/* BPC_PaCo.cu:84-111 */
void orc_find_subband(int x, int y, int size_x, int size_y, int dwt_levels, int* level, int* subband) {
  int aux = 1;
  *level = -1;
  *subband = -1;
  while (*level == -1) {
    if (aux == dwt_levels + 1) {
      *level = dwt_levels;
      *subband = 0;
    } else if (x >= (size_x >> aux) || y >= (size_y >> aux)) {
      *level = aux - 1;
      if (x >= (size_x >> aux))
        *subband = (y >= (size_y >> aux)) ? 2 : 0;
      else
        *subband = 1;
    } else {
      aux++;
    }
  }
}

static int msb_of(uint32_t m) { /* 32 - __ffs(__brev(m)), BPC_PaCo.cu:159 */
  if (m == 0) return 32;
  int b = 31;
  while (!(m >> b)) --b;
  return b;
}

int orc_codeblock_msb(const int32_t* img, int size_x, int size_y, int cb, int cbl) {
  int nx = size_x / 64;
  int tx = cb % nx, ty = cb / nx;
  (void)size_y;
  uint32_t m = 0;
  for (int r = 0; r < cbl; ++r)
    for (int c = 0; c < 64; ++c) {
      int64_t v = img[(int64_t)(ty * cbl + r) * size_x + tx * 64 + c];
      m |= (uint32_t)(v < 0 ? -v : v);
    }
  return msb_of(m);
}

/* ------------------------------------------------------------------------- */
/* warp emulation state                                                       */

typedef struct {
  uint32_t L, S; /* AC_Interval_Lower, AC_Interval_Size (BPC_PaCo.cu:973-975) */
  uint32_t slot; /* Reserved_codeword, relative to the region (encoder)         */
  uint32_t cw;   /* Codeword (decoder, BPC_PaCo.cu:1029)                        */
} orc_ac;

typedef struct {
  uint32_t w[NLANES][2 * MAXROWS]; /* TData per lane: [2*row + col] (BPC_PaCo.cu:1120) */
  orc_ac ac[NLANES];
  uint32_t counter; /* Code_Stream_Shared of this warp (BPC_PaCo.cu:1121,1147)    */
  int32_t* wregion; /* encoder output region (word 0 = MSB)                        */
  const int32_t* rregion;
  int stride;
  int rows;
  int err;
  const int32_t* lut;
  int ref[NLANES], sig[NLANES], sign[NLANES]; /* LUT row pointers at plane bp = 0 */
  int cpoff;                                  /* cp_sig/cp_sign offset (BPC_PaCo.cu:988-989) */
  uint32_t *n0, *n; /* symbol-count mode (non-NULL): per LUT entry, zeros and all symbols */
  int count_nbp;    /* symbol-count mode: planes >= this are modelled but not counted */
  /* per-symbol trace (non-NULL): one record per coded symbol, in coding order */
  orc_trace_rec* trace;
  int trace_cap, trace_n;
  int t_plane, t_pass, t_row, t_odd, t_site; /* where the scan is (set by the passes) */
} orc_cb;

/* One trace record: the printf of BPC_PaCo.cu:266,317,434,445 (THREAD_DEBUG_ID)
 * kept as data for every lane.  L and S are the interval before the symbol, after
 * any reservation the symbol made. */
static void trace_symbol(orc_cb* c, int lane, int idx, int prob, const orc_ac* a, uint32_t sym, int opened) {
  if (!c->trace) return;
  if (c->trace_n < c->trace_cap) {
    orc_trace_rec* r = &c->trace[c->trace_n];
    r->plane = (int16_t)c->t_plane;
    r->pass = (int8_t)c->t_pass;
    r->site = (int8_t)c->t_site;
    r->row = (int16_t)c->t_row;
    r->column = (int8_t)(2 * lane + c->t_odd);
    r->lane = (int8_t)lane;
    r->lut_index = idx;
    r->prob = prob;
    r->L = a->L;
    r->S = a->S;
    r->slot = a->slot;
    r->sym = sym;
    r->opened = (uint32_t)opened;
  }
  c->trace_n++;
}

/* ---- arithmetic coder, BPC_PaCo.cu:261-295 (encoder) ---- */
static void ac_encode(orc_cb* c, int lane, uint32_t sym, int idx) {
  orc_ac* a = &c->ac[lane];
  const int prob = c->lut[idx];
  int opened = 0;
  if (a->S == 0) { /* reservation: BPC_PaCo.cu:268-277 */
    a->L = 0;
    a->S = (1u << 16) - 1;
    a->slot = 1u + c->counter++;
    opened = 1;
  }
  trace_symbol(c, lane, idx, prob, a, sym, opened);
  uint32_t aux = ((a->S * (uint32_t)prob) >> 7) + sym; /* MULT_PRECISION = 7 */
  if (sym == 0) {
    a->S = aux;
  } else {
    a->S -= aux;
    a->L += aux;
  }
  if (a->S == 0) { /* BPC_PaCo.cu:290-294 */
    if (a->slot < (uint32_t)c->stride)
      c->wregion[a->slot] = (int32_t)a->L;
    else
      c->err = ORC_ERR_OVERFLOW;
  }
}

/* Encoder symbol sink: arithmetic-code sym with probability lut[idx], or, in
 * symbol-count mode, count it at LUT entry idx instead (no codeword is formed). */
static void emit(orc_cb* c, int lane, uint32_t sym, int idx, int bp) {
  if (c->n) {
    if (bp < c->count_nbp) {
      c->n[idx]++;
      if (!sym) c->n0[idx]++;
    }
    return;
  }
  ac_encode(c, lane, sym, idx);
}

/* ---- arithmetic coder, BPC_PaCo.cu:297-339 (decoder) ---- */
static uint32_t ac_decode(orc_cb* c, int lane, int idx) {
  orc_ac* a = &c->ac[lane];
  const int prob = c->lut[idx];
  int opened = 0;
  if (a->S == 0) { /* BPC_PaCo.cu:304-319 */
    a->L = 0;
    a->S = (1u << 16) - 1;
    uint32_t slot = 1u + c->counter++;
    a->slot = slot;
    opened = 1;
    if (slot < (uint32_t)c->stride) {
      a->cw = (uint32_t)c->rregion[slot];
    } else {
      a->cw = 0;
      c->err = ORC_ERR_OVERFLOW;
    }
  }
  const orc_ac before = *a;
  uint32_t aux = ((a->S * (uint32_t)prob) >> 7) + 1;
  uint32_t aux2 = a->L + aux;
  uint32_t sym = 0;
  if (a->cw >= aux2) {
    a->S -= aux;
    a->L = aux2;
    sym = 1;
  } else {
    a->S = aux - 1;
  }
  trace_symbol(c, lane, idx, prob, &before, sym, opened);
  return sym;
}

/* ---- neighbour snapshot of one scan step (BPC_PaCo.cu:633-706, 342-410) ----
 * nb[k] in the order UL, U, UR, L, R, DL, D, DR (TData1..TData8 of SPP_Encoder).
 * odd == 0: even column (direction 1, Share_Left_Borders, lane 0 zeroes UL/L/DL);
 * odd == 1: odd column (direction 2, Share_Right_Borders, lane 31 zeroes UR/R/DR).
 * Rows -1 and CBLOCK_LENGTH are zero (the first/last calls of each launcher). */
static void snapshot(const orc_cb* c, int row, int odd, uint32_t nb[NLANES][8]) {
  const int up = row > 0, dn = row < c->rows - 1;
  for (int l = 0; l < NLANES; ++l) {
    const uint32_t* o = c->w[l];
    uint32_t* n = nb[l];
    if (!odd) {
      const uint32_t* lf = l > 0 ? c->w[l - 1] : NULL;
      n[0] = (up && lf) ? lf[2 * (row - 1) + 1] : 0; /* UL: shfl_up(TData[ci-1]) */
      n[1] = up ? o[2 * (row - 1)] : 0;              /* U : TData[ci-2] */
      n[2] = up ? o[2 * (row - 1) + 1] : 0;          /* UR: TData[ci-1] */
      n[3] = lf ? lf[2 * row + 1] : 0;               /* L : shfl_up(TData[ci+1]) */
      n[4] = o[2 * row + 1];                         /* R : TData[ci+1] */
      n[5] = (dn && lf) ? lf[2 * (row + 1) + 1] : 0; /* DL: shfl_up(TData[ci+3]) */
      n[6] = dn ? o[2 * (row + 1)] : 0;              /* D : TData[ci+2] */
      n[7] = dn ? o[2 * (row + 1) + 1] : 0;          /* DR: TData[ci+3] */
    } else {
      const uint32_t* rt = l < NLANES - 1 ? c->w[l + 1] : NULL;
      n[0] = up ? o[2 * (row - 1)] : 0;             /* UL: TData[ci-3] */
      n[1] = up ? o[2 * (row - 1) + 1] : 0;         /* U : TData[ci-2] */
      n[2] = (up && rt) ? rt[2 * (row - 1)] : 0;    /* UR: shfl_down(TData[ci-3]) */
      n[3] = o[2 * row];                            /* L : TData[ci-1] */
      n[4] = rt ? rt[2 * row] : 0;                  /* R : shfl_down(TData[ci-1]) */
      n[5] = dn ? o[2 * (row + 1)] : 0;             /* DL: TData[ci+1] */
      n[6] = dn ? o[2 * (row + 1) + 1] : 0;         /* D : TData[ci+2] */
      n[7] = (dn && rt) ? rt[2 * (row + 1)] : 0;    /* DR: shfl_down(TData[ci+1]) */
    }
  }
}

static int any_sig(const uint32_t nb[8]) {
  uint32_t o = 0;
  for (int k = 0; k < 8; ++k) o |= nb[k];
  return (int)(o >> 31);
}

/* ---- significance / cleanup pass step ----
 * cleanup == 0: SPP_Encoder/SPP_Decoder (BPC_PaCo.cu:412-507)
 * cleanup == 1: CP_Encoder/CP_Decoder   (BPC_PaCo.cu:510-589)
 * Two call sites per step: significance for all participating lanes, then sign
 * for the lanes whose significance symbol was 1 (program order of :431/:443). */
static void sig_step(orc_cb* c, int row, int odd, int bp, int cleanup, int decode, uint32_t mask, int cp3) {
  uint32_t nb[NLANES][8];
  int pending[NLANES];
  snapshot(c, row, odd, nb);
  const int ci = 2 * row + odd;
  for (int l = 0; l < NLANES; ++l) {
    pending[l] = 0;
    uint32_t* w = &c->w[l][ci];
    const int sbase = c->sig[l] + bp * N_SIG + (cleanup ? c->cpoff : 0);
    if (!cleanup) {
      if (*w & S_BIT) continue;
      if (cp3 && !any_sig(nb[l])) { /* BPC_PaCo.cu:424,452 */
        *w |= C_BIT;
        continue;
      }
    } else {
      if (!(*w & C_BIT)) continue; /* BPC_PaCo.cu:515 */
    }
    const int ctx = orc_sig_context(nb[l]);
    uint32_t sym;
    c->t_site = 0;
    if (!decode) {
      sym = (*w >> (bp + 1)) & 1u;
      emit(c, l, sym, sbase + ctx, bp);
    } else {
      sym = ac_decode(c, l, sbase + ctx);
    }
    if (cleanup) *w &= ~C_BIT; /* BPC_PaCo.cu:528,563 */
    if (sym) {
      if (decode) *w |= mask; /* BPC_PaCo.cu:485,568 */
      *w |= S_BIT;
      if (cleanup) *w |= R_BIT; /* BPC_PaCo.cu:534,571 */
      pending[l] = 1;
    }
  }
  for (int l = 0; l < NLANES; ++l) {
    if (!pending[l]) continue;
    uint32_t* w = &c->w[l][ci];
    const uint32_t* n = nb[l];
    const int sctx = orc_sign_context(n[1], n[3], n[4], n[6]); /* (U, L, R, D) */
    const int pidx = c->sign[l] + bp * N_SIGN + (cleanup ? c->cpoff : 0) + (sctx >> 1);
    c->t_site = 1;
    if (!decode) {
      emit(c, l, ((*w & 1u) == (uint32_t)(sctx & 1)) ? 0u : 1u, pidx, bp);
    } else {
      uint32_t s = ac_decode(c, l, pidx);
      s = ((s & 1u) == (uint32_t)(sctx & 1)) ? 0u : 1u;
      *w |= s;
    }
  }
}

/* ---- refinement pass step: MRP_Encoder/MRP_Decoder (BPC_PaCo.cu:593-629) ---- */
static void ref_step(orc_cb* c, int row, int odd, int bp, int decode, uint32_t mask) {
  const int ci = 2 * row + odd;
  for (int l = 0; l < NLANES; ++l) {
    uint32_t* w = &c->w[l][ci];
    const int ridx = c->ref[l] + bp * N_REF;
    c->t_site = 2;
    if (*w & R_BIT) {
      if (!decode) {
        emit(c, l, (*w >> (bp + 1)) & 1u, ridx, bp);
      } else {
        uint32_t s = ac_decode(c, l, ridx);
        *w &= ~mask;
        *w |= mask & (((s << 1) + 1u) << bp);
      }
    } else if (*w & S_BIT) {
      *w |= R_BIT;
    }
  }
}

/* launchers: rows top to bottom, even column then odd column (BPC_PaCo.cu:633-962) */
static void sig_pass(orc_cb* c, int bp, int cleanup, int decode, uint32_t mask, int cp3) {
  c->t_plane = bp;
  c->t_pass = cleanup ? 2 : 0;
  for (int row = 0; row < c->rows; ++row) {
    c->t_row = row;
    c->t_odd = 0;
    sig_step(c, row, 0, bp, cleanup, decode, mask, cp3);
    c->t_odd = 1;
    sig_step(c, row, 1, bp, cleanup, decode, mask, cp3);
  }
}

static void ref_pass(orc_cb* c, int bp, int decode, uint32_t mask) {
  c->t_plane = bp;
  c->t_pass = 1;
  for (int row = 0; row < c->rows; ++row) {
    c->t_row = row;
    c->t_odd = 0;
    ref_step(c, row, 0, bp, decode, mask);
    c->t_odd = 1;
    ref_step(c, row, 1, bp, decode, mask);
  }
}

/* ---- per-lane geometry and LUT pointers ---- */
static int setup_lanes(orc_cb* c, int size_x, int size_y, int cb, const orc_params* p, int lut_len, int msb) {
  const int nx = size_x / 64;
  const int tx = cb % nx, ty = cb / nx;
  const int nbp = p->lut_n_bitplanes;
  const int G = ngroups(p);
  const int off_sig = G * nbp * N_REF;            /* BPC_PaCo.cu:237 */
  const int off_sign = off_sig + G * nbp * N_SIG; /* BPC_PaCo.cu:244 */
  c->cpoff = G * nbp * (N_SIG + N_SIGN);          /* BPC_PaCo.cu:988-989 */
  for (int l = 0; l < NLANES; ++l) {
    int level, sb;
    /* TCoordinate_X / TCoordinate_Y of BPC_PaCo.cu:1126-1127 */
    orc_find_subband(tx * 64 + 2 * l, ty * p->cblock_length, size_x, size_y, p->dwt_levels, &level, &sb);
    const int g = level * N_SUBBANDS + sb;
    /* Initialize_LUT_pointers at MSB followed by one decrement per plane
     * (BPC_PaCo.cu:228-258) puts plane bp at row bp; stored here at bp = 0. */
    c->ref[l] = g * nbp * N_REF;
    c->sig[l] = g * nbp * N_SIG + off_sig;
    c->sign[l] = g * nbp * N_SIGN + off_sign;
    /* every row touched (planes 0..msb) must lie inside the LUT array */
    if (c->ref[l] + msb * N_REF + N_REF > lut_len) return ORC_ERR_RANGE;
    if (c->sig[l] + msb * N_SIG + N_SIG + (p->cpass3 ? c->cpoff : 0) > lut_len) return ORC_ERR_RANGE;
    if (c->sign[l] + msb * N_SIGN + N_SIGN + (p->cpass3 ? c->cpoff : 0) > lut_len) return ORC_ERR_RANGE;
  }
  return ORC_OK;
}

/* Bitplane driver, Encode/Decode of BPC_PaCo.cu:968-1090.  In 3-pass mode the
 * initial cleanup pass at MSB is followed by SPP/MRP/CP for MSB-1 .. 0. */
static void run_planes(orc_cb* c, int msb, int cp3, int decode) {
  int bp = msb;
  uint32_t mask = 0x3u << bp; /* BPC_PaCo.cu:1038-1040 */
  if (bp == 0) mask &= 0x2u;
  if (cp3) {
    sig_pass(c, bp, 1, decode, mask, cp3); /* BPC_PaCo.cu:991,1050 */
    bp--;
    mask >>= 1;
    if (bp == 0) mask = 0x2u;
  }
  for (; bp >= 0; bp--) {
    sig_pass(c, bp, 0, decode, mask, cp3);
    ref_pass(c, bp, decode, mask);
    if (cp3) sig_pass(c, bp, 1, decode, mask, cp3);
    mask >>= 1;
    if (bp == 1) mask = 0x2u; /* BPC_PaCo.cu:1081-1083 */
  }
}

/* Read_Coefficients (BPC_PaCo.cu:14-44) of codeblock cb into the lanes' words and
 * its MSB (Find_MSB, :114-161: OR of (TData >> 1) with the C image cleared).
 * Returns nonzero when a magnitude is >= 2^28 (it would reach the flag bits). */
static int read_coefficients(orc_cb* c, const int32_t* img, int size_x, int cb, const orc_params* p, int* msb) {
  const int nx = size_x / 64;
  const int tx = cb % nx, ty = cb / nx;
  c->rows = p->cblock_length;
  uint32_t m = 0;
  int range_err = 0;
  for (int l = 0; l < NLANES; ++l)
    for (int r = 0; r < c->rows; ++r)
      for (int k = 0; k < 2; ++k) {
        int32_t v = img[(int64_t)(ty * c->rows + r) * size_x + tx * 64 + 2 * l + k];
        uint32_t mag = v < 0 ? (uint32_t)0 - (uint32_t)v : (uint32_t)v;
        if (mag >= (1u << 28)) range_err = 1;
        uint32_t w = (mag << 1) | (v < 0 ? 1u : 0u);
        if (p->cpass3) w |= C_BIT;
        c->w[l][2 * r + k] = w;
        m |= mag;
      }
  *msb = msb_of(m);
  return range_err;
}

static int encode_cb(const int32_t* img, int size_x, int size_y, int cb, const int32_t* lut, int lut_len,
                     const orc_params* p, int32_t* region, int32_t* ncw, orc_trace_rec* trace, int trace_cap,
                     int* trace_n) {
  int st = check_params(size_x, size_y, p, lut_len);
  if (st) return st;
  if (cb < 0 || cb >= (size_x / 64) * (size_y / p->cblock_length)) return ORC_ERR_ARG;
  orc_cb* c = (orc_cb*)calloc(1, sizeof(orc_cb));
  if (!c) return ORC_ERR_ARG;
  c->trace = trace;
  c->trace_cap = trace_cap;
  if (trace_n) *trace_n = 0;
  c->stride = stride_of(p);
  c->lut = lut;
  c->wregion = region;
  int msb;
  if (read_coefficients(c, img, size_x, cb, p, &msb)) {
    free(c);
    return ORC_ERR_RANGE;
  }
  region[0] = msb; /* Output[WCoordinate] = MSB, BPC_PaCo.cu:1155 */
  if (ncw) *ncw = 0;
  if (msb == 32) { /* all-zero codeblock: nothing coded */
    free(c);
    return ORC_OK;
  }
  st = setup_lanes(c, size_x, size_y, cb, p, lut_len, msb);
  if (st) {
    free(c);
    return st;
  }
  run_planes(c, msb, p->cpass3, 0);
  /* flush of the open codeword of every lane, BPC_PaCo.cu:1020 */
  for (int l = 0; l < NLANES; ++l) {
    if (c->ac[l].slot < (uint32_t)c->stride)
      region[c->ac[l].slot] = (int32_t)c->ac[l].L;
    else
      c->err = ORC_ERR_OVERFLOW;
  }
  if (ncw) *ncw = (int32_t)c->counter;
  if (trace_n) *trace_n = c->trace_n;
  st = c->err;
  free(c);
  return st;
}

int orc_encode_cb(const int32_t* img, int size_x, int size_y, int cb, const int32_t* lut, int lut_len,
                  const orc_params* p, int32_t* region, int32_t* ncw) {
  return encode_cb(img, size_x, size_y, cb, lut, lut_len, p, region, ncw, NULL, 0, NULL);
}

int orc_encode_cb_trace(const int32_t* img, int size_x, int size_y, int cb, const int32_t* lut, int lut_len,
                        const orc_params* p, int32_t* region, int32_t* ncw, orc_trace_rec* trace, int cap,
                        int* n) {
  if (!trace || cap < 0 || !n) return ORC_ERR_ARG;
  return encode_cb(img, size_x, size_y, cb, lut, lut_len, p, region, ncw, trace, cap, n);
}

static int decode_cb(const int32_t* region, int size_x, int size_y, int cb, const int32_t* lut, int lut_len,
                     const orc_params* p, int32_t* img, orc_trace_rec* trace, int trace_cap, int* trace_n) {
  if (trace_n) *trace_n = 0;
  int st = check_params(size_x, size_y, p, lut_len);
  if (st) return st;
  if (cb < 0 || cb >= (size_x / 64) * (size_y / p->cblock_length)) return ORC_ERR_ARG;
  const int nx = size_x / 64;
  const int tx = cb % nx, ty = cb / nx;
  const int rows = p->cblock_length;
  const int32_t msb = region[0]; /* BPC_PaCo.cu:1216 */
  int err = ORC_OK;
  orc_cb* c = NULL;
  if (msb != 32 && (msb < 0 || msb > 27)) err = ORC_ERR_MALFORMED;
  if (!err && msb != 32) {
    c = (orc_cb*)calloc(1, sizeof(orc_cb));
    if (!c) return ORC_ERR_ARG;
    c->rows = rows;
    c->stride = stride_of(p);
    c->lut = lut;
    c->rregion = region;
    c->trace = trace;
    c->trace_cap = trace_cap;
    /* Initialize_Coefficients, BPC_PaCo.cu:65-81 */
    for (int l = 0; l < NLANES; ++l)
      for (int k = 0; k < 2 * rows; ++k) c->w[l][k] = p->cpass3 ? C_BIT : 0u;
    err = setup_lanes(c, size_x, size_y, cb, p, lut_len, msb);
    if (!err) {
      run_planes(c, msb, p->cpass3, 1);
      err = c->err;
    }
    if (trace_n) *trace_n = c->trace_n;
  }
  /* Write_Coefficients, BPC_PaCo.cu:46-63 (zeros for the all-zero sentinel or on error) */
  for (int l = 0; l < NLANES; ++l)
    for (int r = 0; r < rows; ++r)
      for (int k = 0; k < 2; ++k) {
        int32_t v = 0;
        if (c && !err) {
          uint32_t w = c->w[l][2 * r + k];
          v = (int32_t)((w & 0x1FFFFFFFu) >> 1);
          if (w & 1u) v = -v;
        }
        img[(int64_t)(ty * rows + r) * size_x + tx * 64 + 2 * l + k] = v;
      }
  free(c);
  return err;
}

int orc_decode_cb(const int32_t* region, int size_x, int size_y, int cb, const int32_t* lut, int lut_len,
                  const orc_params* p, int32_t* img) {
  return decode_cb(region, size_x, size_y, cb, lut, lut_len, p, img, NULL, 0, NULL);
}

int orc_decode_cb_trace(const int32_t* region, int size_x, int size_y, int cb, const int32_t* lut, int lut_len,
                        const orc_params* p, int32_t* img, orc_trace_rec* trace, int cap, int* n) {
  if (!trace || cap < 0 || !n) return ORC_ERR_ARG;
  return decode_cb(region, size_x, size_y, cb, lut, lut_len, p, img, trace, cap, n);
}

/* ------------------------------------------------------------------------- */
/* whole-image drivers (OpenMP over codeblocks, schedule(dynamic))            */

static int merge_err(int a, int b) { return a ? a : b; }

int orc_encode_list(const int32_t* img, int size_x, int size_y, const int32_t* cbs, int ncbs,
                    const int32_t* lut, int lut_len, const orc_params* p, int32_t* stream,
                    int32_t* ncw, int nthreads) {
  int st = check_params(size_x, size_y, p, lut_len);
  if (st) return st;
  const int64_t stride = stride_of(p);
  int err = ORC_OK;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 4)
  for (int k = 0; k < ncbs; ++k) {
    const int cb = cbs ? cbs[k] : k;
    int32_t* region = stream + (int64_t)cb * stride;
    for (int64_t j = 0; j < stride; ++j) region[j] = -1; /* BPC_PaCo.cu:1284 */
    int e = orc_encode_cb(img, size_x, size_y, cb, lut, lut_len, p, region, ncw ? &ncw[cb] : NULL);
    if (e) {
#pragma omp critical
      err = merge_err(err, e);
    }
  }
  return err;
}

int orc_decode_list(const int32_t* stream, int size_x, int size_y, const int32_t* cbs, int ncbs,
                    const int32_t* lut, int lut_len, const orc_params* p, int32_t* img, int nthreads) {
  int st = check_params(size_x, size_y, p, lut_len);
  if (st) return st;
  const int64_t stride = stride_of(p);
  int err = ORC_OK;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 4)
  for (int k = 0; k < ncbs; ++k) {
    const int cb = cbs ? cbs[k] : k;
    int e = orc_decode_cb(stream + (int64_t)cb * stride, size_x, size_y, cb, lut, lut_len, p, img);
    if (e) {
#pragma omp critical
      err = merge_err(err, e);
    }
  }
  return err;
}

int orc_encode(const int32_t* img, int size_x, int size_y, const int32_t* lut, int lut_len,
               const orc_params* p, int32_t* stream, int32_t* ncw, int nthreads) {
  if (!p || p->cblock_length <= 0) return ORC_ERR_ARG;
  const int n = (size_x / 64) * (size_y / p->cblock_length);
  return orc_encode_list(img, size_x, size_y, NULL, n, lut, lut_len, p, stream, ncw, nthreads);
}

int orc_decode(const int32_t* stream, int size_x, int size_y, const int32_t* lut, int lut_len,
               const orc_params* p, int32_t* img, int nthreads) {
  if (!p || p->cblock_length <= 0) return ORC_ERR_ARG;
  const int n = (size_x / 64) * (size_y / p->cblock_length);
  return orc_decode_list(stream, size_x, size_y, NULL, n, lut, lut_len, p, img, nthreads);
}

/* ------------------------------------------------------------------------- */
/* single-lane coder (unit tests)                                             */

int orc_ac_encode_seq(const uint8_t* syms, const int32_t* probs, int n, uint32_t* cw, int cw_cap) {
  uint32_t L = 0, S = 0, slot = 0, used = 0;
  for (int k = 0; k < n; ++k) {
    if (S == 0) {
      L = 0;
      S = 0xFFFFu;
      slot = used++;
    }
    uint32_t sym = syms[k] & 1u;
    uint32_t aux = ((S * (uint32_t)probs[k]) >> 7) + sym;
    if (!sym) {
      S = aux;
    } else {
      S -= aux;
      L += aux;
    }
    if (S == 0 && slot < (uint32_t)cw_cap) cw[slot] = L;
  }
  if (used > 0 && slot < (uint32_t)cw_cap) cw[slot] = L; /* flush, BPC_PaCo.cu:1020 */
  return (int)used;
}

int orc_ac_decode_seq(const uint32_t* cw, int n_cw, const int32_t* probs, int n, uint8_t* syms) {
  uint32_t L = 0, S = 0, C = 0, used = 0;
  for (int k = 0; k < n; ++k) {
    if (S == 0) {
      L = 0;
      S = 0xFFFFu;
      C = used < (uint32_t)n_cw ? cw[used] : 0;
      used++;
    }
    uint32_t aux = ((S * (uint32_t)probs[k]) >> 7) + 1;
    uint32_t aux2 = L + aux;
    if (C >= aux2) {
      S -= aux;
      L = aux2;
      syms[k] = 1;
    } else {
      S = aux - 1;
      syms[k] = 0;
    }
  }
  return (int)used;
}

/* ------------------------------------------------------------------------- */
/* LUT text loader, restating main_example.cu:219-448                          */

static int scan_max_bp(const char* path, int nfields, int* maxbp) {
  FILE* f = fopen(path, "rb");
  if (!f) return ORC_ERR_ARG;
  int lvl, sb, bp;
  char colon[4];
  while (fscanf(f, "%d %d %d %3s", &lvl, &sb, &bp, colon) == 4) {
    int x;
    for (int k = 0; k < nfields; ++k)
      if (fscanf(f, "%d", &x) != 1) {
        fclose(f);
        return ORC_ERR_ARG;
      }
    if (bp > *maxbp) *maxbp = bp;
  }
  fclose(f);
  return ORC_OK;
}

/* One region: main_example.cu:285-302 (ref), :305-335 (sig), :337-364 (sign),
 * :367-397 (cp_sig), :400-426 (cp_sign).  A new group starts whenever a row's
 * bitplane is <= the previous one; then (maxbp - prev)*N entries starting at
 * i + prev*N + 1 are set to 64 and i advances by maxbp*N. */
static int load_region(const char* path, int N, int maxbp, int32_t* lut, int base, int cap) {
  FILE* f = fopen(path, "rb");
  if (!f) return ORC_ERR_ARG;
  int i = base, prev = -1;
  int lvl, sb, bp;
  char colon[4];
  while (fscanf(f, "%d %d %d %3s", &lvl, &sb, &bp, colon) == 4) {
    int v[9];
    for (int k = 0; k < N; ++k)
      if (fscanf(f, "%d", &v[k]) != 1) {
        fclose(f);
        return ORC_ERR_ARG;
      }
    if (bp <= prev) {
      for (int z = 0; z < (maxbp - prev) * N; z++) {
        int idx = i + prev * N + z + 1;
        if (idx >= 0 && idx < cap) lut[idx] = 64;
      }
      i += maxbp * N;
    }
    prev = bp;
    for (int k = 0; k < N; ++k) {
      int idx = i + bp * N + k;
      if (idx < 0 || idx >= cap) {
        fclose(f);
        return ORC_ERR_ARG;
      }
      lut[idx] = v[k];
    }
  }
  fclose(f);
  return ORC_OK;
}

int orc_lut_load_text(const char* folder, int cpass3, int dwt_levels, int32_t* out, int cap, int* nbp,
                      int* len) {
  char path[4096];
  int maxbp = -1;
  const char* files[5] = {"ref.txt", "sig.txt", "sign.txt", "cp_sig.txt", "cp_sign.txt"};
  const int nf[5] = {N_REF, N_SIG, N_SIGN, N_SIG, N_SIGN};
  /* main_example.cu:231-264: the max bitplane is taken over ref, sig and sign only */
  for (int k = 0; k < 3; ++k) {
    snprintf(path, sizeof path, "%s%s", folder, files[k]);
    if (scan_max_bp(path, nf[k], &maxbp)) return ORC_ERR_ARG;
  }
  maxbp++;
  const int G = N_SUBBANDS * dwt_levels + 1;
  const int refN = G * maxbp * N_REF, sigN = G * maxbp * N_SIG, signN = G * maxbp * N_SIGN;
  const int total = refN + sigN + signN + (cpass3 ? sigN + signN : 0);
  if (total > cap) return ORC_ERR_ARG;
  for (int k = 0; k < cap; ++k) out[k] = 64;
  const int bases[5] = {0, refN, refN + sigN, refN + sigN + signN, refN + 2 * sigN + signN};
  for (int k = 0; k < (cpass3 ? 5 : 3); ++k) {
    snprintf(path, sizeof path, "%s%s", folder, files[k]);
    if (load_region(path, nf[k], maxbp, out, bases[k], cap)) return ORC_ERR_ARG;
  }
  if (nbp) *nbp = maxbp;
  if (len) *len = total;
  return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* symbol statistics (the input of LUT training; the reference ships trained   */
/* tables only, main_example.cu:219-448, so the counting rule is this build's) */

int orc_symbol_counts(const int32_t* img, int size_x, int size_y, const orc_params* p, uint32_t* n0, uint32_t* n,
                      int lut_len, int nthreads) {
  int st = check_params(size_x, size_y, p, lut_len);
  if (st) return st;
  const int ncb = (size_x / 64) * (size_y / p->cblock_length);
  memset(n0, 0, sizeof(uint32_t) * (size_t)lut_len);
  memset(n, 0, sizeof(uint32_t) * (size_t)lut_len);
  int err = ORC_OK;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel
  {
    uint32_t* l0 = (uint32_t*)calloc((size_t)lut_len, sizeof(uint32_t));
    uint32_t* ln = (uint32_t*)calloc((size_t)lut_len, sizeof(uint32_t));
    orc_cb* c = (orc_cb*)malloc(sizeof(orc_cb));
    int lerr = (!l0 || !ln || !c) ? ORC_ERR_ARG : ORC_OK;
#pragma omp for schedule(dynamic, 4)
    for (int cb = 0; cb < ncb; ++cb) {
      if (lerr) continue;
      memset(c, 0, sizeof *c);
      int msb;
      /* codeblocks the encoder rejects (a magnitude >= 2^28) or does not code (all zero) count nothing */
      if (read_coefficients(c, img, size_x, cb, p, &msb) || msb == 32) continue;
      /* rows of planes >= lut_n_bitplanes are not counted, so no LUT bound applies */
      (void)setup_lanes(c, size_x, size_y, cb, p, 0x7FFFFFFF, msb);
      c->n0 = l0;
      c->n = ln;
      c->count_nbp = p->lut_n_bitplanes;
      run_planes(c, msb, p->cpass3, 0);
    }
#pragma omp critical
    {
      err = merge_err(err, lerr);
      if (!lerr)
        for (int k = 0; k < lut_len; ++k) {
          n0[k] += l0[k];
          n[k] += ln[k];
        }
    }
    free(l0);
    free(ln);
    free(c);
  }
  return err;
}

/* ------------------------------------------------------------------------- */
/* per-codeblock codestream files (main_example.cu:51-179)                    */

/* load_Data's enumeration (main_example.cu:71-179): file k of the LL band first,
 * then for each level from the coarsest the HL (x = sx, y = 0), LH (x = 0, y = sy)
 * and HH (x = sx, y = sy) bands, each scanned row by row in codeblock steps.
 * Returns the number of files, or ORC_ERR_ARG when a band is not a whole number
 * of codeblocks (the enumeration would then visit a tile twice). */
int orc_cb_file_order(int size_x, int size_y, int dwt_levels, int cblock_length, int32_t* tile_of_file, int cap) {
  if (size_x <= 0 || size_y <= 0 || dwt_levels < 0 || cblock_length <= 0) return ORC_ERR_ARG;
  const int nx = size_x / 64;
  int k = 0;
  /* LL band */
  {
    const int sx = size_x >> dwt_levels, sy = size_y >> dwt_levels;
    if (sx % 64 || sy % cblock_length) return ORC_ERR_ARG;
    for (int j = 0; j < sy; j += cblock_length)
      for (int i = 0; i < sx; i += 64) {
        if (k >= cap) return ORC_ERR_ARG;
        tile_of_file[k++] = (j / cblock_length) * nx + i / 64;
      }
  }
  for (int d = dwt_levels; d > 0; --d) {
    const int sx = size_x >> d, sy = size_y >> d;
    if (sx % 64 || sy % cblock_length) return ORC_ERR_ARG;
    const int x0[3] = {sx, 0, sx}, y0[3] = {0, sy, sy};
    for (int b = 0; b < 3; ++b)
      for (int j = 0; j < sy; j += cblock_length)
        for (int i = 0; i < sx; i += 64) {
          if (k >= cap) return ORC_ERR_ARG;
          tile_of_file[k++] = ((y0[b] + j) / cblock_length) * nx + (x0[b] + i) / 64;
        }
  }
  return k;
}

/* load_CStream (main_example.cu:51-68): `words` big-endian 16-bit values; every
 * value past the end of the file reads as -1 (fgetc's EOF in both bytes). */
void orc_load_cstream(const uint8_t* bytes, size_t nbytes, int32_t* region, int words) {
  for (int i = 0; i < words; ++i) {
    const size_t b = 2 * (size_t)i;
    const int hi = b < nbytes ? bytes[b] : -1;
    const int lo = b + 1 < nbytes ? bytes[b + 1] : -1;
    region[i] = (int32_t)(((uint32_t)hi << 8) | (uint32_t)lo);
  }
}
