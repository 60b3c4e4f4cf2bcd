// BPC-PaCo codeblock encoder / decoder for gfx950 (CDNA4, wave64).
//
// Work mapping.  The codestream format depends on 32 arithmetic coders per
// 64-column codeblock (one per 2-column stripe, BPC_PaCo.cu:1124-1130), so one
// wave64 carries two codeblocks: lanes 0-31 code codeblock 2w, lanes 32-63 code
// codeblock 2w+1 (the per-image kernel gives small images one codeblock per
// wave).  Lane l of a half owns columns 2(l%32) and 2(l%32)+1.
//
// State layout.  Instead of the reference's 128 packed words per lane
// (TData[], BPC_PaCo.cu:1120) each lane keeps its stripe bit-sliced: one 64-bit
// row mask per column for each flag (significant S, refine R, cleanup C, sign),
// so a context is a popcount of three 3-row windows and a flag update is one OR.
// The encoder's magnitude bitplanes (and the decoder's decoded bitplanes) live in
// a per-wave scratch slot, one plane (two row masks per lane) loaded or stored
// per bitplane.
//
// Cross-lane traffic per scan step: one DPP wave shift of the neighbour column's
// 3-row significance window (and, where a coefficient became significant, one of
// its sign bit); codeword slots: one 64-bit ballot per coder call site, ranked per
// 32-lane half with mbcnt, and two wave-uniform (SGPR) slot counters, one per half.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "bpc_internal.h"

namespace bpc {
namespace {

typedef uint64_t u64;
typedef uint32_t u32;
// Image, codestream and LUT pointers in the global address space: the batched
// kernel reads them from its image table, where the compiler cannot tell them
// from LDS or scratch pointers and would emit flat accesses.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
typedef gptr<int32_t> gi32;
typedef gptr<const int32_t> cgi32;

constexpr int N_SIG = 9;   // N_CONTEXT_SIGNIFICANCE (common.h:33)
constexpr int N_SIGN = 4;  // N_CONTEXT_SIGN (common.h:32)

// Sign context of (sgn h, sgn v), BPC_PaCo.cu:178-211, as 4-bit entries at index
// 3*hs + vs with hs, vs in {0: zero, 1: positive, 2: negative}.
constexpr u64 SIGN_CTX_TABLE = 0x715064320ull;

// Sign context per 8-bit neighbour state, as an LDS table built at kernel start.
// Index bits: 0 own other column's sign and 1 its significance (row i); 2 sign and
// 3 significance of the row above, 4 sign and 5 significance of the row below (own
// column); 6 sign and 7 significance of the neighbour lane's column (row i).  A sign
// bit counts only with its significance bit.  Entry: the context (0..3) in bits 3-4,
// the predicted-sign flip in bit 5 (chi = 0 not significant, -1 negative, +1
// positive; h and v are the sums of the horizontal and vertical pairs,
// BPC_PaCo.cu:178-226).
__device__ __forceinline__ int chi(u32 sig, u32 neg) { return sig ? (neg ? -1 : 1) : 0; }
__device__ __forceinline__ uint8_t sign_lut_entry(u32 k) {
  const int h = chi((k >> 1) & 1u, k & 1u) + chi((k >> 7) & 1u, (k >> 6) & 1u);
  const int v = chi((k >> 3) & 1u, (k >> 2) & 1u) + chi((k >> 5) & 1u, (k >> 4) & 1u);
  const u32 hs = h > 0 ? 1u : (h < 0 ? 2u : 0u), vs = v > 0 ? 1u : (v < 0 ? 2u : 0u);
  const u32 s4 = (u32)(SIGN_CTX_TABLE >> (4 * (3 * hs + vs))) & 0xFu;
  return (uint8_t)(((s4 >> 1) << 3) | ((s4 & 1u) << 5));
}
typedef __attribute__((address_space(3))) const uint8_t* lds_u8;
typedef __attribute__((address_space(3))) uint8_t* lds_w8;
typedef __attribute__((address_space(3))) uint32_t* lds_w32;
constexpr int PROW_BYTES = 20;  // a lane's row: 9 significance, 9 cleanup-pass probabilities
__device__ __forceinline__ lds_u8 build_sign_lut(uint8_t* t) {
  for (u32 k = threadIdx.x; k < 256; k += blockDim.x) t[k] = sign_lut_entry(k);
  __syncthreads();
  return (lds_u8)t;
}

struct PlaneLut {  // probabilities of one bitplane, 7 bits each
  u64 sig;         // 9 significance contexts, entry k at bits 7k (per-image kernel; the
  u64 cpsig;       // batched kernel keeps them in the lane's LDS row, load_plane_lut);
                   // 3-pass SPP: entries 1..8 as bytes 0..7 (sig_prob)
  u32 sign;        // 4 sign contexts, entry k at bits 8k
  u32 cpsign;
  u32 ref;         // refinement probability
  // symbol-count mode: LUT entries of context 0 of this plane, cleanup-pass offset,
  // and whether the plane has LUT rows (planes >= the LUT's bitplanes are not counted)
  int isig, isign, iref, icp;
  bool tal;
};

template <int MAXP>
struct LaneState {
  u64 S[2], R[2], C[2], G[2];  // row masks per column (0 even, 1 odd): significant, refine, cleanup, sign
  u64 cur[2];                  // decoder: bits decoded in the current plane
  u32 L, Sz, slot, cw;         // coder: interval lower, interval size, reserved slot, codeword
  u32 cnt0, cnt1;              // next free slot of each half (wave-uniform)
};

// Lane mask of a predicate, taken straight from the compare (no 0/1 round trip).
__device__ __forceinline__ u64 ballot(bool b) { return __builtin_amdgcn_ballot_w64(b); }

// Lane masks of a compare, written out so that the mask stays in an SGPR pair and
// combines with other masks on the scalar unit (a ballot of a combined predicate
// would otherwise be rebuilt through a 0/1 VGPR).  Volatile: the compare must run
// with the whole wave active, where the caller put it.
__device__ __forceinline__ u64 lanes_eq0(u32 v) {
  u64 r;
  asm volatile("v_cmp_eq_u32_e64 %0, 0, %1" : "=s"(r) : "v"(v));
  return r;
}
__device__ __forceinline__ u64 lanes_ne0(u32 v) {
  u64 r;
  asm volatile("v_cmp_ne_u32_e64 %0, 0, %1" : "=s"(r) : "v"(v));
  return r;
}
__device__ __forceinline__ bool lane_in(u64 m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
// t + 1 in the lanes of m, t elsewhere: one add with the mask as carry-in
__device__ __forceinline__ u32 add_if(u32 t, u64 m) {
  u32 r;
  u64 co;
  asm("v_addc_co_u32_e64 %0, %1, %2, 0, %3" : "=v"(r), "=s"(co) : "v"(t), "s"(m));
  return r;
}

// Hides a value from the optimizer so that a later compare on it stays a single
// v_cmp whose lane mask feeds a ballot directly (instead of being folded into a
// combination of masks that must be rebuilt through a 0/1 VGPR).
__device__ __forceinline__ u32 opaque(u32 v) {
  asm volatile("" : "+v"(v));
  return v;
}

__device__ __forceinline__ u32 dpp_from_left(u32 v) {  // lane l receives lane l-1
  return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, true);  // wave_shr:1
}
__device__ __forceinline__ u32 dpp_from_right(u32 v) {  // lane l receives lane l+1
  return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xF, 0xF, true);  // wave_shl:1
}

// Find_Subband (BPC_PaCo.cu:84-111) as a LUT group index level*3 + subband.
__device__ __forceinline__ int find_group(int x, int y, int X, int Y, int levels) {
  for (int aux = 1; aux <= levels; ++aux) {
    const bool bx = x >= (X >> aux), by = y >= (Y >> aux);
    if (bx || by) return (aux - 1) * 3 + (bx ? (by ? 2 : 0) : 1);
  }
  return levels * 3;
}

struct Geo {
  int half;        // 0: lanes 0-31, 1: lanes 32-63
  int hl;          // lane within the half
  int64_t region;  // first word of the codeblock's region
  int stride;
  int rows;
  u32 edge_even;   // 7, or 0 for the first lane of a half (no left neighbour in the codeblock)
  u32 edge_odd;    // 7, or 0 for the last lane of a half (no right neighbour)
  unsigned long long* hist;  // symbol-count mode: per LUT entry, (symbols << 32) + zeros, in LDS
  lds_u8 slut;               // sign contexts (build_sign_lut)
  lds_u8 prow;               // this lane's significance probabilities of the plane: [0, 9) SPP, [9, 18) CP
  // branch-free coder: the wave's codestream window (both halves' regions) and the
  // byte offset of this lane's region in it
  __amdgpu_buffer_rsrc_t cs;
  u32 cs_off;
};
// Buffer offset past every window: a store there is dropped, a load reads 0.
constexpr u32 CS_DISCARD = 0xFFFFFFF0u;

// Symbol-count mode sink: one symbol coded with the probability of LUT entry idx.
__device__ __forceinline__ void tally(unsigned long long* hist, int idx, u32 sym) {
  atomicAdd(hist + idx, (1ull << 32) | (sym ? 0ull : 1ull));
}

// Where the decoder reads codeword slot k (1 <= k < stride) of its codeblock: the
// reference's region (word k after the MSB word, BPC_PaCo.cu:1130,1216), or the
// compact container (payload[offsets[cb] + k - 1] for k up to the codeblock's
// count, -1 past it: the words a region unpacked from the container holds).
struct CompactCw {
  const uint16_t* pay;  // this half's codewords
  u32 n;                // codewords of this half's codeblock
};
__device__ __forceinline__ u32 codeword(cgi32 src, const Geo& g, u32 slot) { return (u32)src[g.region + slot]; }
__device__ __forceinline__ u32 codeword(const CompactCw& c, const Geo&, u32 slot) {
  return slot <= c.n ? (u32)c.pay[slot - 1] : 0xFFFFFFFFu;
}

// Slot reservation for one coder call site (BPC_PaCo.cu:268-277 / :304-316):
// lanes at the site whose interval is exhausted take consecutive slots in lane
// order after the half's counter; the counter advances by their number.  The
// counters stay wave-uniform (SGPRs).
template <int DEC, int MAXP, typename SRC>
__device__ __forceinline__ void reserve(LaneState<MAXP>& s, bool at, u64 atm, const Geo& g, SRC src) {
  // the call site's lanes whose interval is exhausted.  The encoder has the site's
  // lanes as a wave mask and combines it with one compare on the scalar unit; the
  // decoder has them as a predicate and feeds one compare (0 exactly for the lanes
  // that reserve) to the ballot.
  const u32 szr = DEC ? opaque(at ? s.Sz : 1u) : 1u;
  const u64 m = DEC ? ballot(szr == 0) : lanes_eq0(s.Sz) & atm;
  if (m) {
    const u32 lo = (u32)m, hi = (u32)(m >> 32);
    const u32 nlo = __builtin_popcount(lo);
    // mbcnt over the whole wave counts, for an upper-half lane, every lower-half
    // lane too; its base is lowered by that count
    const u32 base = g.half ? s.cnt1 - nlo : s.cnt0;
    const u32 slot = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, base));
    if (DEC ? szr == 0 : lane_in(m)) {
      // A slot past the region (overflow, reported from the final counter) is
      // redirected to the region's last word so that no access leaves the region.
      s.slot = min(slot, (u32)g.stride - 1u);
      s.L = 0;
      s.Sz = 0xFFFFu;
      if (DEC) s.cw = codeword(src, g, s.slot);
    }
    s.cnt0 = __builtin_amdgcn_readfirstlane(s.cnt0 + nlo);
    s.cnt1 = __builtin_amdgcn_readfirstlane(s.cnt1 + __builtin_popcount(hi));
  }
}

// (a << SH) | c as one v_lshl_or_b32 (the compiler would otherwise form two
// shifts and a three-way OR)
template <int SH>
__device__ __forceinline__ u32 lshl_or(u32 a, u32 c) {
  u32 r;
  asm("v_lshl_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "i"(SH), "v"(c));
  return r;
}

// (a & b) | c as one v_and_or_b32
__device__ __forceinline__ u32 and_or(u32 a, u32 b, u32 c) {
  u32 r;
  asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
  return r;
}

// a * b + c with 24-bit a, b (one v_mad_u32_u24; written out because the compiler
// otherwise widens a multiply by a 0/1 value into a 64-bit multiply-add)
__device__ __forceinline__ u32 mad24(u32 a, u32 b, u32 c) {
  u32 r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// Significance probability of context ctx from the plane's packed row.  The 3-pass
// SPP codes only contexts 1..8 (a coefficient without a significant neighbour is
// left to the cleanup pass), which fit one 8-byte row: one byte select (v_perm_b32)
// instead of a 64-bit shift and a mask.  Lanes with context 0 get a value that no
// kept result depends on.
template <int CLEAN, int CP3>
__device__ __forceinline__ u32 sig_prob(const PlaneLut& pl, u32 ctx) {
  if (!CLEAN && CP3) return __builtin_amdgcn_perm((u32)(pl.sig >> 32), (u32)pl.sig, ctx + 0x0C0C0BFFu);
  return (u32)((CLEAN ? pl.cpsig : pl.sig) >> (7u * ctx)) & 127u;
}

// Encoder coder step (BPC_PaCo.cu:281-294).
template <int MAXP>
__device__ __forceinline__ void enc_symbol(LaneState<MAXP>& s, u32 sym, u32 prob, const Geo& g, gi32 dst) {
  const u32 t = __umul24(s.Sz, prob) >> 7;  // Sz < 2^16, prob < 2^7
  const u32 aux = t + sym;                   // sym is 0 or 1
  s.Sz = sym ? s.Sz - aux : t;
  s.L = mad24(aux, sym, s.L);  // L + aux on a 1 (aux < 2^16: exact in 24 bits)
  if (s.Sz == 0) dst[g.region + s.slot] = (int32_t)s.L;
}

// Branch-free forms of the two coder steps, for a wave that is the critical path
// of its launch: every lane computes the step and the lanes at the call site keep
// it by select, so no exec-mask branch splits the scan step.  A closed codeword is
// stored through the codestream window; the other lanes' stores fall outside the
// window and are dropped.
template <int MAXP>
__device__ __forceinline__ void enc_sel(LaneState<MAXP>& s, bool at, u32 sym, u32 prob, const Geo& g) {
  const u32 t = __umul24(s.Sz, prob) >> 7;
  const u32 aux = t + sym;
  const u32 nsz = sym ? s.Sz - aux : t;
  const u32 nl = mad24(aux, sym, s.L);
  s.Sz = at ? nsz : s.Sz;
  s.L = at ? nl : s.L;
  const bool close = at && nsz == 0;
  __builtin_amdgcn_raw_buffer_store_b32(s.L, g.cs, close ? g.cs_off + s.slot * 4u : CS_DISCARD, 0, 0);
}
template <int MAXP>
__device__ __forceinline__ void enc_sel_m(LaneState<MAXP>& s, bool at, u64 atm, u64 pm, u32 prob, const Geo& g) {
  const u32 t = __umul24(s.Sz, prob) >> 7;
  const u32 aux = add_if(t, pm);
  const u32 nsz = lane_in(pm) ? s.Sz - aux : t;
  s.L = lane_in(atm & pm) ? s.L + aux : s.L;
  s.Sz = at ? nsz : s.Sz;
  const bool close = at && nsz == 0;
  __builtin_amdgcn_raw_buffer_store_b32(s.L, g.cs, close ? g.cs_off + s.slot * 4u : CS_DISCARD, 0, 0);
}
template <int MAXP>
__device__ __forceinline__ u32 dec_sel(LaneState<MAXP>& s, bool at, u32 prob) {
  const u32 aux = (__umul24(s.Sz, prob) >> 7) + 1u;
  const u32 aux2 = s.L + aux;
  const bool one = s.cw >= aux2;
  const u32 nsz = one ? s.Sz - aux : aux - 1u;
  const u32 nl = one ? aux2 : s.L;
  s.Sz = at ? nsz : s.Sz;
  s.L = at ? nl : s.L;
  return (at && one) ? 1u : 0u;
}

// Decoder coder step (BPC_PaCo.cu:321-337).
template <int MAXP>
__device__ __forceinline__ u32 dec_symbol(LaneState<MAXP>& s, u32 prob) {
  const u32 aux = (__umul24(s.Sz, prob) >> 7) + 1u;
  const u32 aux2 = s.L + aux;
  if (s.cw >= aux2) {
    s.Sz -= aux;
    s.L = aux2;
    return 1u;
  }
  s.Sz = aux - 1u;
  return 0u;
}

// Row i of a 64-row column mask is bit i - 32 H of its 32-bit word H = i / 32.  The
// scan row is wave-uniform, so each pass runs one loop per word and every bit test
// and update of a step touches one 32-bit word.
template <int H>
__device__ __forceinline__ u32 word(u64 x) {
  return H ? (u32)(x >> 32) : (u32)x;
}
template <int H>
__device__ __forceinline__ void set_bit(u64& x, u32 b) {
  x |= H ? ((u64)b << 32) : (u64)b;
}
template <int H>
__device__ __forceinline__ void clear_bit(u64& x, u32 b) {
  x &= H ? ~((u64)b << 32) : ~(u64)b;
}
// 3-row window of a column mask: rows i-1, i, i+1 at bits 0..2 (rows -1 and 64
// are 0).  W = 0: row 0; W = 1: rows 31 and 32 (one funnel shift across the
// words); W = 2: rows 33..63 (one field extract from the upper word); W = 3: rows
// 1..30 (one field extract from the lower word).
template <int W>
__device__ __forceinline__ u32 window(u64 x, int i) {
  if (W == 0) return ((u32)x << 1) & 7u;
  if (W == 1) return __builtin_amdgcn_alignbit((u32)(x >> 32), (u32)x, (u32)(i - 1)) & 7u;
  if (W == 3) return __builtin_amdgcn_ubfe((u32)x, (u32)(i - 1), 3u);
  return __builtin_amdgcn_ubfe((u32)(x >> 32), (u32)(i - 33), 3u);
}

// One scan step of the significance pass (CLEAN = 0, SPP_Encoder/SPP_Decoder,
// BPC_PaCo.cu:412-507) or of the cleanup pass (CLEAN = 1, CP_Encoder/CP_Decoder,
// BPC_PaCo.cu:510-589) at row i (in word H, window form W) of column ODD.
template <int ODD, int H, int W, int DEC, int CP3, int CLEAN, int TRAIN, int BF, int MAXP, typename SRC>
__device__ __forceinline__ void sig_step(LaneState<MAXP>& s, int i, bool act, u64 actm, u64 plane, const PlaneLut& pl,
                                         const Geo& g, SRC src, gi32 dst) {
  constexpr int c = ODD, o = 1 - ODD;
  const int ib = i - 32 * H;
  const u32 bit = 1u << ib;
  // 3-row windows of the other and the own column
  const u32 wo = window<W>(s.S[o], i);
  const u32 wc = window<W>(s.S[c], i);
  // the neighbour lane's column (left lane's odd column for an even step, right
  // lane's even column for an odd step; BPC_PaCo.cu:342-388); lanes at the
  // codeblock edge see zeros (Correct_CB_borders, BPC_PaCo.cu:390-410)
  const u32 nb = (ODD ? dpp_from_right(wo) : dpp_from_left(wo)) & (ODD ? g.edge_odd : g.edge_even);
  // context = significant neighbours; wc includes the coefficient itself, which is
  // not significant wherever a context is used
  const u32 nbr = lshl_or<6>(nb, lshl_or<3>(wc, wo));
  const u32 ctx = __builtin_popcount(nbr);

  // the lanes that code at this step: the decoder forms the predicate, the encoder
  // the wave mask (compares into SGPR pairs, combined on the scalar unit, which
  // leaves fewer VALU instructions on the encoder's scan step)
  u64 atm = 0;
  bool at;
  if (DEC) {
    if (!CLEAN) {
      const bool cand = act && !(wc & 2u);
      if (CP3) {
        const bool gate = ctx != 0;  // BPC_PaCo.cu:424
        set_bit<H>(s.C[c], (cand && !gate ? 1u : 0u) << ib);  // BPC_PaCo.cu:424, without a branch
        at = cand && gate;
      } else {
        at = cand;
      }
    } else {
      at = act && ((word<H>(s.C[c]) >> ib) & 1u);  // BPC_PaCo.cu:515
      clear_bit<H>(s.C[c], (at ? 1u : 0u) << ib);   // BPC_PaCo.cu:528, without a branch
    }
  } else if (!CLEAN) {
    const u64 candm = actm & lanes_eq0(wc & 2u);
    if (CP3) {
      const u64 gatem = lanes_ne0(nbr);  // context != 0, BPC_PaCo.cu:424
      set_bit<H>(s.C[c], (lane_in(candm & ~gatem) ? 1u : 0u) << ib);  // BPC_PaCo.cu:424, without a branch
      atm = candm & gatem;
    } else {
      atm = candm;
    }
  } else {
    atm = actm & lanes_ne0(word<H>(s.C[c]) & bit);  // BPC_PaCo.cu:515
    clear_bit<H>(s.C[c], (lane_in(atm) ? 1u : 0u) << ib);  // BPC_PaCo.cu:528, without a branch
  }
  if (!DEC) at = lane_in(atm);
  // site 1: significance symbol
  if (!TRAIN) reserve<DEC>(s, at, atm, g, src);
  u32 sym = 0;  // decoder: the decoded symbol
  // encoder: the symbol, known before coding, at its row bit (and, for the
  // branch-free coder, as a lane mask)
  const u32 pb = DEC || !BF ? 0u : word<H>(plane) & bit;
  const u64 pm = DEC || !BF ? 0ull : lanes_ne0(pb);
  if (!DEC) {
    const u32 pbit = (word<H>(plane) >> ib) & 1u;
    if (TRAIN) {
      if (at && pl.tal) tally(g.hist, (CLEAN ? pl.isig + pl.icp : pl.isig) + (int)ctx, pbit);
    } else if (BF) {
      enc_sel_m(s, at, atm, pm, sig_prob<CLEAN, CP3>(pl, ctx), g);
    } else if (at) {
      enc_symbol(s, pbit, (u32)g.prow[(CLEAN ? 9u : 0u) + ctx], g, dst);
    }
  } else if (BF) {
    sym = dec_sel(s, at, sig_prob<CLEAN, CP3>(pl, ctx));
  } else if (at) {
    sym = dec_symbol(s, (u32)g.prow[(CLEAN ? 9u : 0u) + ctx]);
  }
  // a 1 makes the coefficient significant (and, in CP, refined: BPC_PaCo.cu:534;
  // in the decoder it is magnitude bit i of this plane), without a branch.  The
  // encoder forms the symbol at its row bit straight from the plane word (one AND
  // and one select).  The branch-free decoder's symbol is a compare result whose
  // ballot is free; the branchy decoder's is hidden from the optimizer so that its
  // ballot stays one compare.
  const u32 symb = DEC ? sym << ib : (at ? (BF ? pb : word<H>(plane) & bit) : 0u);  // the symbol at its row bit
  set_bit<H>(s.S[c], symb);
  if (CLEAN) set_bit<H>(s.R[c], symb);
  if (DEC) set_bit<H>(s.cur[c], symb);
  const u32 coded1 = DEC ? (BF ? sym : opaque(sym)) : symb;
  const u64 nsm = DEC ? ballot(coded1 != 0) : (BF ? atm & pm : lanes_ne0(symb));
  // site 2: sign symbol of the coefficients that became significant
  if (nsm) {
    const bool nsig = DEC ? coded1 != 0 : lane_in(nsm);
    // neighbour lane's sign bit at row i, moved like its significance window
    const u32 gs = (word<H>(s.G[o]) >> ib) & 1u;
    const u32 gn = ODD ? dpp_from_right(gs) : dpp_from_left(gs);
    // the 8 neighbour bits (build_sign_lut): significance bits straight from the
    // context word, sign bits from the own column's sign window and the two row-i
    // sign bits; the context and flip come from one LDS byte
    const u32 idx = lshl_or<6>(gn, lshl_or<2>(window<W>(s.G[c], i) & 5u, and_or(nbr, 0xAAu, gs)));
    const u32 e = g.slut[idx];
    const u32 ssh = e & 24u;   // 8 x sign context: its probability's bit offset
    const u32 flip = e >> 5;   // predicted sign
    if (!TRAIN) reserve<DEC>(s, nsig, nsm, g, src);
    if (BF && !TRAIN) {
      const u32 prob = ((CLEAN ? pl.cpsign : pl.sign) >> ssh) & 127u;
      const u32 own = (word<H>(s.G[c]) >> ib) & 1u;
      if (!DEC) {
        enc_sel(s, nsig, own ^ flip, prob, g);
      } else {
        const u32 v = dec_sel(s, nsig, prob) ^ flip;
        set_bit<H>(s.G[c], (nsig ? v : 0u) << ib);
      }
    } else {
      u32 v = 0;  // decoder: the sign bit, set outside the branch (no merge of the mask)
      if (nsig) {
        const u32 prob = ((CLEAN ? pl.cpsign : pl.sign) >> ssh) & 127u;
        const u32 own = (word<H>(s.G[c]) >> ib) & 1u;
        if (TRAIN) {
          if (pl.tal) tally(g.hist, (CLEAN ? pl.isign + pl.icp : pl.isign) + (int)(ssh >> 3), own ^ flip);
        } else if (!DEC) {
          enc_symbol(s, own ^ flip, prob, g, dst);
        } else {
          v = dec_symbol(s, prob) ^ flip;
        }
      }
      if (DEC) set_bit<H>(s.G[c], v << ib);
    }
  }
}

// One scan step of the refinement pass (MRP_Encoder/MRP_Decoder, BPC_PaCo.cu:593-629):
// the coefficients with the refine flag code one bit each.  The reference's
// "significant without refine flag -> set it" (BPC_PaCo.cu:604) touches only the
// step's own coefficient, whose flag no later step of the pass reads, so the pass
// applies it to all coefficients at its end (ref_pass).
template <int ODD, int H, int DEC, int TRAIN, int BF, int MAXP, typename SRC>
__device__ __forceinline__ void ref_step(LaneState<MAXP>& s, int i, bool act, u64 actm, u64 plane, const PlaneLut& pl,
                                         const Geo& g, SRC src, gi32 dst) {
  constexpr int c = ODD;
  const int ib = i - 32 * H;
  const u64 atm = DEC ? 0 : actm & lanes_ne0(word<H>(s.R[c]) & (1u << ib));
  const bool at = DEC ? act && ((word<H>(s.R[c]) >> ib) & 1u) : lane_in(atm);
  if (!TRAIN) reserve<DEC>(s, at, atm, g, src);
  if (BF && !TRAIN) {
    if (!DEC)
      enc_sel(s, at, (word<H>(plane) >> ib) & 1u, pl.ref, g);
    else
      set_bit<H>(s.cur[c], dec_sel(s, at, pl.ref) << ib);
  } else {
    u32 d = 0;  // decoder: the refinement bit, set outside the branch
    if (at) {
      if (TRAIN) {
        if (pl.tal) tally(g.hist, pl.iref, (word<H>(plane) >> ib) & 1u);
      } else if (!DEC) {
        enc_symbol(s, (word<H>(plane) >> ib) & 1u, pl.ref, g, dst);
      } else {
        d = dec_symbol(s, pl.ref);
      }
    }
    if (DEC) set_bit<H>(s.cur[c], d << ib);
  }
}

// OR of a 32-bit value over the wave, wave-uniform: row_shr 1/2/4/8 leave each
// 16-lane row's OR in its lane 15, row_bcast:15 folds rows 0 and 2 into rows 1
// and 3, row_bcast:31 folds row 1 into rows 2 and 3; lane 63 then holds the OR of
// every lane.  All DPP, no LDS round trips (lanes without a source read 0).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ u32 dpp_or(u32 v) {
  return v | (u32)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, true);
}
__device__ __forceinline__ u64 wave_or(u64 x) {
  u32 lo = (u32)x, hi = (u32)(x >> 32);
  lo = dpp_or<0x111, 0xF>(lo);  // row_shr:1
  hi = dpp_or<0x111, 0xF>(hi);
  lo = dpp_or<0x112, 0xF>(lo);  // row_shr:2
  hi = dpp_or<0x112, 0xF>(hi);
  lo = dpp_or<0x114, 0xF>(lo);  // row_shr:4
  hi = dpp_or<0x114, 0xF>(hi);
  lo = dpp_or<0x118, 0xF>(lo);  // row_shr:8
  hi = dpp_or<0x118, 0xF>(hi);
  lo = dpp_or<0x142, 0xA>(lo);  // row_bcast:15 into rows 1 and 3
  hi = dpp_or<0x142, 0xA>(hi);
  lo = dpp_or<0x143, 0xC>(lo);  // row_bcast:31 into rows 2 and 3
  hi = dpp_or<0x143, 0xC>(hi);
  return ((u64)__builtin_amdgcn_readlane(hi, 63) << 32) | (u32)__builtin_amdgcn_readlane(lo, 63);
}

// Steps whose row holds no coefficient that can take part in the pass, in any lane
// of the wave, change nothing and are skipped.  Row sets are taken at pass start
// and are supersets of the rows that take part: SPP codes only coefficients that
// are not significant (significance only grows, one coefficient per step); CP
// only those with the cleanup flag (never set during CP); MRP only those with the
// refine flag (never set during MRP).  Rows 0-31 and 32-63 run as two loops.
template <int DEC, int CP3, int CLEAN, int TRAIN, int BF, int MAXP, typename SRC>
__device__ __forceinline__ void sig_pass(LaneState<MAXP>& s, bool act, const u64* plane, const PlaneLut& pl,
                                         const Geo& g, SRC src, gi32 dst) {
  const u64 rowmask = g.rows >= 64 ? ~0ull : ((1ull << g.rows) - 1ull);
  const u64 actm = DEC ? 0 : ballot(act);  // the encoder's steps take the active lanes as a mask
  const u64 mE = wave_or(act ? (CLEAN ? s.C[0] : (~s.S[0] & rowmask)) : 0ull);
  const u64 mO = wave_or(act ? (CLEAN ? s.C[1] : (~s.S[1] & rowmask)) : 0ull);
  const u32 eL = (u32)mE, oL = (u32)mO, eH = (u32)(mE >> 32), oH = (u32)(mO >> 32);
  if ((eL | oL) & 1u) {
    if (eL & 1u) sig_step<0, 0, 0, DEC, CP3, CLEAN, TRAIN, BF>(s, 0, act, actm, plane[0], pl, g, src, dst);
    if (oL & 1u) sig_step<1, 0, 0, DEC, CP3, CLEAN, TRAIN, BF>(s, 0, act, actm, plane[1], pl, g, src, dst);
  }
  for (u32 r = (eL | oL) & 0x7FFFFFFEu; r; r &= r - 1u) {
    const int i = __builtin_ctz(r);
    if ((eL >> i) & 1u) sig_step<0, 0, 3, DEC, CP3, CLEAN, TRAIN, BF>(s, i, act, actm, plane[0], pl, g, src, dst);
    if ((oL >> i) & 1u) sig_step<1, 0, 3, DEC, CP3, CLEAN, TRAIN, BF>(s, i, act, actm, plane[1], pl, g, src, dst);
  }
  if ((eL | oL) >> 31) {  // row 31: its window reaches row 32 in the upper word
    if (eL >> 31) sig_step<0, 0, 1, DEC, CP3, CLEAN, TRAIN, BF>(s, 31, act, actm, plane[0], pl, g, src, dst);
    if (oL >> 31) sig_step<1, 0, 1, DEC, CP3, CLEAN, TRAIN, BF>(s, 31, act, actm, plane[1], pl, g, src, dst);
  }
  if ((eH | oH) & 1u) {
    if (eH & 1u) sig_step<0, 1, 1, DEC, CP3, CLEAN, TRAIN, BF>(s, 32, act, actm, plane[0], pl, g, src, dst);
    if (oH & 1u) sig_step<1, 1, 1, DEC, CP3, CLEAN, TRAIN, BF>(s, 32, act, actm, plane[1], pl, g, src, dst);
  }
  for (u32 r = (eH | oH) & ~1u; r; r &= r - 1u) {
    const int i = 32 + __builtin_ctz(r);
    if ((eH >> (i - 32)) & 1u) sig_step<0, 1, 2, DEC, CP3, CLEAN, TRAIN, BF>(s, i, act, actm, plane[0], pl, g, src, dst);
    if ((oH >> (i - 32)) & 1u) sig_step<1, 1, 2, DEC, CP3, CLEAN, TRAIN, BF>(s, i, act, actm, plane[1], pl, g, src, dst);
  }
}

template <int DEC, int TRAIN, int BF, int MAXP, typename SRC>
__device__ __forceinline__ void ref_pass(LaneState<MAXP>& s, bool act, const u64* plane, const PlaneLut& pl,
                                         const Geo& g, SRC src, gi32 dst) {
  const u64 actm = DEC ? 0 : ballot(act);  // the encoder's steps take the active lanes as a mask
  const u64 mE = wave_or(act ? s.R[0] : 0ull);  // rows with a refinement symbol in some lane
  const u64 mO = wave_or(act ? s.R[1] : 0ull);
  const u32 eL = (u32)mE, oL = (u32)mO, eH = (u32)(mE >> 32), oH = (u32)(mO >> 32);
  for (u32 r = eL | oL; r; r &= r - 1u) {
    const int i = __builtin_ctz(r);
    if ((eL >> i) & 1u) ref_step<0, 0, DEC, TRAIN, BF>(s, i, act, actm, plane[0], pl, g, src, dst);
    if ((oL >> i) & 1u) ref_step<1, 0, DEC, TRAIN, BF>(s, i, act, actm, plane[1], pl, g, src, dst);
  }
  for (u32 r = eH | oH; r; r &= r - 1u) {
    const int j = __builtin_ctz(r);
    if ((eH >> j) & 1u) ref_step<0, 1, DEC, TRAIN, BF>(s, 32 + j, act, actm, plane[0], pl, g, src, dst);
    if ((oH >> j) & 1u) ref_step<1, 1, DEC, TRAIN, BF>(s, 32 + j, act, actm, plane[1], pl, g, src, dst);
  }
  if (act) {  // every significant coefficient refines from the next plane on (BPC_PaCo.cu:604)
    s.R[0] |= s.S[0];
    s.R[1] |= s.S[1];
  }
}

__device__ __forceinline__ u32 lut_at(cgi32 lut, int idx, bool& bad) {
  const int32_t v = lut[idx];
  bad |= (u32)v > 127u;
  return (u32)v & 127u;
}

// The plane's probabilities of the lane's LUT group (lanes of one codeblock can
// sit in different subbands).  ROWS = 1 (batched kernel): the significance
// probabilities go to the lane's LDS row as bytes (entry k of SPP at byte k, of CP
// at byte 9 + k), where a scan step reads one with a byte load; with 8 waves per
// SIMD that latency is hidden and the three VALU of a register lookup are saved.
// ROWS = 0 (per-image kernel, latency-bound): they stay packed in registers.
template <int CP3, int ROWS>
__device__ __forceinline__ PlaneLut load_plane_lut(cgi32 lut, int ref0, int sig0, int sign0, int cpoff,
                                                   int bp, bool& bad, lds_w8 row) {
  PlaneLut pl;
  pl.sig = 0;
  pl.cpsig = 0;
  pl.sign = 0;
  pl.cpsign = 0;
  u32 w[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < N_SIG; ++k) {
    const u32 v = lut_at(lut, sig0 + bp * N_SIG + k, bad);
    if (ROWS)
      w[k >> 2] |= v << (8 * (k & 3));
    else if (CP3)
      pl.sig |= k ? (u64)v << (8 * (k - 1)) : 0ull;  // context 0 never codes in the 3-pass SPP
    else
      pl.sig |= (u64)v << (7 * k);
  }
#pragma unroll
  for (int k = 0; k < N_SIGN; ++k) pl.sign |= lut_at(lut, sign0 + bp * N_SIGN + k, bad) << (8 * k);
  pl.ref = lut_at(lut, ref0 + bp, bad);
  if (CP3) {
#pragma unroll
    for (int k = 0; k < N_SIG; ++k) {
      const u32 v = lut_at(lut, cpoff + sig0 + bp * N_SIG + k, bad);
      if (ROWS)
        w[(N_SIG + k) >> 2] |= v << (8 * ((N_SIG + k) & 3));
      else
        pl.cpsig |= (u64)v << (7 * k);
    }
#pragma unroll
    for (int k = 0; k < N_SIGN; ++k) pl.cpsign |= lut_at(lut, cpoff + sign0 + bp * N_SIGN + k, bad) << (8 * k);
  }
  if (ROWS) {
    const lds_w32 r = (lds_w32)row;
#pragma unroll
    for (int q = 0; q < (CP3 ? 5 : 3); ++q) r[q] = w[q];
  }
  return pl;
}

// The image a wave codes and the codeblock of each half (both halves of a wave
// always code the same image, in the per-image and in the batched kernel).
struct HalfDesc {
  cgi32 src;
  gi32 dst;
  cgi32 lut;
  gi32 ncw;
  int size_x, size_y, nx, stride, nbp, levels, lut_len, layout, fill;
  int in16;  // coefficients as int16 instead of int32
  int cb;    // codeblock of this half in its image, -1 for an empty half
  // decoder reading the compact container (code_pair's CMP): MSB bytes, offsets,
  // payload and its length in codewords
  const uint8_t* cmsb;
  const unsigned long long* coff;
  const uint16_t* cpay;
  unsigned long long ccap;
  lds_u8 slut;  // sign contexts in LDS (build_sign_lut)
  lds_w8 prows; // this wave's LDS rows, PROW_BYTES per lane (unused in symbol-count mode)
};

// Codes the two codeblocks of one wave (Kernel_BPC_CODER / Kernel_BPC_DECODER
// per warp, BPC_PaCo.cu:1097-1224).
// The wave's bitplane scratch, addressed through a buffer resource: the plane
// offset goes in the scalar offset and the lane offset (lane * 8) in the vector
// offset, so no per-plane 64-bit addresses occupy VGPRs.  Plane p of rows 0-31 is
// at byte (p * 64) * 8, rows 32-63 at ((MAXP + p) * 64) * 8; one uint2 (even
// column, odd column) per lane.  Accesses outside the slot are range-checked.
constexpr int BUFFER_RSRC_WORD3 = 0x00020000;  // raw buffer, gfx9 family

template <int MAXP>
struct Scratch {
  __amdgpu_buffer_rsrc_t rsrc;
  u32 voff;
  __device__ __forceinline__ Scratch(void* base, int lane)
      : rsrc(__builtin_amdgcn_make_buffer_rsrc(base, (short)0, 2 * MAXP * 64 * 8, BUFFER_RSRC_WORD3)),
        voff((u32)lane * 8u) {}
  __device__ __forceinline__ void store(int slot, u32 even, u32 odd) const {
    __attribute__((ext_vector_type(2))) u32 v = {even, odd};
    __builtin_amdgcn_raw_buffer_store_b64(v, rsrc, voff, slot * 512, 0);
  }
  __device__ __forceinline__ uint2 load(int slot) const {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(rsrc, voff, slot * 512, 0);
    return make_uint2(v[0], v[1]);
  }
  // magnitude plane bp of both columns, all rows
  __device__ __forceinline__ void plane(int bp, u64* out) const {
    const uint2 lo = load(bp), hi = load(MAXP + bp);
    out[0] = (u64)lo.x | ((u64)hi.x << 32);
    out[1] = (u64)lo.y | ((u64)hi.y << 32);
  }
  // rows of half h (0: rows 0-31, 1: rows 32-63) of plane p, both columns
  __device__ __forceinline__ void put_half(int h, int p, u32 even, u32 odd) { store(h * MAXP + p, even, odd); }
  __device__ __forceinline__ uint2 half_plane(int h, int p) const { return load(h * MAXP + p); }
  // all rows of plane bp, both columns
  __device__ __forceinline__ void put_plane(int bp, u64 even, u64 odd) {
    store(bp, (u32)even, (u32)odd);
    store(MAXP + bp, (u32)(even >> 32), (u32)(odd >> 32));
  }
};

// Plane k of a register array, read by a wave-uniform index with a switch (static
// register indices in every case).
template <int MAXP>
__device__ __forceinline__ void plane_get(const u64 (&pl)[MAXP][2], int k, u64& e, u64& o) {
  static_assert(MAXP <= 20, "register planes hold at most 20 planes");
  e = o = 0;
  switch (k) {
      case 0: if (0 < MAXP) { e = pl[0 < MAXP ? 0 : 0][0]; o = pl[0 < MAXP ? 0 : 0][1]; } break;
      case 1: if (1 < MAXP) { e = pl[1 < MAXP ? 1 : 0][0]; o = pl[1 < MAXP ? 1 : 0][1]; } break;
      case 2: if (2 < MAXP) { e = pl[2 < MAXP ? 2 : 0][0]; o = pl[2 < MAXP ? 2 : 0][1]; } break;
      case 3: if (3 < MAXP) { e = pl[3 < MAXP ? 3 : 0][0]; o = pl[3 < MAXP ? 3 : 0][1]; } break;
      case 4: if (4 < MAXP) { e = pl[4 < MAXP ? 4 : 0][0]; o = pl[4 < MAXP ? 4 : 0][1]; } break;
      case 5: if (5 < MAXP) { e = pl[5 < MAXP ? 5 : 0][0]; o = pl[5 < MAXP ? 5 : 0][1]; } break;
      case 6: if (6 < MAXP) { e = pl[6 < MAXP ? 6 : 0][0]; o = pl[6 < MAXP ? 6 : 0][1]; } break;
      case 7: if (7 < MAXP) { e = pl[7 < MAXP ? 7 : 0][0]; o = pl[7 < MAXP ? 7 : 0][1]; } break;
      case 8: if (8 < MAXP) { e = pl[8 < MAXP ? 8 : 0][0]; o = pl[8 < MAXP ? 8 : 0][1]; } break;
      case 9: if (9 < MAXP) { e = pl[9 < MAXP ? 9 : 0][0]; o = pl[9 < MAXP ? 9 : 0][1]; } break;
      case 10: if (10 < MAXP) { e = pl[10 < MAXP ? 10 : 0][0]; o = pl[10 < MAXP ? 10 : 0][1]; } break;
      case 11: if (11 < MAXP) { e = pl[11 < MAXP ? 11 : 0][0]; o = pl[11 < MAXP ? 11 : 0][1]; } break;
      case 12: if (12 < MAXP) { e = pl[12 < MAXP ? 12 : 0][0]; o = pl[12 < MAXP ? 12 : 0][1]; } break;
      case 13: if (13 < MAXP) { e = pl[13 < MAXP ? 13 : 0][0]; o = pl[13 < MAXP ? 13 : 0][1]; } break;
      case 14: if (14 < MAXP) { e = pl[14 < MAXP ? 14 : 0][0]; o = pl[14 < MAXP ? 14 : 0][1]; } break;
      case 15: if (15 < MAXP) { e = pl[15 < MAXP ? 15 : 0][0]; o = pl[15 < MAXP ? 15 : 0][1]; } break;
      case 16: if (16 < MAXP) { e = pl[16 < MAXP ? 16 : 0][0]; o = pl[16 < MAXP ? 16 : 0][1]; } break;
      case 17: if (17 < MAXP) { e = pl[17 < MAXP ? 17 : 0][0]; o = pl[17 < MAXP ? 17 : 0][1]; } break;
      case 18: if (18 < MAXP) { e = pl[18 < MAXP ? 18 : 0][0]; o = pl[18 < MAXP ? 18 : 0][1]; } break;
      case 19: if (19 < MAXP) { e = pl[19 < MAXP ? 19 : 0][0]; o = pl[19 < MAXP ? 19 : 0][1]; } break;
      default: break;
  }
}
// Writes plane k by a compare on every index (a switch here would move the array
// to the stack).
template <int MAXP>
__device__ __forceinline__ void plane_put(u64 (&pl)[MAXP][2], int k, u64 e, u64 o) {
#pragma unroll
  for (int j = 0; j < MAXP; ++j)
    if (j == k) {
      pl[j][0] = e;
      pl[j][1] = o;
    }
}

// The same planes in registers: one 64-bit row mask per plane and column (64
// VGPRs at 16 planes), for the per-image launch when every wave it needs is
// resident at once (it then runs at two waves per SIMD at most, with registers to
// spare).  Planes are read and written by a wave-uniform index (plane_get /
// plane_put), so that the array keeps static register indices.
template <int MAXP>
struct RegPlanes {
  u64 pl[MAXP][2];
  __device__ __forceinline__ RegPlanes(void*, int) {
#pragma unroll
    for (int p = 0; p < MAXP; ++p) pl[p][0] = pl[p][1] = 0;
  }
  __device__ __forceinline__ void put_half(int h, int p, u32 even, u32 odd) {
    pl[p][0] |= (u64)even << (32 * h);
    pl[p][1] |= (u64)odd << (32 * h);
  }
  __device__ __forceinline__ uint2 half_plane(int h, int p) const {
    return make_uint2((u32)(pl[p][0] >> (32 * h)), (u32)(pl[p][1] >> (32 * h)));
  }
  __device__ __forceinline__ void plane(int bp, u64* out) const { plane_get<MAXP>(pl, bp, out[0], out[1]); }
  __device__ __forceinline__ void put_plane(int bp, u64 even, u64 odd) { plane_put<MAXP>(pl, bp, even, odd); }
};

// Coefficients of one lane's stripe (int32 or int16, two's complement) into the
// wave's bitplane scratch and the sign masks; returns the OR of the magnitudes.
template <typename T, int MAXP, typename PS>
__device__ __forceinline__ u32 read_planes(cgi32 src32, int64_t cbase, int rs, int rows, bool valid, PS& ps,
                                           LaneState<MAXP>& s) {
  const gptr<const T> src = reinterpret_cast<gptr<const T>>(src32);
  u32 mor = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int r0 = 32 * h, r1 = min(rows, 32 * h + 32);
    u32 plo[MAXP][2];
#pragma unroll
    for (int p = 0; p < MAXP; ++p) plo[p][0] = plo[p][1] = 0;
    u32 glo0 = 0, glo1 = 0;
#pragma unroll 4
    for (int r = r0; r < r1; ++r) {
      int v0 = 0, v1 = 0;
      if (valid) {
        const gptr<const T> q = src + cbase + (int64_t)r * rs;
        if (sizeof(T) == 2) {
          const u32 w = *reinterpret_cast<gptr<const u32>>(q);
          v0 = (int16_t)(w & 0xFFFFu);
          v1 = (int16_t)(w >> 16);
        } else {
          const u64 w = *reinterpret_cast<gptr<const u64>>(q);
          v0 = (int32_t)(u32)w;
          v1 = (int32_t)(u32)(w >> 32);
        }
      }
      const u32 m0 = v0 < 0 ? 0u - (u32)v0 : (u32)v0;
      const u32 m1 = v1 < 0 ? 0u - (u32)v1 : (u32)v1;
      mor |= m0 | m1;
      const int sh = r - r0;
      glo0 |= (u32)(v0 < 0) << sh;
      glo1 |= (u32)(v1 < 0) << sh;
#pragma unroll
      for (int p = 0; p < MAXP; ++p) {
        plo[p][0] |= ((m0 >> p) & 1u) << sh;
        plo[p][1] |= ((m1 >> p) & 1u) << sh;
      }
    }
    s.G[0] |= (u64)glo0 << (32 * h);
    s.G[1] |= (u64)glo1 << (32 * h);
    // plane p, rows of this half, both columns -> the plane store
#pragma unroll
    for (int p = 0; p < MAXP; ++p) ps.put_half(h, p, plo[p][0], plo[p][1]);
  }
  return mor;
}

// BF = 1: the branch-free coder steps (enc_sel / dec_sel).
// TRAIN = 1 (with DEC = 0): symbol-count mode.  The encoder's modelling runs
// unchanged, but every symbol is counted at the LUT entry whose probability would
// code it (hist, in LDS) instead of being coded; nothing is written to dst.
// RP = 1: the bitplanes in registers (RegPlanes) instead of the wave's scratch slot.
// CMP = 1 (decoder): the codestream comes from the compact container (MSB byte,
// codewords at payload[offsets[cb] ...]) instead of the codeblock regions.
template <int DEC, int CP3, int MAXP, int TRAIN = 0, int BF = 0, int RP = 0, int CMP = 0>
__device__ __forceinline__ void code_pair(const HalfDesc& a, const int rows, uint32_t* status_word, void* scr_base,
                                          unsigned long long* hist = nullptr) {
  const int lane = threadIdx.x & 63;
  typename std::conditional<RP != 0, RegPlanes<MAXP>, Scratch<MAXP>>::type scr(scr_base, lane);
  const int half = lane >> 5;
  const int hl = lane & 31;
  const int cb = a.cb;
  const bool valid = cb >= 0;
  const int cbc = valid ? cb : 0;
  const int tx = cbc % a.nx, ty = cbc / a.nx;
  const int x = tx * 64 + 2 * hl;  // TCoordinate_X, BPC_PaCo.cu:1126
  const int y0 = ty * rows;        // TCoordinate_Y, BPC_PaCo.cu:1127
  Geo g;
  g.half = half;
  g.hl = hl;
  g.region = (int64_t)cbc * a.stride;  // WCoordinate, BPC_PaCo.cu:1130
  g.stride = a.stride;
  g.rows = rows;
  g.edge_even = hl == 0 ? 0u : 7u;
  g.edge_odd = hl == 31 ? 0u : 7u;
  g.hist = hist;
  g.slut = a.slut;
  const lds_w8 myrow = a.prows + PROW_BYTES * lane;
  g.prow = (lds_u8)myrow;
  if (BF) {
    // the window spans both halves' regions (two codeblocks of one image)
    const bool v0 = __builtin_amdgcn_readlane((int)valid, 0), v1 = __builtin_amdgcn_readlane((int)valid, 32);
    const int64_t q0 = (int64_t)__builtin_amdgcn_readlane(cbc, 0) * a.stride;
    const int64_t q1 = (int64_t)__builtin_amdgcn_readlane(cbc, 32) * a.stride;
    const int64_t lo = !v0 ? q1 : (!v1 ? q0 : min(q0, q1));
    const int64_t hi = !v0 ? q1 : (!v1 ? q0 : max(q0, q1));
    int32_t* base = (int32_t*)(DEC ? const_cast<gi32>(a.src) : a.dst) + lo;
    g.cs = __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)((hi - lo + a.stride) * 4), BUFFER_RSRC_WORD3);
    g.cs_off = (u32)((g.region - lo) * 4);
  }
  int64_t cbase;
  int rs;
  if (a.layout) {
    cbase = (int64_t)cbc * 64 * rows + 2 * hl;
    rs = 64;
  } else {
    cbase = (int64_t)y0 * a.size_x + x;
    rs = a.size_x;
  }

  LaneState<MAXP> s;
#pragma unroll
  for (int c = 0; c < 2; ++c) s.S[c] = s.R[c] = s.C[c] = s.G[c] = s.cur[c] = 0;
  s.L = 0;
  s.Sz = 0;
  s.slot = 0;
  s.cw = 0;
  s.cnt0 = 1;
  s.cnt1 = 1;
  if (CP3) {  // every coefficient starts with the cleanup flag (BPC_PaCo.cu:34-39, 73-78)
    const u64 rm = rows >= 64 ? ~0ull : ((1ull << rows) - 1ull);
    s.C[0] = s.C[1] = rm;
  }

  u32 status = 0;
  int msb;
  bool good = valid;
  CompactCw ccw{nullptr, 0u};  // CMP: this half's codewords
  if (!DEC) {
    // Read_Coefficients (BPC_PaCo.cu:14-44) straight into bit-sliced planes.
    const u32 mor = a.in16 ? read_planes<int16_t>(a.src, cbase, rs, rows, valid, scr, s)
                           : read_planes<int32_t>(a.src, cbase, rs, rows, valid, scr, s);
    // Find_MSB (BPC_PaCo.cu:114-161): OR over the 32 lanes of the half.
    u32 mall = mor;
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) mall |= (u32)__shfl_xor((int)mall, off, 64);
    msb = mall ? 31 - __clz(mall) : 32;
    if (!TRAIN && valid && hl == 0) a.dst[g.region] = msb;  // BPC_PaCo.cu:1155
    if (msb != 32 && msb > 27) {
      if (!TRAIN) status |= ST_RANGE;  // symbol counting skips such a codeblock
      good = false;
    }
  } else if (CMP) {
    // the compact container's checks of compact_unpack_kernel: a count past the
    // region or the payload is an overflow and leaves no codewords
    if (valid) {
      const unsigned long long o = a.coff[cbc];
      long long n = (long long)(a.coff[cbc + 1] - o);
      if (n < 0 || n > a.stride - 1 || (unsigned long long)n > a.ccap || o > a.ccap - (unsigned long long)n) {
        status |= ST_OVERFLOW;
        n = 0;
      }
      ccw.pay = n ? a.cpay + o : nullptr;  // read only below slot n + 1
      ccw.n = (u32)n;
    }
    msb = valid ? (int)a.cmsb[cbc] : 32;  // 255 marks a malformed MSB word
    if (msb != 32 && (msb < 0 || msb > 27)) {
      status |= ST_MALFORMED;
      good = false;
    }
  } else {
    msb = valid ? a.src[g.region] : 32;  // BPC_PaCo.cu:1216
    if (msb != 32 && (msb < 0 || msb > 27)) {
      status |= ST_MALFORMED;
      good = false;
    }
  }
  if (msb == 32) good = false;  // all-zero codeblock: nothing coded
  if (good && msb >= MAXP) {
    status |= ST_WIDE;
    good = false;
  }

  // LUT row pointers (Initialize_LUT_pointers, BPC_PaCo.cu:228-250), plane 0 rows.
  const int G = 3 * a.levels + 1;
  const int grp = find_group(x, y0, a.size_x, a.size_y, a.levels);
  const int ref0 = grp * a.nbp;
  const int sig0 = G * a.nbp + grp * a.nbp * N_SIG;
  const int sign0 = G * a.nbp * (1 + N_SIG) + grp * a.nbp * N_SIGN;
  const int cpoff = G * a.nbp * (N_SIG + N_SIGN);  // BPC_PaCo.cu:988-989
  if (good && !TRAIN) {
    const int cp = CP3 ? cpoff : 0;
    if (ref0 + msb + 1 > a.lut_len || sig0 + cp + msb * N_SIG + N_SIG > a.lut_len ||
        sign0 + cp + msb * N_SIGN + N_SIGN > a.lut_len) {
      status |= ST_RANGE;
      good = false;
    }
  }
  // a codeblock is coded only if every lane of its half is good
  {
    const u64 ok = ballot(good);
    const u64 hm = half ? 0xFFFFFFFF00000000ull : 0x00000000FFFFFFFFull;
    const u64 vm = ballot(valid);
    good = good && ((ok & hm) == (vm & hm));
  }
  const int msb_eff = good ? msb : -1;
  const int top = max(__builtin_amdgcn_readlane(msb_eff, 0), __builtin_amdgcn_readlane(msb_eff, 32));

  // encoder: magnitude plane `top`, and the next one prefetched during each plane
  u64 plane[2] = {0, 0};
  // (the per-image kernel prefetches each next plane; the batched kernel, at 8 waves
  // per SIMD, loads each plane when it starts and spends no registers on a prefetch)
  if (!DEC && BF && !RP && top >= 0) scr.plane(top, plane);

  // Bitplane driver (Encode / Decode, BPC_PaCo.cu:968-1090).  In 3-pass mode the
  // reference's opening cleanup pass at MSB equals plane MSB with SPP and MRP
  // skipped (nothing is significant yet, so they code nothing).
  for (int bp = top; bp >= 0; --bp) {
    const bool act = good && bp <= msb;
    bool lbad = false;
    PlaneLut pl;
    pl.sig = pl.cpsig = 0;
    pl.sign = pl.cpsign = pl.ref = 0;
    pl.isig = pl.isign = pl.iref = pl.icp = 0;
    pl.tal = false;
    if (TRAIN) {
      pl.isig = sig0 + bp * N_SIG;
      pl.isign = sign0 + bp * N_SIGN;
      pl.iref = ref0 + bp;
      pl.icp = cpoff;
      pl.tal = bp < a.nbp;
    } else if (act) {
      pl = load_plane_lut<CP3, BF ? 0 : 1>(a.lut, ref0, sig0, sign0, cpoff, bp, lbad, myrow);
    }
    if (lbad) status |= ST_RANGE;
    u64 next[2] = {0, 0};
    if (!DEC && BF && !RP && bp > 0) scr.plane(bp - 1, next);
    if (!DEC && (!BF || RP)) scr.plane(bp, plane);
    const bool act_sr = act && !(CP3 && bp == msb);
    if constexpr (CMP != 0) {
      if (ballot(act_sr)) {
        sig_pass<DEC, CP3, 0, TRAIN, BF>(s, act_sr, plane, pl, g, ccw, a.dst);
        ref_pass<DEC, TRAIN, BF>(s, act_sr, plane, pl, g, ccw, a.dst);
      }
      if (CP3) sig_pass<DEC, CP3, 1, TRAIN, BF>(s, act, plane, pl, g, ccw, a.dst);
    } else {
      if (ballot(act_sr)) {
        sig_pass<DEC, CP3, 0, TRAIN, BF>(s, act_sr, plane, pl, g, a.src, a.dst);
        ref_pass<DEC, TRAIN, BF>(s, act_sr, plane, pl, g, a.src, a.dst);
      }
      if (CP3) sig_pass<DEC, CP3, 1, TRAIN, BF>(s, act, plane, pl, g, a.src, a.dst);
    }
    if (!DEC) {
      if (BF && !RP) {
        plane[0] = next[0];
        plane[1] = next[1];
      }
    } else {  // decoded magnitude bits of plane bp -> the wave's scratch
      scr.put_plane(bp, s.cur[0], s.cur[1]);
      s.cur[0] = s.cur[1] = 0;
    }
  }

  const u32 cnt = half ? s.cnt1 : s.cnt0;  // next free slot of this codeblock
  const bool overflow = !TRAIN && good && cnt > (u32)a.stride;
  if (overflow) status |= ST_OVERFLOW;
  if (!DEC && !TRAIN) {
    // flush of each lane's open codeword (BPC_PaCo.cu:1020)
    if (good) a.dst[g.region + s.slot] = (int32_t)s.L;
    // -1 in every word past the used slots (BPC_PaCo.cu:1284); a codeblock that was
    // not coded (all zero, or rejected) keeps only its MSB word
    if (valid && a.fill)
      for (u32 j = (good ? cnt : 1u) + hl; j < (u32)a.stride; j += 32) a.dst[g.region + j] = -1;
    if (valid && hl == 0 && a.ncw) a.ncw[cb] = good ? (int32_t)(cnt - 1) : 0;
  } else if (DEC) {
    // Write_Coefficients (BPC_PaCo.cu:46-63): magnitude from the decoded planes
    const bool emit = good && !overflow;
    if (valid) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r0 = 32 * h, r1 = min(rows, 32 * h + 32);
        if (r0 >= r1) break;
        uint2 pw[MAXP];  // planes 0..top of this half's rows; unwritten planes read as 0
#pragma unroll
        for (int p = 0; p < MAXP; ++p) pw[p] = p <= top ? scr.half_plane(h, p) : make_uint2(0u, 0u);
        const u32 g0 = (u32)(s.G[0] >> (32 * h)), g1 = (u32)(s.G[1] >> (32 * h));
        for (int r = r0; r < r1; ++r) {
          const int sh = r - r0;
          u32 m0 = 0, m1 = 0;
#pragma unroll
          for (int p = 0; p < MAXP; ++p) {
            m0 |= ((pw[p].x >> sh) & 1u) << p;
            m1 |= ((pw[p].y >> sh) & 1u) << p;
          }
          int v0 = (int)m0, v1 = (int)m1;
          if ((g0 >> sh) & 1u) v0 = -v0;
          if ((g1 >> sh) & 1u) v1 = -v1;
          if (!emit) v0 = v1 = 0;
          if (a.in16) {
            // int16 output: out-of-range values saturate and are reported
            const int s0 = max(-32768, min(32767, v0)), s1 = max(-32768, min(32767, v1));
            if (s0 != v0 || s1 != v1) status |= ST_NARROW;
            *reinterpret_cast<gptr<u32>>(reinterpret_cast<gptr<int16_t>>(a.dst) + cbase + (int64_t)r * rs) =
                ((u32)s0 & 0xFFFFu) | ((u32)s1 << 16);
          } else {
            *reinterpret_cast<gptr<u64>>(a.dst + cbase + (int64_t)r * rs) = (u64)(u32)v0 | ((u64)(u32)v1 << 32);
          }
        }
      }
    }
  }
  if (status && status_word) atomicOr(status_word, status);
}


// RP = 1: bitplanes in registers (no scratch slot; a.scratch unused).  CMP = 1: the
// decoder reads the compact container (a.cmsb, a.coff, a.cpay) instead of regions.
template <int DEC, int CP3, int MAXP, int RP, int CMP = 0>
__global__ __launch_bounds__(64) void bpc_codec_kernel(const KArgs a) {
  __shared__ uint8_t slut[256];
  const lds_u8 sl = build_sign_lut(slut);
  const int wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
  void* scr = RP ? nullptr : reinterpret_cast<char*>(a.scratch) + (size_t)wave * (2 * MAXP * 64 * 8);
  HalfDesc d;
  d.src = (cgi32)a.src;
  d.dst = (gi32)a.dst;
  d.lut = (cgi32)a.lut;
  d.ncw = (gi32)a.ncw;
  d.size_x = a.size_x;
  d.size_y = a.size_y;
  d.nx = a.nx;
  d.stride = a.stride;
  d.nbp = a.nbp;
  d.levels = a.levels;
  d.lut_len = a.lut_len;
  d.layout = a.layout;
  d.fill = a.fill;
  d.in16 = a.in16;
  d.cmsb = a.cmsb;
  d.coff = a.coff;
  d.cpay = a.cpay;
  d.ccap = a.ccap;
  d.slut = sl;
  d.prows = nullptr;  // the branch-free form keeps its probabilities in registers
  const int half = (threadIdx.x & 63) >> 5;
  const int nunits = a.single ? a.ncb : (a.ncb + 1) / 2;
  for (int unit = wave; unit < nunits; unit += nwaves) {
    const int cb = a.single ? (half ? a.ncb : unit) : 2 * unit + half;
    d.cb = cb < a.ncb ? cb : -1;
    // every wave of this launch may be the critical path: branch-free coder steps
    code_pair<DEC, CP3, MAXP, 0, 1, RP, CMP>(d, a.rows, a.status, scr);
  }
}

// ---- batched path ---------------------------------------------------------
// Codeblocks are sorted by MSB inside each image and paired there (both halves of
// a wave always code the same image, so its descriptor is wave-uniform); pairs
// are then ordered by (pass mode, MSB) with the costliest first.
// Counts the calling wave out of its launch.  When the launch ends the run
// (a.fold), the grid's last wave folds the run's status: an atomic OR of its bits
// into the caller's device word and a store of them into the host-visible word,
// with ST_WIDE_PENDING when deferred pairs wait for a 28-plane pass that the run
// left out.  No trailing kernel: under concurrent runs a small dispatch after the
// main pass waits for a CU behind other runs' persistent kernels.
__device__ __forceinline__ void batch_wave_done(const BatchArgs& a, bool wide_pass) {
  if (!a.fold) return;
  const int lane = threadIdx.x & 63;
  u32 done = 0;
  if (lane == 0) {
    __threadfence();  // the wave's status and deferral atomics come before its count
    done = atomicAdd(&a.ctl[wide_pass ? CTL_DONE_WIDE : CTL_DONE], 1u);
  }
  done = (u32)__builtin_amdgcn_readfirstlane((int)done);
  if (done != gridDim.x * (blockDim.x >> 6) - 1u || lane != 0) return;
  __threadfence();
  const u32 bits = atomicOr(&a.ctl[CTL_STATUS], 0u);
  const BatchRun* run = a.run;
  if (bits && run->status) atomicOr(run->status, bits);
  if (run->host_status) {
    const u32 pending = a.report_wide && atomicOr(&a.ctl[CTL_WIDE_COUNT], 0u) ? ST_WIDE_PENDING : 0u;
    *reinterpret_cast<volatile u32*>(run->host_status) = bits | pending;
    __threadfence_system();
  }
}

// Persistent kernel: every wave takes the next pair off the queue until it is drained.
// OCC = 0 leaves the register budget to the compiler (up to 8 waves per SIMD);
// OCC = 5..8 caps it so that that many waves fit on a SIMD, at the price of
// spills (the launch tuner compares them).
template <int DEC, int MAXP, int OCC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC ? OCC : 1, OCC ? OCC : 8)))
void bpc_batch_kernel(const BatchArgs a) {
  // The main pass drains the pair queue and defers the pairs whose MSB is beyond
  // its planes; a 28-plane pass then drains the deferred pairs.  Sizes come from
  // the run words (one captured sequence serves batches of any shape).
  const bool wide_pass = a.wide_pass != 0;
  const u32 n_items =
      (u32)__builtin_amdgcn_readfirstlane((int)(wide_pass ? a.ctl[CTL_WIDE_COUNT] : (u32)a.run->total_pairs));
  if (n_items == 0) {  // uniform over the workgroup: before its first barrier
    batch_wave_done(a, wide_pass);
    return;
  }
  const int rows = __builtin_amdgcn_readfirstlane(a.run->rows);
  __shared__ uint8_t slut[256];
  __shared__ uint32_t prows[4 * 64 * PROW_BYTES / 4];  // per wave of the workgroup (at most 4)
  const lds_u8 sl = build_sign_lut(slut);
  const lds_w8 pr = (lds_w8)prows + 64 * PROW_BYTES * (threadIdx.x >> 6);
  const int half = (threadIdx.x & 63) >> 5;
  const size_t wave = (size_t)__builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  void* scr = reinterpret_cast<char*>(a.scratch) + wave * (2 * MAXP * 64 * 8);
  const u32* queue = wide_pass ? a.wide : a.pair_queue;
  u32* head = &a.ctl[wide_pass ? CTL_WIDE_HEAD : CTL_HEAD];
  for (;;) {
    u32 slot = 0;
    if ((threadIdx.x & 63) == 0) slot = atomicAdd(head, 1u);
    slot = __builtin_amdgcn_readfirstlane(slot);
    if (slot >= n_items) break;
    const u32 pair = __builtin_amdgcn_readfirstlane(queue[slot]);
    if (MAXP < 28 && !wide_pass) {
      // the pair's costlier codeblock comes first: its MSB bucket is the pair's key
      const u32 b = (u32)__builtin_amdgcn_readfirstlane((int)a.pair_keys[pair]) % (u32)NBUCKET;
      if (b < (u32)NBUCKET - 1u && 27 - (int)b >= MAXP) {
        if ((threadIdx.x & 63) == 0) a.wide[atomicAdd(&a.ctl[CTL_WIDE_COUNT], 1u)] = pair;
        continue;
      }
    }
    const u32 first = __builtin_amdgcn_readfirstlane(a.sorted[2 * (size_t)pair]);
    const u32 e = a.sorted[2 * (size_t)pair + half];
    const BatchImage& im = a.images[queue_entry_image(first)];
    HalfDesc d;
    d.src = (cgi32)im.src;
    d.dst = (gi32)im.dst;
    d.lut = (cgi32)im.lut;
    d.ncw = (gi32)im.ncw;
    d.size_x = im.size_x;
    d.size_y = im.size_y;
    d.nx = im.nx;
    d.stride = im.stride;
    d.nbp = im.nbp;
    d.levels = im.levels;
    d.lut_len = im.lut_len;
    d.layout = im.layout;
    d.fill = im.fill;
    d.in16 = im.in16;
    d.slut = sl;
    d.prows = pr;
    d.cb = e == QUEUE_EMPTY ? -1 : (int)queue_entry_cb(e);
    if (im.cpass3)
      code_pair<DEC, 1, MAXP>(d, rows, &a.ctl[CTL_STATUS], scr);
    else
      code_pair<DEC, 0, MAXP>(d, rows, &a.ctl[CTL_STATUS], scr);
  }
  batch_wave_done(a, wide_pass);
}

// ---- LUT trainer ------------------------------------------------------------
// Symbol statistics of a training image: the encoder's modelling in symbol-count
// mode, every plane in registers (no codeblock is too wide).  The waves of a
// workgroup share one LDS histogram over the LUT entries; each workgroup writes
// its histogram to its own partial row, and lut_from_counts_kernel sums the rows.
template <int CP3>
__global__ __launch_bounds__(64 * TRAIN_WAVES) void bpc_train_kernel(const TrainArgs a) {
  extern __shared__ unsigned long long hist[];
  for (int k = threadIdx.x; k < a.lut_len; k += blockDim.x) hist[k] = 0;
  __shared__ uint8_t slut[256];
  const lds_u8 sl = build_sign_lut(slut);  // its barrier also covers the histogram clear
  const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * TRAIN_WAVES + (threadIdx.x >> 6)));
  const int nwaves = (int)gridDim.x * TRAIN_WAVES;
  void* scr = reinterpret_cast<char*>(a.scratch) + (size_t)wave * scratch_bytes_per_wave(TRAIN_MAXP);
  HalfDesc d;
  d.src = (cgi32)a.img;
  d.dst = nullptr;
  d.lut = nullptr;
  d.ncw = nullptr;
  d.size_x = a.size_x;
  d.size_y = a.size_y;
  d.nx = a.nx;
  d.stride = 1;
  d.nbp = a.nbp;
  d.levels = a.levels;
  d.lut_len = a.lut_len;
  d.layout = a.layout;
  d.fill = 0;
  d.in16 = 0;
  d.slut = sl;
  d.prows = nullptr;  // symbol counting reads no probabilities
  const int npairs = (a.ncb + 1) / 2;
  for (int pair = wave; pair < npairs; pair += nwaves) {
    const int cb = 2 * pair + ((threadIdx.x & 63) >> 5);
    d.cb = cb < a.ncb ? cb : -1;
    code_pair<0, CP3, TRAIN_MAXP, 1>(d, a.rows, nullptr, scr, hist);
  }
  __syncthreads();
  unsigned long long* row = a.partial + (size_t)blockIdx.x * a.lut_len;
  for (int k = threadIdx.x; k < a.lut_len; k += blockDim.x) row[k] = hist[k];
}

// Sums the partial rows per LUT entry; P = clamp(round(128 * zeros / symbols), 1, 127),
// 64 for an entry that codes nothing.  Optional outputs: the summed counts.
__global__ __launch_bounds__(256) void lut_from_counts_kernel(const unsigned long long* partial, int rows_n,
                                                               int lut_len, int32_t* lut, uint32_t* n0_out,
                                                               uint32_t* n_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= lut_len) return;
  unsigned long long t = 0;
  for (int b = 0; b < rows_n; ++b) t += partial[(size_t)b * lut_len + k];
  const unsigned long long n = t >> 32, n0 = t & 0xFFFFFFFFull;
  if (lut) {
    int v = 64;
    if (n) v = min(127, max(1, (int)((n0 * 128 + n / 2) / n)));
    lut[k] = v;
  }
  if (n0_out) n0_out[k] = (uint32_t)n0;
  if (n_out) n_out[k] = (uint32_t)n;
}
}  // namespace

int train_resident_blocks_per_cu(int cpass3, int lds_bytes) {
  int blocks = 0;
  if (hipFuncSetAttribute(cpass3 ? (const void*)bpc_train_kernel<1> : (const void*)bpc_train_kernel<0>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes) != hipSuccess)
    return 0;
  hipError_t e = cpass3 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, bpc_train_kernel<1>, 64 * TRAIN_WAVES,
                                                                       lds_bytes)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, bpc_train_kernel<0>, 64 * TRAIN_WAVES,
                                                                       lds_bytes);
  return e == hipSuccess ? blocks : 0;
}

hipError_t launch_train(int cpass3, const TrainArgs& a, int grid_blocks, int32_t* lut, uint32_t* n0, uint32_t* n,
                        hipStream_t stream) {
  const size_t lds = (size_t)a.lut_len * sizeof(unsigned long long);
  if (cpass3) {
    hipError_t e = hipFuncSetAttribute((const void*)bpc_train_kernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bpc_train_kernel<1>, dim3(grid_blocks), dim3(64 * TRAIN_WAVES), lds, stream, a);
  } else {
    hipError_t e = hipFuncSetAttribute((const void*)bpc_train_kernel<0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bpc_train_kernel<0>, dim3(grid_blocks), dim3(64 * TRAIN_WAVES), lds, stream, a);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(lut_from_counts_kernel, dim3((a.lut_len + 255) / 256), dim3(256), 0, stream, a.partial,
                     grid_blocks, a.lut_len, lut, n0, n);
  return hipGetLastError();
}

// Device-to-device copy, 16 bytes per lane per iteration (grid-stride).
__global__ __launch_bounds__(256) void copy_kernel(const int4* __restrict__ src, int4* __restrict__ dst, size_t n16) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += (size_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream) {
  if (bytes % 16 || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) return hipErrorInvalidValue;
  const size_t n16 = bytes / 16;
  if (!n16) return hipSuccess;
  const size_t blocks = std::min<size_t>((n16 + 255) / 256, 256 * 64);
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, (const int4*)src, (int4*)dst, n16);
  return hipGetLastError();
}

// ---- fixed bitplane scratch of the entry point that allocates nothing ---------
// A file-scope device array: every device that loads this code object holds its own
// copy, at its own address, until the process ends.  In host code the array's name
// is only a stand-in for the symbol; its device address comes from
// hipGetSymbolAddress, asked once per device with that device current.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

hipError_t noalloc_scratch(void** ptr) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* a = nullptr;
    e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = a;
  }
  *ptr = addr[dev];
  return hipSuccess;
}

int planes_for(int nbp) {
  if (nbp <= 16) return 16;
  if (nbp <= 20) return 20;
  return 28;
}

// The compact-container variants exist for the decoder only.
template <int DEC, int CP3, int MAXP, int RP>
static hipError_t launch_one(const KArgs& a, int grid_waves, hipStream_t stream) {
  if (DEC && a.cmsb)
    hipLaunchKernelGGL((bpc_codec_kernel<DEC, CP3, MAXP, RP, DEC>), dim3(grid_waves), dim3(64), 0, stream, a);
  else
    hipLaunchKernelGGL((bpc_codec_kernel<DEC, CP3, MAXP, RP>), dim3(grid_waves), dim3(64), 0, stream, a);
  return hipGetLastError();
}

template <int DEC, int CP3, int MAXP, int RP>
static int plain_resident(int compact) {
  int blocks = 0;
  const hipError_t e =
      DEC && compact
          ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, bpc_codec_kernel<DEC, CP3, MAXP, RP, DEC>, 64, 0)
          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, bpc_codec_kernel<DEC, CP3, MAXP, RP>, 64, 0);
  return e == hipSuccess ? blocks : 0;
}

// Register planes exist for 16 and 20 planes (28 would not stay in registers).
template <int DEC, int CP3, int RP>
static hipError_t launch_p(int maxp, const KArgs& a, int grid_waves, hipStream_t stream) {
  switch (maxp) {
    case 16: return launch_one<DEC, CP3, 16, RP>(a, grid_waves, stream);
    case 20: return launch_one<DEC, CP3, 20, RP>(a, grid_waves, stream);
    case 28: return RP ? hipErrorInvalidValue : launch_one<DEC, CP3, 28, 0>(a, grid_waves, stream);
    default: return hipErrorInvalidValue;
  }
}

template <int DEC, int CP3, int RP>
static int resident_p(int maxp, int compact) {
  switch (maxp) {
    case 16: return plain_resident<DEC, CP3, 16, RP>(compact);
    case 20: return plain_resident<DEC, CP3, 20, RP>(compact);
    case 28: return RP ? 0 : plain_resident<DEC, CP3, 28, 0>(compact);
    default: return 0;
  }
}

int codec_resident_waves_per_cu(int decode, int cpass3, int maxp, int reg_planes, int compact) {
  if (reg_planes) {
    if (decode) return cpass3 ? resident_p<1, 1, 1>(maxp, compact) : resident_p<1, 0, 1>(maxp, compact);
    return cpass3 ? resident_p<0, 1, 1>(maxp, 0) : resident_p<0, 0, 1>(maxp, 0);
  }
  if (decode) return cpass3 ? resident_p<1, 1, 0>(maxp, compact) : resident_p<1, 0, 0>(maxp, compact);
  return cpass3 ? resident_p<0, 1, 0>(maxp, 0) : resident_p<0, 0, 0>(maxp, 0);
}

template <int DEC, int MAXP, int OCC>
static hipError_t launch_batch_one(const BatchArgs& a, int grid_blocks, int waves_per_block, hipStream_t stream) {
  hipLaunchKernelGGL((bpc_batch_kernel<DEC, MAXP, OCC>), dim3(grid_blocks), dim3(64 * waves_per_block), 0, stream, a);
  return hipGetLastError();
}

template <int DEC, int MAXP, int OCC>
static int resident_one(int waves_per_block) {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, bpc_batch_kernel<DEC, MAXP, OCC>, 64 * waves_per_block,
                                                   0) != hipSuccess)
    return 0;
  return blocks;
}

// Register-capped variants exist for the 16- and 20-plane kernels.
static int occ_of(int maxp, int vgpr_cap_waves) {
  return (vgpr_cap_waves >= 5 && vgpr_cap_waves <= 8 && maxp <= 20) ? vgpr_cap_waves : 0;
}

int batch_resident_blocks_per_cu(int decode, int maxp, int waves_per_block, int vgpr_cap_waves) {
  const int occ = occ_of(maxp, vgpr_cap_waves);
  switch ((maxp * 2 + decode) * 16 + occ) {
    case 32 * 16: return resident_one<0, 16, 0>(waves_per_block);
    case 33 * 16: return resident_one<1, 16, 0>(waves_per_block);
    case 40 * 16: return resident_one<0, 20, 0>(waves_per_block);
    case 41 * 16: return resident_one<1, 20, 0>(waves_per_block);
    case 56 * 16: return resident_one<0, 28, 0>(waves_per_block);
    case 57 * 16: return resident_one<1, 28, 0>(waves_per_block);
    case 32 * 16 + 5: return resident_one<0, 16, 5>(waves_per_block);
    case 33 * 16 + 5: return resident_one<1, 16, 5>(waves_per_block);
    case 40 * 16 + 5: return resident_one<0, 20, 5>(waves_per_block);
    case 41 * 16 + 5: return resident_one<1, 20, 5>(waves_per_block);
    case 32 * 16 + 6: return resident_one<0, 16, 6>(waves_per_block);
    case 33 * 16 + 6: return resident_one<1, 16, 6>(waves_per_block);
    case 40 * 16 + 6: return resident_one<0, 20, 6>(waves_per_block);
    case 41 * 16 + 6: return resident_one<1, 20, 6>(waves_per_block);
    case 32 * 16 + 7: return resident_one<0, 16, 7>(waves_per_block);
    case 33 * 16 + 7: return resident_one<1, 16, 7>(waves_per_block);
    case 40 * 16 + 7: return resident_one<0, 20, 7>(waves_per_block);
    case 41 * 16 + 7: return resident_one<1, 20, 7>(waves_per_block);
    case 32 * 16 + 8: return resident_one<0, 16, 8>(waves_per_block);
    case 33 * 16 + 8: return resident_one<1, 16, 8>(waves_per_block);
    case 40 * 16 + 8: return resident_one<0, 20, 8>(waves_per_block);
    case 41 * 16 + 8: return resident_one<1, 20, 8>(waves_per_block);
    default: return 0;
  }
}

hipError_t launch_batch(int decode, int maxp, const BatchArgs& a, int grid_blocks, int waves_per_block,
                        int vgpr_cap_waves, hipStream_t stream) {
  if (waves_per_block < 1 || waves_per_block > 4 || grid_blocks < 1) return hipErrorInvalidValue;
  const int occ = occ_of(maxp, vgpr_cap_waves);
  switch ((maxp * 2 + decode) * 16 + occ) {
    case 32 * 16: return launch_batch_one<0, 16, 0>(a, grid_blocks, waves_per_block, stream);
    case 33 * 16: return launch_batch_one<1, 16, 0>(a, grid_blocks, waves_per_block, stream);
    case 40 * 16: return launch_batch_one<0, 20, 0>(a, grid_blocks, waves_per_block, stream);
    case 41 * 16: return launch_batch_one<1, 20, 0>(a, grid_blocks, waves_per_block, stream);
    case 56 * 16: return launch_batch_one<0, 28, 0>(a, grid_blocks, waves_per_block, stream);
    case 57 * 16: return launch_batch_one<1, 28, 0>(a, grid_blocks, waves_per_block, stream);
    case 32 * 16 + 5: return launch_batch_one<0, 16, 5>(a, grid_blocks, waves_per_block, stream);
    case 33 * 16 + 5: return launch_batch_one<1, 16, 5>(a, grid_blocks, waves_per_block, stream);
    case 40 * 16 + 5: return launch_batch_one<0, 20, 5>(a, grid_blocks, waves_per_block, stream);
    case 41 * 16 + 5: return launch_batch_one<1, 20, 5>(a, grid_blocks, waves_per_block, stream);
    case 32 * 16 + 6: return launch_batch_one<0, 16, 6>(a, grid_blocks, waves_per_block, stream);
    case 33 * 16 + 6: return launch_batch_one<1, 16, 6>(a, grid_blocks, waves_per_block, stream);
    case 40 * 16 + 6: return launch_batch_one<0, 20, 6>(a, grid_blocks, waves_per_block, stream);
    case 41 * 16 + 6: return launch_batch_one<1, 20, 6>(a, grid_blocks, waves_per_block, stream);
    case 32 * 16 + 7: return launch_batch_one<0, 16, 7>(a, grid_blocks, waves_per_block, stream);
    case 33 * 16 + 7: return launch_batch_one<1, 16, 7>(a, grid_blocks, waves_per_block, stream);
    case 40 * 16 + 7: return launch_batch_one<0, 20, 7>(a, grid_blocks, waves_per_block, stream);
    case 41 * 16 + 7: return launch_batch_one<1, 20, 7>(a, grid_blocks, waves_per_block, stream);
    case 32 * 16 + 8: return launch_batch_one<0, 16, 8>(a, grid_blocks, waves_per_block, stream);
    case 33 * 16 + 8: return launch_batch_one<1, 16, 8>(a, grid_blocks, waves_per_block, stream);
    case 40 * 16 + 8: return launch_batch_one<0, 20, 8>(a, grid_blocks, waves_per_block, stream);
    case 41 * 16 + 8: return launch_batch_one<1, 20, 8>(a, grid_blocks, waves_per_block, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_codec(int decode, int cpass3, int maxp, int reg_planes, const KArgs& a, int grid_waves,
                        hipStream_t stream) {
  if (a.ncb <= 0) return hipSuccess;
  if (reg_planes) {
    if (decode)
      return cpass3 ? launch_p<1, 1, 1>(maxp, a, grid_waves, stream) : launch_p<1, 0, 1>(maxp, a, grid_waves, stream);
    return cpass3 ? launch_p<0, 1, 1>(maxp, a, grid_waves, stream) : launch_p<0, 0, 1>(maxp, a, grid_waves, stream);
  }
  if (decode)
    return cpass3 ? launch_p<1, 1, 0>(maxp, a, grid_waves, stream) : launch_p<1, 0, 0>(maxp, a, grid_waves, stream);
  return cpass3 ? launch_p<0, 1, 0>(maxp, a, grid_waves, stream) : launch_p<0, 0, 0>(maxp, a, grid_waves, stream);
}

}  // namespace bpc
