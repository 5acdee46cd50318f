// H·v in the SK-compressed and matrix-free layouts: per bond only the
// direction cosines and the 21 strain-scaled two-centre integrals are
// needed (24 values), stored once per bond (cation-owned) or recomputed
// from the atoms' displacements (MfSrc); the 10x10 block is rebuilt row by
// row in registers on every product.  The onsite record (diagonal with
// Boykin corrections, surface couplings, spin-orbit) is stored per atom as
// in the block layout.
//
// Mapping: 4 lanes per atom, lane n owns neighbour slot n.  A workgroup
// holds only cation-role or only anion-role atoms (lists `cations` /
// `anions`), so whether the lane applies T (anion row) or its transpose
// (cation row) is uniform across the wavefront.  Each lane forms its bond's
// 20 (10) complex products with R-precision arithmetic inside the block,
// adds them to FP64 accumulators, a 4-lane butterfly sums the four bonds,
// and lane n then finishes and writes rows r = n, n+4, ... with the onsite
// terms.
#pragma once

#include "nh_internal.h"
#include "nh_sk_row.h"

namespace nh {

constexpr int kSkThreads = 256;
constexpr int kSkAtomsPerBlock = kSkThreads / 4;

// Workgroup -> atoms.  Consecutive workgroups go to different XCDs (round
// robin over 8): workgroup b runs on XCD b % 8, and XCD x takes the blocks
// [x * ppx, (x + 1) * ppx) of the atom lists, so each XCD works on a
// compact region (the lists are ordered by bricks of cells) and reuses the
// neighbours' vector entries from its L2.  Cation block k and anion block k
// cover the same cells.  Two mappings:
//  * one role per workgroup (mixed = 0): 64 atoms, cation and anion blocks
//    alternating on each XCD, so the two roles of the same cells run back
//    to back; grid 16 * ppx, ppx = ceil(max(blocks of 64) / 8);
//  * both roles (mixed = 1): waves 0-1 hold 32 cations and waves 2-3 the 32
//    anions of the same list position, so both rows of a bond read its
//    record within one workgroup; grid 8 * ppx, ppx = ceil(max(blocks of
//    32) / 8).
// Either way the role is uniform across a wavefront, so whether a lane
// applies T (anion row) or its transpose (cation row) is too.
struct SkGrid {
  int ppx;
  int mixed;
};
struct SkLane {
  bool cat;
  int li;  // index into the cation or anion list
};
// LANES: lanes per atom (4: lane = bond; 8: lane = bond and spin, SkSplit)
template <int LANES = 4>
__device__ __forceinline__ SkLane sk_lane_of(SkGrid g, int b, int t) {
  constexpr int kAtoms = kSkThreads / LANES;
  const int x = b & 7, s = b >> 3;
  if (g.mixed) {
    constexpr int kHalf = kSkThreads / 2;
    return {t < kHalf, (x * g.ppx + s) * (kAtoms / 2) + (t & (kHalf - 1)) / LANES};
  }
  return {(s & 1) == 0, (x * g.ppx + (s >> 1)) * kAtoms + t / LANES};
}
// the grid of a launch and its mapping
inline SkGrid sk_grid(int n_cations, int n_anions, bool mixed, int *blocks,
                      int lanes = 4) {
  const int atoms = kSkThreads / lanes;
  const int per = mixed ? atoms / 2 : atoms;
  const int nbc = (n_cations + per - 1) / per;
  const int nba = (n_anions + per - 1) / per;
  const int ppx = ((nbc > nba ? nbc : nba) + 7) / 8;
  *blocks = (mixed ? 8 : 16) * ppx;
  return SkGrid{ppx, mixed ? 1 : 0};
}

// Complex entries of the FP64 vector, or of an FP32 one widened on load
__device__ __forceinline__ double2 widen2(double2 v) { return v; }
__device__ __forceinline__ double2 widen2(float2 v) { return make_double2(v.x, v.y); }

// Onsite contribution (diagonal - shift, surface s-p couplings, spin-orbit)
// to row (s, o) of atom with vector base pointer xa.
template <int NSP, typename XA = double2>
__device__ __forceinline__ void onsite_row(const double *__restrict__ os,
                                           const XA *__restrict__ xa,
                                           int s, int o, double shift,
                                           double &re, double &im) {
  const double2 xo = widen2(xa[s * 10 + o]);
  const double d = os[o] - shift;
  re += d * xo.x;
  im += d * xo.y;
  if (o >= 1 && o <= 4) {
    // partners and stored coupling index: (1,2)0 (1,3)1 (1,4)2 (2,3)3 (2,4)4 (3,4)5
    const int p0 = o == 1 ? 2 : 1, k0 = o == 1 ? 0 : o == 2 ? 0 : o == 3 ? 1 : 2;
    const int p1 = o <= 2 ? 3 : 2, k1 = o == 1 ? 1 : o == 2 ? 3 : o == 3 ? 3 : 4;
    const int p2 = o <= 3 ? 4 : 3, k2 = o == 1 ? 2 : o == 2 ? 4 : 5;
    const double h0 = os[10 + k0], h1 = os[10 + k1], h2 = os[10 + k2];
    const double2 a0 = widen2(xa[s * 10 + p0]), a1 = widen2(xa[s * 10 + p1]), a2 = widen2(xa[s * 10 + p2]);
    re += h0 * a0.x + h1 * a1.x + h2 * a2.x;
    im += h0 * a0.y + h1 * a1.y + h2 * a2.y;
  }
  if (NSP == 2 && o >= 2 && o <= 4) {
    const double lam = os[16];
    if (s == 0) {
      if (o == 2) {
        re += lam * widen2(xa[3]).y + lam * widen2(xa[14]).x;
        im += -lam * widen2(xa[3]).x + lam * widen2(xa[14]).y;
      } else if (o == 3) {
        re += -lam * widen2(xa[2]).y + lam * widen2(xa[14]).y;
        im += lam * widen2(xa[2]).x - lam * widen2(xa[14]).x;
      } else {
        re += -lam * widen2(xa[12]).x - lam * widen2(xa[13]).y;
        im += -lam * widen2(xa[12]).y + lam * widen2(xa[13]).x;
      }
    } else {
      if (o == 2) {
        re += -lam * widen2(xa[13]).y - lam * widen2(xa[4]).x;
        im += lam * widen2(xa[13]).x - lam * widen2(xa[4]).y;
      } else if (o == 3) {
        re += lam * widen2(xa[12]).y + lam * widen2(xa[4]).y;
        im += -lam * widen2(xa[12]).x - lam * widen2(xa[4]).x;
      } else {
        re += lam * widen2(xa[2]).x - lam * widen2(xa[3]).y;
        im += lam * widen2(xa[2]).y + lam * widen2(xa[3]).x;
      }
    }
  }
}

// Block product of one bond, rows applied by orientation: anion rows
// acc[s][a] += sum_c T[a][c] x[s][c]; cation rows acc[s][c] += sum_a
// T[a][c] x[s][a].
template <int NSP, typename R, bool CATION>
struct BondApply {
  template <int A>
  __device__ __forceinline__ static void run(R l, R m, R n, const R *v,
                                             const R (&xr)[NSP][10],
                                             const R (&xi)[NSP][10],
                                             R (&br)[NSP][10],
                                             R (&bi)[NSP][10]) {
    if constexpr (A < 10) {
      R row[10];
      sk_row<A, R>(l, m, n, v, row);
#pragma unroll
      for (int s = 0; s < NSP; ++s) {
        if constexpr (!CATION) {
          R sr = R(0), si = R(0);
#pragma unroll
          for (int c = 0; c < 10; ++c) {
            sr += row[c] * xr[s][c];
            si += row[c] * xi[s][c];
          }
          br[s][A] += sr;
          bi[s][A] += si;
        } else {
#pragma unroll
          for (int c = 0; c < 10; ++c) {
            br[s][c] += row[c] * xr[s][A];
            bi[s][c] += row[c] * xi[s][A];
          }
        }
      }
      run<A + 1>(l, m, n, v, xr, xi, br, bi);
    }
  }
};

// The 24 record values of a bond as 16-byte loads (12 for FP64, 6 for
// FP32): the records are 16-byte aligned (24 values per bond).  Across a
// wavefront every lane reads its own record, so the number of load
// instructions, each touching 64 cache lines, is what the texture
// addresser pays for.
__device__ __forceinline__ void sk_load_record(const double *rec, double (&v)[24]) {
  const double2 *p = reinterpret_cast<const double2 *>(rec);
#pragma unroll
  for (int q = 0; q < 12; ++q) {
    const double2 t = p[q];
    v[2 * q] = t.x;
    v[2 * q + 1] = t.y;
  }
}
__device__ __forceinline__ void sk_load_record(const float *rec, float (&v)[24]) {
  const float4 *p = reinterpret_cast<const float4 *>(rec);
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const float4 t = p[q];
    v[4 * q] = t.x;
    v[4 * q + 1] = t.y;
    v[4 * q + 2] = t.z;
    v[4 * q + 3] = t.w;
  }
}

// Where a lane's bond record comes from.
//
// SkSrc: the stored SK layout, 24 values per bond (each bond once), found
// through bond_of.
template <typename R>
struct SkSrc {
  const R *__restrict__ sk;
  const int32_t *__restrict__ bond_of;
  // FP64: the record is requested before the neighbour gather, whose
  // barrier its loads then overlap (measured: plain solve -4%); FP32: after
  // it (requested early, the 24 values spill at 128 VGPRs: +3% per product)
  static constexpr bool kEarly = sizeof(R) == 8;
  template <bool CATION>
  __device__ __forceinline__ void fetch(int atom, int slot, int /*j*/,
                                        R (&rv)[24]) const {
    sk_load_record(sk + (size_t)bond_of[atom * 4 + slot] * kSkStride, rv);
  }
};

// MfSrc: the matrix-free layout (the reference's HamCompute = Recompute,
// Hmult_spds_col, NEMO_3D/nemo3d/src/base/h_cvectr_mult.c:1134-1886): the
// record is recomputed on every product from the two atoms' displacements
// and species and the slot's ideal bond vector, as the assembly computes it
// (strain_corr_prep, mb_ham_spds_micro.c:2848-2978): direction cosines of
// the strained bond and the 21 integrals scaled by (d0/d)^eta, with
// (d0/d)^eta = exp(eta log(d0/d)) (one log and 15 exp per bond).  The bond
// vector is formed in the cation's frame from either end, so both rows of
// a bond get bit-identical records and H stays exactly Hermitian.  With
// R = float the integrals are scaled in FP32 (geometry and the log in FP64).
template <typename R>
struct MfSrc {
  const double4 *__restrict__ geo;  // [A] displacement x, y, z (nm), species
  const int32_t *__restrict__ q;    // [A][4] packed int8 ideal bond vector,
                                    // cation frame
  const double *__restrict__ pair;  // [S][S][kMfPair]
  int S;
  double lx, ly, lz;
  static constexpr bool kEarly = true;

  __device__ __forceinline__ static void scale(const double *__restrict__ P,
                                               double lb, R *__restrict__ v) {
    R f[15];
#pragma unroll
    for (int e = 0; e < 15; ++e) {
      if constexpr (sizeof(R) == 8) f[e] = exp(P[kMfEta + e] * lb);
      else f[e] = __expf((float)(P[kMfEta + e] * lb));
    }
#pragma unroll
    for (int k = 0; k < 21; ++k) v[k] = R(P[k]) * f[eta_index(k)];
  }

  template <bool CATION>
  __device__ __forceinline__ void fetch(int atom, int slot, int j,
                                        R (&rv)[24]) const {
    const double4 gi = geo[atom], gj = geo[j];
    const int qp = q[atom * 4 + slot];
    const double4 gc = CATION ? gi : gj, ga = CATION ? gj : gi;
    const double qx = (double)(int8_t)(qp & 0xff);
    const double qy = (double)(int8_t)((qp >> 8) & 0xff);
    const double qz = (double)(int8_t)((qp >> 16) & 0xff);
    // anion -> cation (getRelativePosVec, geo_struct.c:2278-2294, with the
    // cation-row sign of h_cvectr_mult.c:1073-1090)
    const double nx = -(lx * (0.25 * qx) + ga.x - gc.x);
    const double ny = -(ly * (0.25 * qy) + ga.y - gc.y);
    const double nz = -(lz * (0.25 * qz) + ga.z - gc.z);
    const double den = sqrt(nx * nx + ny * ny + nz * nz);
    rv[0] = R(nx / den);
    rv[1] = R(ny / den);
    rv[2] = R(nz / den);
    const int p = (int)gc.w * S + (int)ga.w;
    // the common case, one species pair across the wavefront, reads the
    // pair constants through the scalar cache
    const int pu = __builtin_amdgcn_readfirstlane(p);
    if (__all(p == pu)) {
      const double *P = pair + pu * kMfPair;
      scale(P, log((0.25 * kSqrt3Ref * P[kMfA0]) / den), rv + 3);
    } else {
      const double *P = pair + p * kMfPair;
      scale(P, log((0.25 * kSqrt3Ref * P[kMfA0]) / den), rv + 3);
    }
  }
};

template <typename R>
inline MfSrc<R> mf_src(const nh_context *c) {
  return MfSrc<R>{c->mf_geo, c->mf_q, c->mf_pair, c->n_species,
                  c->lattice[0], c->lattice[1], c->lattice[2]};
}

// The neighbour's NB complex entries, from the FP64 vector (one double2
// per entry) or from its FP32 copy (one float4 per two entries; the entries
// of an atom start on a 16-byte boundary, NB * 8 bytes per atom).
template <int NSP, typename R>
__device__ __forceinline__ void sk_load_x(const double2 *__restrict__ x,
                                          int64_t j, R (&xr)[NSP][10],
                                          R (&xi)[NSP][10]) {
  const double2 *xj = x + j * (NSP * 10);
#pragma unroll
  for (int s = 0; s < NSP; ++s)
#pragma unroll
    for (int c = 0; c < 10; ++c) {
      const double2 t = xj[s * 10 + c];
      xr[s][c] = R(t.x);
      xi[s][c] = R(t.y);
    }
}
template <int NSP, typename R>
__device__ __forceinline__ void sk_load_x(const float2 *__restrict__ x,
                                          int64_t j, R (&xr)[NSP][10],
                                          R (&xi)[NSP][10]) {
  const float4 *xj = reinterpret_cast<const float4 *>(x + j * (NSP * 10));
#pragma unroll
  for (int s = 0; s < NSP; ++s)
#pragma unroll
    for (int c = 0; c < 10; c += 2) {
      const float4 t = xj[(s * 10 + c) / 2];
      xr[s][c] = R(t.x);
      xi[s][c] = R(t.y);
      xr[s][c + 1] = R(t.z);
      xi[s][c + 1] = R(t.w);
    }
}

// Quad-cooperative gather of the neighbours' entries through LDS: the four
// lanes of an atom load the entries of each of its four neighbours in turn,
// lane `slot` taking the 16-byte chunks slot, slot + 4, ..., so one load
// instruction touches one 64-byte piece per quad (16 cache lines per
// wavefront instead of 64 when every lane reads its own neighbour); every
// lane then reads its own neighbour's entries back from LDS.  Entries are
// those of XT (double2: the FP64 vector; float2: its FP32 copy).
template <int NSP, typename XT>
struct SkQuadStage {
  static constexpr int kChunks = NSP * 10 * (int)sizeof(XT) / 16;  // 20/10/5
  // Lane t keeps its item in chunks [t kChunks, (t + 1) kChunks), no
  // padding.  A multiple of 4 chunks: chunk index XOR-swizzled by
  // (t >> 2) & 3 (16 lanes then cover all 64 banks); 2 mod 4: chunk index
  // rotated by one for lanes with bit 2 ^ bit 3 set, so the reads of 8 or
  // 16 consecutive lanes hit distinct banks (the quad writes stay 2-way);
  // odd: already conflict-free.
  static constexpr bool kSwizzle = kChunks % 4 == 0;
  static constexpr bool kRotate = kChunks % 4 == 2;
  static constexpr int kStride = kChunks;
  __device__ static int pos(int t, int c) {
    if constexpr (kSwizzle) return t * kStride + (c ^ ((t >> 2) & 3));
    if constexpr (kRotate) {
      const int r = c + (((t >> 2) ^ (t >> 3)) & 1);
      return t * kStride + (r >= kChunks ? r - kChunks : r);
    }
    return t * kStride + c;
  }
  uint4 buf[kSkThreads * kStride];
};

// FP32 Lanczos vectors: r_k = u_{k-1} - a r_{k-1}, a = alpha_{k-1} /
// beta_{k-1} (so that a r_{k-1} = alpha_{k-1} q_{k-1}), in FP32 with a
// split into two FP32 parts (a = hi + lo to 2^-48 relative): two fmas,
// rounding entry by entry.  The coefficient must not be rounded to FP32
// alone: its error is the same for every entry, i.e. an error along
// q_{k-1} in every step, which cost a 786,520-atom dot 10,251 steps
// instead of 7,901 (and a level, from another start); with the split it
// takes the same 7,826 steps as one FP64 fma per entry, 6% faster in the
// product.  Every kernel that forms r_k (the products that read u_{k-1}
// and r_{k-1}, the eigenvector pass that writes r_k) evaluates these
// expressions on the same operands, so they agree bit for bit.
struct LzCoef {
  float hi, lo;
};
__device__ __forceinline__ LzCoef lz_defer_coef(double al, double inv) {
  const double a = al * inv;
  const float hi = (float)a;
  return LzCoef{hi, (float)(a - (double)hi)};
}
__device__ __forceinline__ float lz_defer1(float u, float rp, LzCoef a) {
  // the cancellation u - hi rp happens inside the first fma (one rounding
  // of the exact difference), then the small correction: about one ulp of
  // r in all (the other order adds half an ulp of u, which can be larger)
  return __builtin_fmaf(-a.lo, rp, __builtin_fmaf(-a.hi, rp, u));
}
// where the products read r_k from: x holds u_{k-1}, rp r_{k-1}
struct LzDeferSrc {
  const float2 *rp;
  LzCoef a;  // lz_defer_coef(alpha_{k-1}, 1 / beta_{k-1})
};
__device__ __forceinline__ uint4 lz_defer_chunk(uint4 u, uint4 rp, const LzDeferSrc &d) {
  uint4 o;
  o.x = __float_as_uint(lz_defer1(__uint_as_float(u.x), __uint_as_float(rp.x), d.a));
  o.y = __float_as_uint(lz_defer1(__uint_as_float(u.y), __uint_as_float(rp.y), d.a));
  o.z = __float_as_uint(lz_defer1(__uint_as_float(u.z), __uint_as_float(rp.z), d.a));
  o.w = __float_as_uint(lz_defer1(__uint_as_float(u.w), __uint_as_float(rp.w), d.a));
  return o;
}

// The NSP * 10 entries of an item start at entry `off` of the neighbour's
// `stride` entries (stride = NSP * 10, off = 0: the whole atom; stride =
// 20, off = 10 s: spin block s of a 20-band atom).  DEFER (XT = float2):
// the staged entries are r_k formed from x = u_{k-1} and d.rp = r_{k-1}.
template <int NSP, typename R, typename XT, bool DEFER = false>
__device__ __forceinline__ void sk_quad_gather(int j, int slot,
                                               const XT *__restrict__ x,
                                               SkQuadStage<NSP, XT> &L,
                                               R (&xr)[NSP][10],
                                               R (&xi)[NSP][10],
                                               int stride = NSP * 10, int off = 0,
                                               LzDeferSrc d = {nullptr, {0.0f, 0.0f}}) {
  static_assert(!DEFER || sizeof(XT) == 8, "deferred r_k needs FP32 vectors");
  using S = SkQuadStage<NSP, XT>;
  constexpr int NB = NSP * 10;
  constexpr int C = S::kChunks;
  const int t = threadIdx.x, q0 = t & ~3;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int jn = __shfl(j, (t & 63 & ~3) | n, 64);
    if (jn >= 0) {
      const uint4 *src = reinterpret_cast<const uint4 *>(x + (int64_t)jn * stride + off);
      if constexpr (DEFER) {
        const uint4 *srp =
            reinterpret_cast<const uint4 *>(d.rp + (int64_t)jn * stride + off);
#pragma unroll
        for (int c0 = 0; c0 < C; c0 += 4) {
          const int c = c0 + slot;
          if (c < C) L.buf[S::pos(q0 + n, c)] = lz_defer_chunk(src[c], srp[c], d);
        }
      } else {
#pragma unroll
        for (int c0 = 0; c0 < C; c0 += 4) {
          const int c = c0 + slot;
          if (c < C) L.buf[S::pos(q0 + n, c)] = src[c];
        }
      }
    }
  }
  __syncthreads();
  if (j >= 0) {
    if constexpr (sizeof(XT) == 16) {
#pragma unroll
      for (int e = 0; e < NB; ++e) {
        const double2 v = *reinterpret_cast<const double2 *>(&L.buf[S::pos(t, e)]);
        xr[e / 10][e % 10] = R(v.x);
        xi[e / 10][e % 10] = R(v.y);
      }
    } else {
#pragma unroll
      for (int e = 0; e < NB; e += 2) {
        const float4 v = *reinterpret_cast<const float4 *>(&L.buf[S::pos(t, e / 2)]);
        xr[e / 10][e % 10] = R(v.x);
        xi[e / 10][e % 10] = R(v.y);
        xr[(e + 1) / 10][(e + 1) % 10] = R(v.z);
        xi[(e + 1) / 10][(e + 1) % 10] = R(v.w);
      }
    }
  }
}

// Epilogue staging (after the quad gather, in the same LDS): every atom's
// own entries and onsite record, loaded by its quad with lane `slot` taking
// rows slot, slot + 4, ... (64 contiguous bytes per quad and load), so that
// the onsite couplings read their partner entries from LDS instead of
// issuing scattered loads per row.
template <int NSP, int ATOMS = kSkAtomsPerBlock>
struct SkOwnStage {
  double2 x[ATOMS][NSP * 10];
  double2 os[ATOMS][kOnsiteStride / 2];
};

// Own entries and onsite record loaded into registers (before the barrier
// that frees the staging LDS, so the loads overlap it) ...
// (LANES lanes per atom; lane `lane` takes entries lane, lane + LANES, ...)
template <int NSP, int LANES = 4>
struct SkOwnRegs {
  static constexpr int kOs = (kOnsiteStride / 2 + LANES - 1) / LANES;  // onsite chunks per lane
  double2 x[(NSP * 10 + LANES - 1) / LANES];
  double2 os[kOs];
};
// (xa: the FP64 vector, or the FP32 one whose entries are widened on load)
template <int NSP, int LANES = 4, typename XO = double2>
__device__ __forceinline__ void sk_load_own(int lane, const XO *__restrict__ xa,
                                            const double *__restrict__ os,
                                            SkOwnRegs<NSP, LANES> &g) {
  constexpr int NB = NSP * 10;
#pragma unroll
  for (int i = 0; i < (NB + LANES - 1) / LANES; ++i) {
    const int r = LANES * i + lane;
    g.x[i] = r < NB ? widen2(xa[r]) : make_double2(0.0, 0.0);
  }
  const double2 *o2 = reinterpret_cast<const double2 *>(os);  // 144-byte records
#pragma unroll
  for (int i = 0; i < SkOwnRegs<NSP, LANES>::kOs; ++i) {
    const int c = LANES * i + lane;
    g.os[i] = c < kOnsiteStride / 2 ? o2[c] : make_double2(0.0, 0.0);
  }
}
// ... and stored to the epilogue staging after it
template <int NSP, int LANES = 4, int ATOMS = kSkAtomsPerBlock>
__device__ __forceinline__ void sk_store_own(int la, int lane,
                                             const SkOwnRegs<NSP, LANES> &g,
                                             SkOwnStage<NSP, ATOMS> &S) {
  constexpr int NB = NSP * 10;
#pragma unroll
  for (int i = 0; i < (NB + LANES - 1) / LANES; ++i)
    if (LANES * i + lane < NB) S.x[la][LANES * i + lane] = g.x[i];
#pragma unroll
  for (int i = 0; i < SkOwnRegs<NSP, LANES>::kOs; ++i)
    if (LANES * i + lane < kOnsiteStride / 2) S.os[la][LANES * i + lane] = g.os[i];
}

// Exchange of a double with lane ^ 1 / lane ^ 2 inside the quad through
// DPP quad permutations (VALU moves; a __shfl_xor is an LDS permute).
template <int CTRL>
__device__ __forceinline__ double quad_perm_f64(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                          (unsigned)lo));
}
constexpr int kQuadXor1 = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int kQuadXor2 = 0x4E;  // quad_perm [2, 3, 0, 1]

// The 4 bonds of `atom` (lane = slot), summed over the quad in FP64 (fixed
// order); lane `slot` keeps rows r = 4 i + slot in own_re[i] / own_im[i].
// SRC: where the bond records come from (SkSrc, MfSrc).
// XT: double2 (the FP64 vector) or float2 (its FP32 copy, R = float).
// QUAD: gather through LDS (sk_quad_gather; every thread of the workgroup
// must call, there is a barrier) in the caller's LDS area L, or directly
// from the caches.  Callers branch on the role with a wave-uniform
// (scalar) condition, so every wave runs exactly one instantiation and
// reaches the same number of barriers whichever role it holds.
template <int NSP, typename R, bool CATION, typename XT, bool QUAD, bool DEFER = false,
          typename SRC>
__device__ __forceinline__ void sk_bonds(int atom, int slot, bool active,
                                         const SRC &src,
                                         const int32_t *__restrict__ nbr,
                                         const XT *__restrict__ xc,
                                         double (&own_re)[(NSP * 10 + 3) / 4],
                                         double (&own_im)[(NSP * 10 + 3) / 4],
                                         SkQuadStage<NSP, XT> *L = nullptr,
                                         LzDeferSrc d = {nullptr, {0.0f, 0.0f}}) {
  static_assert(!DEFER || QUAD, "deferred r_k needs the LDS gather");
  R br[NSP][10], bi[NSP][10];
#pragma unroll
  for (int s = 0; s < NSP; ++s)
#pragma unroll
    for (int c = 0; c < 10; ++c) br[s][c] = bi[s][c] = R(0);
  const int j = active ? nbr[atom * 4 + slot] : -1;
  R xr[NSP][10], xi[NSP][10];
  if constexpr (QUAD) {
    R rv[24];
    if (SRC::kEarly && j >= 0) src.template fetch<CATION>(atom, slot, j, rv);
    sk_quad_gather<NSP, R, XT, DEFER>(j, slot, xc, *L, xr, xi, NSP * 10, 0, d);
    if (!SRC::kEarly && j >= 0) src.template fetch<CATION>(atom, slot, j, rv);
    if (j >= 0)
      BondApply<NSP, R, CATION>::template run<0>(rv[0], rv[1], rv[2], rv + 3,
                                                  xr, xi, br, bi);
  } else if (j >= 0) {
    R rv[24];
    src.template fetch<CATION>(atom, slot, j, rv);
    sk_load_x<NSP, R>(xc, j, xr, xi);
    BondApply<NSP, R, CATION>::template run<0>(rv[0], rv[1], rv[2], rv + 3, xr,
                                                xi, br, bi);
  }
  // sum over the quad (lanes 4k..4k+3); the owner lane keeps the row
#pragma unroll
  for (int s = 0; s < NSP; ++s)
#pragma unroll
    for (int c = 0; c < 10; ++c) {
      const int r = s * 10 + c;
      double re = double(br[s][c]), im = double(bi[s][c]);
      re += quad_perm_f64<kQuadXor1>(re);
      im += quad_perm_f64<kQuadXor1>(im);
      re += quad_perm_f64<kQuadXor2>(re);
      im += quad_perm_f64<kQuadXor2>(im);
      if ((r & 3) == slot) {
        own_re[r >> 2] = re;
        own_im[r >> 2] = im;
      }
    }
}

// SkSplit: 8 lanes per atom of a 20-band model, lane = (spin s, bond n):
// lanes 4 s .. 4 s + 3 of an atom hold the four bonds for spin block s, so
// each lane applies its bond's 10 x 10 block to one spin block only (half
// the vector registers of sk_bonds), the record is read by both spin lanes
// of a bond, and the four-lane sums stay within the spin group.  Lane
// `slot` of spin group s keeps rows s * 10 + 4 i + slot (i < 3) in
// own_re[i] / own_im[i].  QUAD only (LDS gather of the spin block).
struct SkSplit {
  static constexpr int kLanes = 8;
  static constexpr int kAtoms = kSkThreads / kLanes;
  static constexpr int kOwn = 3;  // rows per lane
};

template <typename R, bool CATION, typename XT, bool DEFER = false, typename SRC>
__device__ __forceinline__ void sk_bonds_split(int atom, int slot, int spin,
                                               bool active, const SRC &src,
                                               const int32_t *__restrict__ nbr,
                                               const XT *__restrict__ xc,
                                               double (&own_re)[SkSplit::kOwn],
                                               double (&own_im)[SkSplit::kOwn],
                                               SkQuadStage<1, XT> &L,
                                               LzDeferSrc d = {nullptr, {0.0f, 0.0f}}) {
  R br[1][10], bi[1][10];
#pragma unroll
  for (int c = 0; c < 10; ++c) br[0][c] = bi[0][c] = R(0);
  const int j = active ? nbr[atom * 4 + slot] : -1;
  R xr[1][10], xi[1][10];
  R rv[24];
  if (SRC::kEarly && j >= 0) src.template fetch<CATION>(atom, slot, j, rv);
  sk_quad_gather<1, R, XT, DEFER>(j, slot, xc, L, xr, xi, 20, 10 * spin, d);
  if (!SRC::kEarly && j >= 0) src.template fetch<CATION>(atom, slot, j, rv);
  if (j >= 0)
    BondApply<1, R, CATION>::template run<0>(rv[0], rv[1], rv[2], rv + 3, xr,
                                              xi, br, bi);
#pragma unroll
  for (int c = 0; c < 10; ++c) {
    double re = double(br[0][c]), im = double(bi[0][c]);
    re += quad_perm_f64<kQuadXor1>(re);
    im += quad_perm_f64<kQuadXor1>(im);
    re += quad_perm_f64<kQuadXor2>(re);
    im += quad_perm_f64<kQuadXor2>(im);
    if ((c & 3) == slot) {
      own_re[c >> 2] = re;
      own_im[c >> 2] = im;
    }
  }
}

// The same with the bond record already in registers (rv; j < 0: no bond)
// and the quad LDS gather of XT entries: the block products of several
// vectors reuse one record.
template <int NSP, typename R, bool CATION, typename XT>
__device__ __forceinline__ void sk_bonds_rec(int j, int slot, const R (&rv)[24],
                                             const XT *__restrict__ xc,
                                             double (&own_re)[(NSP * 10 + 3) / 4],
                                             double (&own_im)[(NSP * 10 + 3) / 4],
                                             SkQuadStage<NSP, XT> &L) {
  R br[NSP][10], bi[NSP][10];
#pragma unroll
  for (int s = 0; s < NSP; ++s)
#pragma unroll
    for (int c = 0; c < 10; ++c) br[s][c] = bi[s][c] = R(0);
  R xr[NSP][10], xi[NSP][10];
  sk_quad_gather<NSP, R, XT>(j, slot, xc, L, xr, xi);
  if (j >= 0)
    BondApply<NSP, R, CATION>::template run<0>(rv[0], rv[1], rv[2], rv + 3, xr,
                                                xi, br, bi);
#pragma unroll
  for (int s = 0; s < NSP; ++s)
#pragma unroll
    for (int c = 0; c < 10; ++c) {
      const int r = s * 10 + c;
      double re = double(br[s][c]), im = double(bi[s][c]);
      re += quad_perm_f64<kQuadXor1>(re);
      im += quad_perm_f64<kQuadXor1>(im);
      re += quad_perm_f64<kQuadXor2>(re);
      im += quad_perm_f64<kQuadXor2>(im);
      if ((r & 3) == slot) {
        own_re[r >> 2] = re;
        own_im[r >> 2] = im;
      }
    }
}

}  // namespace nh
