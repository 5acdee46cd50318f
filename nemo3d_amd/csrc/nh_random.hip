// The reference's random start vectors, reproduced number for number on the
// host (no device work):
//
// * libc rand() as glibc implements it (the generator every Linux build of
//   the reference uses): the additive feedback generator of glibc's
//   random_r (TYPE_3, r[i] = r[i-3] + r[i-31] mod 2^32, output r[i] >> 1,
//   the table seeded by the Park-Miller step 16807 r mod (2^31 - 1) and the
//   first 310 outputs discarded; as described by P. Selinger, "The GLIBC
//   pseudo-random number generator").  Used by the Lanczos Random start
//   (lanczos_initialize, NEMO_3D/nemo3d/src/base/eigsys3d_par.c:125-131:
//   r0[j] = (rand() / RAND_MAX, 0)) and the block Lanczos start block
//   (NEMO_3D/eigen/src/lanczos_blk.c:2220-2223, the p vectors in turn).
// * LAPACK zlarnv(idist = 1): real and imaginary parts uniform on (0, 1)
//   from dlaruv, the multiplicative congruential generator with modulus
//   2^48 and multiplier 33952834046453 (Fishman 1990), number i of a call
//   being (a^i seed mod 2^48) / 2^48 and the seed advancing to the last
//   product; zlarnv draws 64 complex numbers per dlaruv call.  Used for the
//   TRACEMIN start block (NEMO_3D/nemo3d/src/base/qtracemin_driver.c:
//   105-116, iseed = {103 + rank, 101 rank, 7, 57}).
#include <cstdint>

#include "nh_internal.h"

namespace nh {

GlibcRand::GlibcRand(uint32_t seed) {
  // srand(0) behaves as srand(1)
  int32_t word = seed == 0 ? 1 : (int32_t)seed;
  r_[0] = (uint32_t)word;
  for (int i = 1; i < 31; ++i) {
    // 16807 word mod (2^31 - 1) without overflow (Schrage's method)
    const int32_t hi = word / 127773, lo = word % 127773;
    word = 16807 * lo - 2836 * hi;
    if (word < 0) word += 2147483647;
    r_[i] = (uint32_t)word;
  }
  for (int i = 31; i < 34; ++i) r_[i] = r_[i - 31];
  k_ = 34;
  for (int i = 34; i < 344; ++i) step();
}

uint32_t GlibcRand::step() {
  const uint32_t v = r_[(k_ - 31) % 34] + r_[(k_ - 3) % 34];
  r_[k_ % 34] = v;
  ++k_;
  if (k_ >= 34 * 1024) k_ -= 34 * 1000;  // keep the index small, same residues
  return v;
}

void dlaruv(int32_t iseed[4], int n, double *x) {
  constexpr uint64_t kMask = (1ull << 48) - 1;
  constexpr uint64_t kA = 33952834046453ull;
  constexpr double R = 1.0 / 4096.0;
  const uint64_t s = ((uint64_t)iseed[0] << 36) | ((uint64_t)iseed[1] << 24) |
                     ((uint64_t)iseed[2] << 12) | (uint64_t)iseed[3];
  uint64_t m = 1, p = s;
  for (int i = 0; i < n && i < 128; ++i) {
    m = (uint64_t)(((unsigned __int128)m * kA) & kMask);
    p = (uint64_t)(((unsigned __int128)m * s) & kMask);
    const int it1 = (int)(p >> 36), it2 = (int)((p >> 24) & 4095),
              it3 = (int)((p >> 12) & 4095), it4 = (int)(p & 4095);
    // the conversion of dlaruv, in its order of operations
    x[i] = R * ((double)it1 + R * ((double)it2 + R * ((double)it3 + R * (double)it4)));
  }
  iseed[0] = (int32_t)(p >> 36);
  iseed[1] = (int32_t)((p >> 24) & 4095);
  iseed[2] = (int32_t)((p >> 12) & 4095);
  iseed[3] = (int32_t)(p & 4095);
}

void zlarnv_uniform(int32_t iseed[4], int64_t n, nh_cplx *x) {
  double u[128];
  for (int64_t iv = 0; iv < n; iv += 64) {
    const int il = (int)(n - iv < 64 ? n - iv : 64);
    dlaruv(iseed, 2 * il, u);
    for (int i = 0; i < il; ++i) x[iv + i] = nh_cplx{u[2 * i], u[2 * i + 1]};
  }
}

void rand_start(uint32_t seed, int64_t n, nh_cplx *x) {
  GlibcRand g(seed);
  for (int64_t j = 0; j < n; ++j) x[j] = nh_cplx{1.0 * g.next() / 2147483647.0, 0.0};
}

}  // namespace nh

extern "C" nh_status nh_random_start(int kind, const int32_t *seed, int64_t n,
                                     nh_cplx *out) {
  NH_CHECK(seed && out && n >= 0, NH_ERR_ARG, "nh_random_start: bad arguments");
  if (kind == NH_RANDOM_RAND) {
    nh::rand_start((uint32_t)seed[0], n, out);
    return NH_OK;
  }
  NH_CHECK(kind == NH_RANDOM_ZLARNV, NH_ERR_ARG, "nh_random_start: kind %d", kind);
  int32_t is[4] = {seed[0], seed[1], seed[2], seed[3]};
  for (int k = 0; k < 4; ++k)
    NH_CHECK(is[k] >= 0 && is[k] < 4096, NH_ERR_ARG,
             "zlarnv: seed entries must lie in 0..4095");
  NH_CHECK(is[3] % 2 == 1, NH_ERR_ARG, "zlarnv: seed[3] must be odd");
  nh::zlarnv_uniform(is, n, out);
  return NH_OK;
}
