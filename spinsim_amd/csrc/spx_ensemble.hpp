// Host-side disorder draws shared by the ensemble entry points and the
// surrogate's bulk prediction.
#pragma once

#include <cstddef>
#include <cstdint>

namespace spx {

struct DipolarSpec {
  uint64_t seed;
  size_t nspins;
  double scale;
  double min_separation;
  bool angular;
  double disorder;
};

int check_spec(const DipolarSpec& s);
// One realisation: pairs (lo < hi, in generation order) with the secular
// coupling jz (jx = jy = -jz / 2), then the Sz disorder hz.
int draw_dipolar(const DipolarSpec& s, uint64_t index, uint32_t* lo, uint32_t* hi, double* jz, double* hz);
// The same draws followed by the realisation's typical initial state
// (psi: 2 * 2^L doubles, interleaved, normalised); SPX_EDOMAIN for a zero
// state.
int draw_realisation(const DipolarSpec& s, uint64_t index, uint32_t* lo, uint32_t* hi, double* jz, double* hz,
                     double* psi);
// Realisations first .. first + count - 1 as dense symmetric J [count][L][L]
// and fields h [count][L], in FP64 and/or FP32 (null outputs are skipped).
// threads <= 0 picks the pool size.
int draw_graphs(const DipolarSpec& s, uint64_t first, size_t count, double* J, double* h, float* Jf, float* hf,
                int threads);

}  // namespace spx
