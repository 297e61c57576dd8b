// Vector.hpp -- the HPCG Vector (the reference's Vector.hpp:27-113: same
// struct, same helpers, same semantics).
//
// Host `values` are the authoritative copy between calls.  The optimized
// Compute* functions copy their inputs to the device problem's scratch
// vectors and their outputs back (CG keeps its whole solve on the device and
// copies only b and x), so code that reads or writes `values` directly --
// the reference's tests, ComputeResidual, user code -- always sees current
// data.  `optimizationData` is left unused.
#ifndef HPCG_AMD_VECTOR_HPP
#define HPCG_AMD_VECTOR_HPP

#include <cstdint>
#include <cassert>
#include <cstdlib>

#include "Geometry.hpp"

struct Vector_STRUCT {
  local_int_t localLength;  // length of the local part
  double *values;           // host values
  void *optimizationData;
};
typedef struct Vector_STRUCT Vector;

inline void InitializeVector(Vector &v, local_int_t localLength) {
  v.localLength = localLength;
  v.values = new double[localLength];
  v.optimizationData = 0;
}

inline void ZeroVector(Vector &v) {
  for (local_int_t i = 0; i < v.localLength; ++i) v.values[i] = 0.0;
}

inline void ScaleVectorValue(Vector &v, local_int_t index, double value) {
  assert(index >= 0 && index < v.localLength);
  v.values[index] *= value;
}

namespace hpcg_amd {
// The values glibc's rand() returns from its default seed (srand(1)), from a
// private state: other threads of the process (the GPU runtime's among
// them) may call rand() at times that vary from run to run, which would
// shift the shared sequence.  glibc's generator: r[0] = seed,
// r[i] = 16807 r[i-1] mod (2^31 - 1) for i < 31, r[i] = r[i-31] for
// i = 31..33, then r[i] = r[i-31] + r[i-3] mod 2^32; the outputs are
// r[i] >> 1 from i = 344 on.
inline int reference_rand() {
  static uint32_t s[34];
  static int k = -1;  // index of the next r[i], modulo 34
  auto step = [&]() {
    const uint32_t v = s[(k + 34 - 31) % 34] + s[(k + 34 - 3) % 34];
    s[k] = v;
    k = (k + 1) % 34;
    return v;
  };
  if (k < 0) {
    int64_t r = 1;
    s[0] = 1;
    for (int i = 1; i < 31; ++i) s[i] = (uint32_t)(r = (16807 * r) % 2147483647);
    for (int i = 31; i < 34; ++i) s[i] = s[i - 31];
    k = 0;  // r[34] goes to slot 0
    for (int i = 34; i < 344; ++i) step();
  }
  return (int)(step() >> 1);
}
}  // namespace hpcg_amd

// rand()/RAND_MAX + 1, a value in [1, 2] (the default-seed rand() sequence,
// hpcg_amd::reference_rand).
inline void FillRandomVector(Vector &v) {
  for (local_int_t i = 0; i < v.localLength; ++i)
    v.values[i] = hpcg_amd::reference_rand() / (double)(RAND_MAX) + 1.0;
}

inline void CopyVector(const Vector &v, Vector &w) {
  assert(w.localLength >= v.localLength);
  for (local_int_t i = 0; i < v.localLength; ++i) w.values[i] = v.values[i];
}

inline void DeleteVector(Vector &v) {
  delete[] v.values;
  v.values = 0;
  v.localLength = 0;
}

#endif
