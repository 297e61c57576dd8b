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

// rand()/RAND_MAX + 1, a value in [1, 2].
inline void FillRandomVector(Vector &v) {
  for (local_int_t i = 0; i < v.localLength; ++i)
    v.values[i] = rand() / (double)(RAND_MAX) + 1.0;
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
