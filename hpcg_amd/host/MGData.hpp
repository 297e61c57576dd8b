// MGData.hpp -- coarse-grid transfer data of one level (the reference's
// MGData.hpp:28-75: same fields and helpers).
#ifndef HPCG_AMD_MGDATA_HPP
#define HPCG_AMD_MGDATA_HPP

#include "Vector.hpp"

struct MGData_STRUCT {
  int numberOfPresmootherSteps;   // SYMGS calls before coarsening
  int numberOfPostsmootherSteps;  // SYMGS calls after coarsening
  local_int_t *f2cOperator;       // coarse row -> injected fine row
  Vector *rc;                     // coarse residual
  Vector *xc;                     // coarse correction
  Vector *Axf;                    // fine A*x
  void *optimizationData;         // unused: the device problem keeps its own
};
typedef struct MGData_STRUCT MGData;

inline void InitializeMGData(local_int_t *f2cOperator, Vector *rc, Vector *xc,
                             Vector *Axf, MGData &data) {
  data.numberOfPresmootherSteps = 1;
  data.numberOfPostsmootherSteps = 1;
  data.f2cOperator = f2cOperator;
  data.rc = rc;
  data.xc = xc;
  data.Axf = Axf;
  data.optimizationData = 0;
}

inline void DeleteMGData(MGData &data) {
  delete[] data.f2cOperator;
  DeleteVector(*data.Axf);
  DeleteVector(*data.rc);
  DeleteVector(*data.xc);
  delete data.Axf;
  delete data.rc;
  delete data.xc;
}

#endif
