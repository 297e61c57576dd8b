// Geometry.hpp -- problem geometry of the HPCG driver (single rank).
//
// Same type names and fields as the reference's Geometry.hpp:29-75, so that
// code written against the reference's headers compiles against these.  With
// one rank the process grid is 1x1x1 and global == local sizes.
#ifndef HPCG_AMD_GEOMETRY_HPP
#define HPCG_AMD_GEOMETRY_HPP

typedef int local_int_t;
typedef long long global_int_t;

struct Geometry_STRUCT {
  int size;        // ranks (1)
  int rank;        // this rank (0)
  int numThreads;  // host threads
  local_int_t nx, ny, nz;  // local grid
  int npx, npy, npz;       // process grid (1, 1, 1)
  int pz;
  int npartz;
  int *partz_ids;
  local_int_t *partz_nz;
  int ipx, ipy, ipz;
  global_int_t gnx, gny, gnz;     // global grid (== local)
  global_int_t gix0, giy0, giz0;  // this rank's offset (0)
};
typedef struct Geometry_STRUCT Geometry;

// GenerateGeometry of the reference (GenerateGeometry.cpp:49-169) for one rank.
void GenerateGeometry(int size, int rank, int numThreads, int pz,
                      local_int_t zl, local_int_t zu, local_int_t nx,
                      local_int_t ny, local_int_t nz, int npx, int npy,
                      int npz, Geometry *geom);

inline void DeleteGeometry(Geometry &geom) {
  delete[] geom.partz_nz;
  delete[] geom.partz_ids;
  geom.partz_nz = 0;
  geom.partz_ids = 0;
}

#endif
