// Problem file for building the reference CPU code as a checker.
//
// This is a user problem file in the reference's own plugin interface
// (Init / InitDomain / Analysis / UserDefBoundary, template at
// Src/Templates/init.cpp:17-191).  It does not replace any reference
// source: it only feeds a seeded initial condition, written by
// gpluto_dev_amd.ic, into the reference through InitDomain()
// (called at Src/startup.cpp:344, after Init() and before the staggered
// field average at Src/startup.cpp:388).
//
// IC file "gpl_ic.bin" in the working directory, little endian:
//   char[8]  "GPLIC01"            magic (NUL terminated)
//   int32    nx, ny, nz           active zones per direction
//   int32    nvar                 cell-centred primitive variables
//   int32    nstag                staggered face components (0 for HD)
//   int32    periodic             1: ghosts wrap, 0: ghosts clamp
//   double   gamma                EOS gamma (run_config.gamma)
//   double   Vc[nvar][nz][ny][nx]
//   double   Vs[nstag][nz][ny][nx]  face d of cell (i,j,k) sits at the
//                                  lower face (i-1/2 for d=0, ...)
// Vs is stored periodically (one face per cell); the reference's extra
// upper face (index iend+1) receives the value of the lower face ibeg,
// exactly what PeriodicBoundary does (Src/boundary.cpp:518-533).

#include "gpluto.hpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct ICFile {
  int nx = 0, ny = 0, nz = 0, nvar = 0, nstag = 0, periodic = 1;
  double gamma = 5.0/3.0;
  std::vector<double> vc, vs;
  bool loaded = false;
};

ICFile &ic()
{
  static ICFile f;
  if (f.loaded) return f;
  FILE *fp = fopen("gpl_ic.bin", "rb");
  if (!fp) { fprintf(stderr, "init_ic: cannot open gpl_ic.bin\n"); exit(2); }
  char magic[8];
  int hdr[6];
  if (fread(magic, 1, 8, fp) != 8 || strcmp(magic, "GPLIC01") != 0) {
    fprintf(stderr, "init_ic: bad magic\n"); exit(2);
  }
  if (fread(hdr, sizeof(int), 6, fp) != 6 || fread(&f.gamma, sizeof(double), 1, fp) != 1) {
    fprintf(stderr, "init_ic: short header\n"); exit(2);
  }
  f.nx = hdr[0]; f.ny = hdr[1]; f.nz = hdr[2];
  f.nvar = hdr[3]; f.nstag = hdr[4]; f.periodic = hdr[5];
  size_t ncell = (size_t)f.nx*f.ny*f.nz;
  f.vc.resize(ncell*f.nvar);
  f.vs.resize(ncell*f.nstag);
  if (fread(f.vc.data(), sizeof(double), f.vc.size(), fp) != f.vc.size() ||
      fread(f.vs.data(), sizeof(double), f.vs.size(), fp) != f.vs.size()) {
    fprintf(stderr, "init_ic: short data\n"); exit(2);
  }
  fclose(fp);
  f.loaded = true;
  return f;
}

int wrap(int a, int n, int periodic)
{
  if (periodic) { a %= n; return a < 0 ? a + n : a; }
  return a < 0 ? 0 : (a >= n ? n - 1 : a);
}

} // namespace

void Init (double *v, double x1, double x2, double x3, RunConfig &run_config)
{
  run_config.gamma = ic().gamma;
  for (int nv = 0; nv < NVAR; nv++) v[nv] = 0.0;
  v[RHO] = 1.0;
  #if HAVE_ENERGY
  v[PRS] = 1.0;
  #endif
  #if PHYSICS == MHD
  v[AX1] = v[AX2] = v[AX3] = 0.0;
  #endif
}

void InitDomain (Data *data, Grid *grid)
{
  ICFile &f = ic();
  const int ib = grid->lbeg[IDIR], jb = grid->lbeg[JDIR], kb = grid->lbeg[KDIR];
  const int n1 = grid->ntot[IDIR], n2 = grid->ntot[JDIR], n3 = grid->ntot[KDIR];
  if (grid->nint[IDIR] != f.nx || grid->nint[JDIR] != f.ny || grid->nint[KDIR] != f.nz) {
    fprintf(stderr, "init_ic: grid %d %d %d != IC %d %d %d\n", grid->nint[IDIR],
            grid->nint[JDIR], grid->nint[KDIR], f.nx, f.ny, f.nz);
    exit(2);
  }
  if (f.nvar != NVAR) { fprintf(stderr, "init_ic: nvar mismatch\n"); exit(2); }
  auto cell = [&](int nv, int a, int b, int c) {
    return f.vc[(((size_t)nv*f.nz + c)*f.ny + b)*f.nx + a];
  };
  for (int k = 0; k < n3; k++)
  for (int j = 0; j < n2; j++)
  for (int i = 0; i < n1; i++) {
    int a = wrap(i - ib, f.nx, f.periodic);
    int b = wrap(j - jb, f.ny, f.periodic);
    int c = wrap(k - kb, f.nz, f.periodic);
    for (int nv = 0; nv < NVAR; nv++) data->Vc(i,j,k,nv) = cell(nv, a, b, c);
  }
  #ifdef STAGGERED_MHD
  auto face = [&](int d, int a, int b, int c) {
    return f.vs[(((size_t)d*f.nz + c)*f.ny + b)*f.nx + a];
  };
  for (int d = 0; d < f.nstag && d < DIMENSIONS; d++) {
    Ary3D &Q = data->Vs[d];
    int m1 = n1 + (d == 0), m2 = n2 + (d == 1 && INCLUDE_JDIR), m3 = n3 + (d == 2 && INCLUDE_KDIR);
    for (int k = 0; k < m3; k++)
    for (int j = 0; j < m2; j++)
    for (int i = 0; i < m1; i++) {
      int a = wrap(i - ib, f.nx, 1);
      int b = wrap(j - jb, f.ny, 1);
      int c = wrap(k - kb, f.nz, 1);
      Q(i,j,k) = face(d, a, b, c);
    }
  }
  #endif
}

void Analysis (Data *data, Grid *grid) {}

#if PHYSICS == MHD
void BackgroundField (State &state, RunConfig &run_config)
{
  state.Bgf[IDIR] = 0.0;
  state.Bgf[JDIR] = 0.0;
  state.Bgf[KDIR] = 0.0;
}
#endif

void UserDefBoundary (const Data *data, RBox *box, int side, Grid *grid) {}
