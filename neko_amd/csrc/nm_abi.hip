// Drop-in entry points with the signatures Neko's Fortran already binds
// through iso_c_binding for its HIP backend (device pointers by value,
// integers by reference, 1-based gs maps, FP64 "real").  Linking this
// library in place of Neko's HIP kernel objects swaps the kernels without
// touching the Fortran side.  Like Neko's HIP_CHECK these entry points have
// no error return: a HIP failure aborts with a message.
//
// Not re-entrant: the reduction scratch buffers are shared per process, as
// Neko's own redbuf (math.hip) is.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "neko_amd.h"
#include "nm_common.h"
#include "nm_kernels.h"

using namespace nm;

// Neko's global command queue (device.F90:52, bind(c)).  Weak, so the
// Fortran definition wins when this library is linked into Neko; used as the
// stream of the entry points that take none (hip_ax_helm, hip_jacobi_update,
// hip_gmres_part2), exactly as in the reference.
extern "C" __attribute__((weak)) void *glb_cmd_queue = nullptr;

namespace {

struct RedBuf {
  double *part = nullptr;   // RED_BLOCKS * maxj
  double *out = nullptr;    // maxj
  double *host = nullptr;   // pinned, maxj
  const double **vptr = nullptr;  // device array of pointers
  int cap = 0;
};
RedBuf g_red;

void red_reserve(int nv) {
  if (nv <= g_red.cap) return;
  if (g_red.part) {
    NM_HIP_CHECK(hipFree(g_red.part));
    NM_HIP_CHECK(hipFree(g_red.out));
    NM_HIP_CHECK(hipHostFree(g_red.host));
    NM_HIP_CHECK(hipFree(g_red.vptr));
  }
  int cap = nv < 32 ? 32 : nv;
  NM_HIP_CHECK(hipMalloc(&g_red.part, sizeof(double) * RED_BLOCKS * cap));
  NM_HIP_CHECK(hipMalloc(&g_red.out, sizeof(double) * cap));
  NM_HIP_CHECK(hipHostMalloc(&g_red.host, sizeof(double) * cap, hipHostMallocDefault));
  NM_HIP_CHECK(hipMalloc(&g_red.vptr, sizeof(double *) * cap));
  g_red.cap = cap;
}

inline hipStream_t glb() { return (hipStream_t)glb_cmd_queue; }

void check(int st) {
  if (st != NM_OK) {
    std::fprintf(stderr, "neko_amd: %s\n", last_error());
    std::abort();
  }
}

}  // namespace

extern "C" {

// ax_helm.hip:64-67 -> w = sum D^T (h1 G) D u, FP64, on glb_cmd_queue.
// dy/dz/dxt/dyt/dzt are accepted for signature compatibility; the space is
// isotropic (lx = ly = lz, space.f90:163-170), so D = dx throughout.
void hip_ax_helm(void *w, void *u, void *dx, void *dy, void *dz, void *dxt,
                 void *dyt, void *dzt, void *h1, void *g11, void *g22,
                 void *g33, void *g12, void *g13, void *g23, int *nelv,
                 int *lx) {
  (void)dy; (void)dz; (void)dxt; (void)dyt; (void)dzt;
  AxArgs<double> a;
  a.w = (double *)w;
  a.u = (const double *)u;
  a.dx = (const double *)dx;
  a.h1 = (const double *)h1;
  a.h1c = 1.0;
  a.g[0] = (const double *)g11; a.g[1] = (const double *)g22;
  a.g[2] = (const double *)g33; a.g[3] = (const double *)g12;
  a.g[4] = (const double *)g13; a.g[5] = (const double *)g23;
  a.h2b = nullptr;
  a.nel = *nelv;
  a.lx = *lx;
  a.skip = nullptr;
  check(launch_ax<double>(a, 0, glb()));
}

// FP32 twin of hip_ax_helm (Neko's sp build passes float fields).
void hip_ax_helm_sp(void *w, void *u, void *dx, void *h1, void *g11,
                    void *g22, void *g33, void *g12, void *g13, void *g23,
                    int *nelv, int *lx, hipStream_t strm) {
  AxArgs<float> a;
  a.w = (float *)w;
  a.u = (const float *)u;
  a.dx = (const float *)dx;
  a.h1 = (const float *)h1;
  a.h1c = 1.0f;
  a.g[0] = (const float *)g11; a.g[1] = (const float *)g22;
  a.g[2] = (const float *)g33; a.g[3] = (const float *)g12;
  a.g[4] = (const float *)g13; a.g[5] = (const float *)g23;
  a.h2b = nullptr;
  a.nel = *nelv;
  a.lx = *lx;
  a.skip = nullptr;
  check(launch_ax<float>(a, 0, strm));
}

// gs.hip:52-93 (gather) and :98-118 (scatter), Neko map layout.
void hip_gather_kernel(void *v, int *m, int *o, void *dg, void *u, int *n,
                       void *gd, int *nb, void *b, void *bo, int *op,
                       hipStream_t stream) {
  (void)n;
  check(launch_neko_gather<double>((double *)v, *m, *o, (const int *)dg,
                                   (const double *)u, (const int *)gd, *nb,
                                   (const int *)b, (const int *)bo, *op,
                                   stream));
}

void hip_scatter_kernel(void *v, int *m, void *dg, void *u, int *n, void *gd,
                        int *nb, void *b, void *bo, hipStream_t stream) {
  (void)n;
  check(launch_neko_scatter<double>((const double *)v, *m, (const int *)dg,
                                    (double *)u, (const int *)gd, *nb,
                                    (const int *)b, (const int *)bo, stream));
}

// math.hip:891 glsc3: sum a*b*c, host scalar (one synchronisation).
double hip_glsc3(void *a, void *b, void *c, int *n, hipStream_t stream) {
  red_reserve(1);
  check(launch_glsc3_partial((const double *)a, (const double *)b,
                             (const double *)c, *n, g_red.part, stream));
  check(launch_combine(g_red.part, 1, g_red.out, stream));
  NM_HIP_CHECK(hipMemcpyAsync(g_red.host, g_red.out, sizeof(double),
                              hipMemcpyDeviceToHost, stream));
  NM_HIP_CHECK(hipStreamSynchronize(stream));
  return g_red.host[0];
}

// math.hip:922 glsc3_many: h(l) = sum w * v_l * mult, l = 1..j
void hip_glsc3_many(double *h, void *w, void *v, void *mult, int *j, int *n,
                    hipStream_t stream) {
  red_reserve(*j);
  NM_HIP_CHECK(hipMemcpyAsync(g_red.vptr, v, sizeof(double *) * (*j),
                              hipMemcpyDeviceToDevice, stream));
  check(launch_glsc3_many_partial((const double *)w, g_red.vptr,
                                  (const double *)mult, *j, *n, g_red.part,
                                  stream));
  check(launch_combine(g_red.part, *j, g_red.out, stream));
  NM_HIP_CHECK(hipMemcpyAsync(g_red.host, g_red.out, sizeof(double) * (*j),
                              hipMemcpyDeviceToHost, stream));
  NM_HIP_CHECK(hipStreamSynchronize(stream));
  for (int l = 0; l < *j; l++) h[l] = g_red.host[l];
}

// BLAS-1 updates of math.hip (real scalars by reference, as bound there).
#define NM_VEC1(name, OP)                                                   \
  void name(void *a, void *b, double *c1, int *n, hipStream_t strm) {       \
    check(launch_vec<double>(OP, (double *)a, (const double *)b, nullptr,   \
                             nullptr, *c1, *n, strm));                      \
  }
NM_VEC1(hip_add2s1, V_ADD2S1)
NM_VEC1(hip_add2s2, V_ADD2S2)
NM_VEC1(hip_cmult2, V_CMULT2)
#undef NM_VEC1

void hip_copy(void *a, void *b, int *n, hipStream_t strm) {
  check(launch_vec<double>(V_COPY, (double *)a, (const double *)b, nullptr, nullptr, 0.0, *n, strm));
}
void hip_rzero(void *a, int *n, hipStream_t strm) {
  check(launch_vec<double>(V_RZERO, (double *)a, nullptr, nullptr, nullptr, 0.0, *n, strm));
}
void hip_cmult(void *a, double *c, int *n, hipStream_t strm) {
  check(launch_vec<double>(V_CMULT, (double *)a, nullptr, nullptr, nullptr, *c, *n, strm));
}
void hip_cfill(void *a, double *c, int *n, hipStream_t strm) {
  check(launch_vec<double>(V_CFILL, (double *)a, nullptr, nullptr, nullptr, *c, *n, strm));
}
void hip_add2(void *a, void *b, int *n, hipStream_t strm) {
  check(launch_vec<double>(V_ADD2, (double *)a, (const double *)b, nullptr, nullptr, 0.0, *n, strm));
}
void hip_col2(void *a, void *b, int *n, hipStream_t strm) {
  check(launch_vec<double>(V_COL2, (double *)a, (const double *)b, nullptr, nullptr, 0.0, *n, strm));
}
void hip_col3(void *a, void *b, void *c, int *n, hipStream_t strm) {
  check(launch_vec<double>(V_COL3, (double *)a, (const double *)b, (const double *)c, nullptr, 0.0, *n, strm));
}
void hip_sub2(void *a, void *b, int *n, hipStream_t strm) {
  check(launch_vec<double>(V_SUB2, (double *)a, (const double *)b, nullptr, nullptr, 0.0, *n, strm));
}
void hip_invcol1(void *a, int *n, hipStream_t strm) {
  check(launch_vec<double>(V_INVCOL1, (double *)a, nullptr, nullptr, nullptr, 0.0, *n, strm));
}
void hip_addcol3(void *a, void *b, void *c, int *n, hipStream_t strm) {
  check(launch_vec<double>(V_ADDCOL3, (double *)a, (const double *)b, (const double *)c, nullptr, 0.0, *n, strm));
}
void hip_addcol4(void *a, void *b, void *c, void *d, int *n, hipStream_t strm) {
  check(launch_vec<double>(V_ADDCOL4, (double *)a, (const double *)b, (const double *)c, (const double *)d, 0.0, *n, strm));
}

// pc_jacobi.hip:108 signature and formula: the diagonal of the reference's
// device kernel (cross terms G12/G13/G23 at every face point, which makes
// it the exact diagonal of the element stiffness), on glb_cmd_queue.  The
// CPU formula (pc_jacobi.f90:168-264: cross terms only at the corners of
// deformed elements) is nm_jacobi_update.
void hip_jacobi_update(void *d, void *dxt, void *dyt, void *dzt, void *G11,
                       void *G22, void *G33, void *G12, void *G13, void *G23,
                       int *nel, int *lxp) {
  (void)dyt; (void)dzt;
  const double *g[6] = {(const double *)G11, (const double *)G22, (const double *)G33,
                        (const double *)G12, (const double *)G13, (const double *)G23};
  check(launch_jacobi_diag(*nel, *lxp, (const double *)dxt, g, nullptr, (double *)d, glb()));
}

// Same with Neko's per-element deformation flags (dfrmd_el, device int32).
void nm_jacobi_update(void *d, void *dxt, void *G11, void *G22, void *G33,
                      void *G12, void *G13, void *G23, void *dfrmd, int *nel,
                      int *lxp, hipStream_t strm) {
  const double *g[6] = {(const double *)G11, (const double *)G22, (const double *)G33,
                        (const double *)G12, (const double *)G13, (const double *)G23};
  check(launch_jacobi_diag(*nel, *lxp, (const double *)dxt, g, (const int *)dfrmd,
                           (double *)d, strm));
}

// Ordered device copy helper for host harnesses (kind: hipMemcpyKind).
int nm_memcpy(void *dst, const void *src, size_t bytes, int kind,
              hipStream_t strm) {
  NM_HIP_TRY(hipMemcpyAsync(dst, src, bytes, (hipMemcpyKind)kind, strm));
  NM_HIP_TRY(hipStreamSynchronize(strm));
  return NM_OK;
}

}  // extern "C"
