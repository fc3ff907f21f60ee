// FP64 C++/OpenMP restatement of QGN's GetRHS and the ARK43 / ARK43_QGLeith
// time step, with the reference's FFT library (MKL DFTI, loaded at run time
// from libmkl_rt).
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product links or loads this
// library: tests/ check it against the numpy oracle (oracle/qgn_oracle.py),
// and bench.py times it as the CPU baseline on the host cores.
//
// It follows the reference statement by statement, with its arrays in the
// reference's memory layout (C order (nz, ny, nx/2+1) complex = Fortran
// complex(nx/2+1, ny, nz)):
//   GetPsi          QGN_Module.f90:315-341
//   Jacobian        QGN_Module.f90:215-249 (flux form)
//   GetRHS          QGN_Module.f90:258-296
//   DeAlias         QGN_Module.f90:194-204
//   ARK43*          ARK43.f90 / ARK43_QGLeith.f90 (stages, error norm over all
//                   layers, reject factor 0.75 / 0.5, PI.3.4 with 0.75 / 0.2 TOL)
//   DFTI set-up     MKL_DFTI_Setup.f90:11-50: batched 2-D r2c / c2r, CCE half
//                   spectrum along x, out of place, unnormalised both ways.
// Loops over layers and rows run on OpenMP threads; MKL threads its
// transforms with the same (GNU) OpenMP runtime.
#include <dlfcn.h>
#include <omp.h>

#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

using cd = std::complex<double>;
typedef void* DftiHandle;
typedef long MklLong;

// DFTI configuration parameters and values of MKL's C interface
enum {
  kForwardScale = 4, kNumberOfTransforms = 7, kConjugateEvenStorage = 10, kPlacement = 11, kInputStrides = 12,
  kOutputStrides = 13, kInputDistance = 14, kOutputDistance = 15, kPackedFormat = 21
};
enum { kReal = 33, kComplexComplex = 39, kNotInplace = 44, kCce = 57 };

struct Mkl {
  void* so = nullptr;
  MklLong (*create_d_md)(DftiHandle*, int, MklLong, MklLong*) = nullptr;
  MklLong (*set_value)(DftiHandle, int, ...) = nullptr;
  MklLong (*commit)(DftiHandle) = nullptr;
  MklLong (*forward)(DftiHandle, void*, ...) = nullptr;
  MklLong (*backward)(DftiHandle, void*, ...) = nullptr;
  MklLong (*free_desc)(DftiHandle*) = nullptr;
};

Mkl g_mkl;
std::string g_err;

bool load_mkl() {
  if (g_mkl.so) return true;
  // MKL threads through the GNU OpenMP runtime this library uses
  setenv("MKL_THREADING_LAYER", "GNU", 0);
  const char* names[] = {"libmkl_rt.so.1", "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1"};
  for (const char* n : names) {
    g_mkl.so = dlopen(n, RTLD_NOW | RTLD_GLOBAL);
    if (g_mkl.so) break;
  }
  if (!g_mkl.so) {
    g_err = "libmkl_rt not found";
    return false;
  }
  g_mkl.create_d_md = reinterpret_cast<decltype(g_mkl.create_d_md)>(dlsym(g_mkl.so, "DftiCreateDescriptor_d_md"));
  g_mkl.set_value = reinterpret_cast<decltype(g_mkl.set_value)>(dlsym(g_mkl.so, "DftiSetValue"));
  g_mkl.commit = reinterpret_cast<decltype(g_mkl.commit)>(dlsym(g_mkl.so, "DftiCommitDescriptor"));
  g_mkl.forward = reinterpret_cast<decltype(g_mkl.forward)>(dlsym(g_mkl.so, "DftiComputeForward"));
  g_mkl.backward = reinterpret_cast<decltype(g_mkl.backward)>(dlsym(g_mkl.so, "DftiComputeBackward"));
  g_mkl.free_desc = reinterpret_cast<decltype(g_mkl.free_desc)>(dlsym(g_mkl.so, "DftiFreeDescriptor"));
  if (!g_mkl.create_d_md || !g_mkl.set_value || !g_mkl.commit || !g_mkl.forward || !g_mkl.backward ||
      !g_mkl.free_desc) {
    g_err = "libmkl_rt lacks the DFTI C interface";
    dlclose(g_mkl.so);
    g_mkl.so = nullptr;
    return false;
  }
  return true;
}

// The descriptors of MKL_DFTI_Setup.f90 in the C interface's row-major
// terms: lengths {ny, nx}; real rows of nx, complex rows of nx/2+1.
bool make_desc(DftiHandle* h, int nx, int ny, int howmany, bool fwd) {
  MklLong len[2] = {ny, nx};
  if (g_mkl.create_d_md(h, kReal, 2, len)) return false;
  MklLong rs[3] = {0, nx, 1}, cs[3] = {0, nx / 2 + 1, 1};
  bool ok = !g_mkl.set_value(*h, kNumberOfTransforms, (MklLong)howmany) &&
            !g_mkl.set_value(*h, kConjugateEvenStorage, (int)kComplexComplex) &&
            !g_mkl.set_value(*h, kPackedFormat, (int)kCce) && !g_mkl.set_value(*h, kPlacement, (int)kNotInplace) &&
            !g_mkl.set_value(*h, kInputStrides, fwd ? rs : cs) && !g_mkl.set_value(*h, kOutputStrides, fwd ? cs : rs) &&
            !g_mkl.set_value(*h, kInputDistance, (MklLong)(fwd ? (long)nx * ny : (long)(nx / 2 + 1) * ny)) &&
            !g_mkl.set_value(*h, kOutputDistance, (MklLong)(fwd ? (long)(nx / 2 + 1) * ny : (long)nx * ny));
  return ok && !g_mkl.commit(*h);
}

// ARK4(3)6L[2]SA (ARK43.f90:17-57, each a._dp/b._dp in double)
struct Ark {
  double ae[7][7] = {}, ai[7][7] = {}, b[7] = {}, be[7] = {};
  Ark() {
    ae[2][1] = 0.5;
    ae[3][1] = 13861.0 / 62500.0;
    ae[3][2] = 6889.0 / 62500.0;
    ae[4][1] = -116923316275.0 / 2393684061468.0;
    ae[4][2] = -2731218467317.0 / 15368042101831.0;
    ae[4][3] = 9408046702089.0 / 11113171139209.0;
    ae[5][1] = -451086348788.0 / 2902428689909.0;
    ae[5][2] = -2682348792572.0 / 7519795681897.0;
    ae[5][3] = 12662868775082.0 / 11960479115383.0;
    ae[5][4] = 3355817975965.0 / 11060851509271.0;
    ae[6][1] = 647845179188.0 / 3216320057751.0;
    ae[6][2] = 73281519250.0 / 8382639484533.0;
    ae[6][3] = 552539513391.0 / 3454668386233.0;
    ae[6][4] = 3354512671639.0 / 8306763924573.0;
    ae[6][5] = 4040.0 / 17871.0;
    b[1] = 82889.0 / 524892.0;
    b[3] = 15625.0 / 83664.0;
    b[4] = 69875.0 / 102672.0;
    b[5] = -2260.0 / 8211.0;
    b[6] = 0.25;
    ai[2][1] = 0.25;
    ai[3][1] = 8611.0 / 62500.0;
    ai[3][2] = -1743.0 / 31250.0;
    ai[4][1] = 5012029.0 / 34652500.0;
    ai[4][2] = -654441.0 / 2922500.0;
    ai[4][3] = 174375.0 / 388108.0;
    ai[5][1] = 15267082809.0 / 155376265600.0;
    ai[5][2] = -71443401.0 / 120774400.0;
    ai[5][3] = 730878875.0 / 902184768.0;
    ai[5][4] = 2285395.0 / 8070912.0;
    for (int j = 1; j <= 6; ++j) ai[6][j] = b[j];
    be[1] = 31666707.0 / 9881966720.0;
    be[3] = -256875.0 / 105007616.0;
    be[4] = -2768025.0 / 128864768.0;
    be[5] = 169839.0 / 3864644.0;
    be[6] = -5247.0 / 225920.0;
  }
};
const Ark kArk;

}  // namespace

struct qcpu {
  int nx, ny, nz, nxh;
  long plane, spec, phys;   // (ny, nxh) complex, nz planes; nz (ny, nx) reals
  double Lx, Ly, Htot, beta, A2, A8, r0, Cd, TOL, dt_max;
  int expl;
  std::vector<double> H, Modes, EVals, uBar, qyBar;   // Modes[layer * nz + mode]
  std::vector<double> kx, ky, k2;                     // kx (nxh), ky (ny), k2 (plane)
  DftiHandle fwd = nullptr, bwd = nullptr, fwd1 = nullptr;
  std::vector<cd> psi, uh, vh, w1, w2, qtmp, N[7], L[7];
  std::vector<double> qp, up, vp, g1, g2;
  double err0 = 0.0;
  bool first = true;
};

namespace {

void dealias(const qcpu* c, cd* f) {
  // QGN_Module.f90:194-204: kx index >= nx/3, ky rows ny/3 .. 2(ny/3), and (0,0)
  const int ia = c->nx / 3, ja = c->ny / 3, jb = 2 * (c->ny / 3);
#pragma omp parallel for collapse(2)
  for (int z = 0; z < c->nz; ++z)
    for (int j = 0; j < c->ny; ++j) {
      cd* row = f + z * c->plane + (long)j * c->nxh;
      if (j >= ja && j <= jb) {
        for (int i = 0; i < c->nxh; ++i) row[i] = 0.0;
      } else {
        for (int i = ia; i < c->nxh; ++i) row[i] = 0.0;
      }
      if (j == 0) row[0] = 0.0;
    }
}

void c2r(qcpu* c, const cd* in, double* out) { g_mkl.backward(c->bwd, (void*)in, (void*)out); }
void r2c(qcpu* c, const double* in, cd* out) { g_mkl.forward(c->fwd, (void*)in, (void*)out); }

// GetPsi: psi = Modes diag(1/(-k2 + EVals)) (H Modes^T / Htot) q per wavenumber
void get_psi(qcpu* c, const cd* q, cd* psi) {
  const int nz = c->nz;
#pragma omp parallel
  {
    std::vector<cd> qm(nz);
#pragma omp for
    for (long p = 0; p < c->plane; ++p) {
      for (int k = 0; k < nz; ++k) {
        cd s = 0.0;
        for (int kk = 0; kk < nz; ++kk) s += c->H[kk] * c->Modes[kk * nz + k] * q[kk * c->plane + p];
        qm[k] = (p == 0) ? cd(0.0) : s / c->Htot / (-c->k2[p] + c->EVals[k]);
      }
      for (int k = 0; k < nz; ++k) {
        cd s = 0.0;
        for (int kk = 0; kk < nz; ++kk) s += c->Modes[k * nz + kk] * qm[kk];
        psi[k * c->plane + p] = s;
      }
    }
  }
}

void get_rhs(qcpu* c, const cd* q, cd* rhs) {
  const int nz = c->nz, nxh = c->nxh, ny = c->ny;
  const long plane = c->plane;
  const double nn = 1.0 / ((double)c->nx * c->ny);
  cd* psi = c->psi.data();
  get_psi(c, q, psi);
  // u = -i ky psi, v = i kx psi  (QGN_Module.f90:222-227)
#pragma omp parallel for collapse(2)
  for (int z = 0; z < nz; ++z)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nxh; ++i) {
        const long e = z * plane + (long)j * nxh + i;
        c->uh[e] = cd(0.0, -c->ky[j]) * psi[e];
        c->vh[e] = cd(0.0, c->kx[i]) * psi[e];
      }
  c2r(c, q, c->qp.data());
  c2r(c, c->uh.data(), c->up.data());
  c2r(c, c->vh.data(), c->vp.data());
#pragma omp parallel for
  for (long e = 0; e < c->phys; ++e) {
    c->qp[e] *= nn;
    c->up[e] *= nn;
    c->vp[e] *= nn;
    c->g1[e] = c->up[e] * c->qp[e];
    c->g2[e] = c->vp[e] * c->qp[e];
  }
  r2c(c, c->g1.data(), c->w1.data());
  r2c(c, c->g2.data(), c->w2.data());
  // RHS = -jaco - uBar i kx q - (beta + qyBar) v  [- (A2 k2 + A8 k8) q] (:266-280)
#pragma omp parallel for collapse(2)
  for (int z = 0; z < nz; ++z)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nxh; ++i) {
        const long p = (long)j * nxh + i, e = z * plane + p;
        const cd jac = cd(0.0, c->kx[i]) * c->w1[e] + cd(0.0, c->ky[j]) * c->w2[e];
        cd r = -jac - c->uBar[z] * cd(0.0, c->kx[i]) * q[e] - (c->beta + c->qyBar[z]) * c->vh[e];
        if (c->expl) {
          const double k2 = c->k2[p], k4 = k2 * k2;
          r -= (c->A2 * k2 + c->A8 * k4 * k4) * q[e];
        }
        rhs[e] = r;
      }
  // Ekman on layer nz (:282)
  const double ek = c->r0 * c->Htot / c->H[nz - 1];
  cd* rb = rhs + (long)(nz - 1) * plane;
  const cd* pb = psi + (long)(nz - 1) * plane;
#pragma omp parallel for
  for (long p = 0; p < plane; ++p) rb[p] += ek * c->k2[p] * pb[p];
  // quadratic drag from layer nz only (:284-292; the other layers' transforms are unused)
  if (c->Cd != 0.0) {
    const long nxy = (long)c->nx * ny;
    const double* ub = c->up.data() + (long)(nz - 1) * nxy;
    const double* vb = c->vp.data() + (long)(nz - 1) * nxy;
#pragma omp parallel for
    for (long e = 0; e < nxy; ++e) {
      const double urms = std::sqrt(ub[e] * ub[e] + vb[e] * vb[e]);
      c->g1[e] = urms * ub[e];
      c->g2[e] = urms * vb[e];
    }
    g_mkl.forward(c->fwd1, c->g1.data(), c->w1.data());
    g_mkl.forward(c->fwd1, c->g2.data(), c->w2.data());
    const double cdr = c->Cd * c->Htot / c->H[nz - 1];
#pragma omp parallel for collapse(2)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nxh; ++i) {
        const long p = (long)j * nxh + i;
        rb[p] += cdr * (cd(0.0, c->ky[j]) * c->w1[p] - cd(0.0, c->kx[i]) * c->w2[p]);
      }
  }
  dealias(c, rhs);
}

// A8L(k) = (1.3 Lx/(nx pi))^12 sqrt(sum q d / (nx ny)), d = c2r(k8 q)/(nx ny),
// the 1.3 a default-real literal (ARK43_QGLeith.f90:62-73)
void leith(qcpu* c, const cd* q, double* a8l) {
  const double nn = 1.0 / ((double)c->nx * c->ny);
  cd* d = c->qtmp.data();
#pragma omp parallel for
  for (long e = 0; e < c->spec; ++e) {
    const double k2 = c->k2[e % c->plane], k4 = k2 * k2;
    d[e] = (k4 * k4) * q[e];
  }
  c2r(c, d, c->g1.data());
  c2r(c, q, c->g2.data());
  const double base = (double)1.3f * c->Lx / ((double)c->nx * 3.14159265358979323846);
  const double b2 = base * base, b4 = b2 * b2, c12 = b4 * b4 * b4;
  const long nxy = (long)c->nx * c->ny;
  for (int z = 0; z < c->nz; ++z) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s)
    for (long e = 0; e < nxy; ++e) s += (c->g1[z * nxy + e] * nn) * (c->g2[z * nxy + e] * nn);
    a8l[z] = c12 * std::sqrt(s * nn);
  }
}

int ark43_step(qcpu* c, cd* q, double* t, double* dt, bool qgleith, long* rejects) {
  const int nz = c->nz;
  const long plane = c->plane, spec = c->spec;
  if (c->first) {
    c->err0 = c->TOL;
    c->first = false;
  }
  std::vector<double> a8(nz, c->A8);
  if (qgleith) leith(c, q, a8.data());
  auto lam = [&](long e) {
    const double k2 = c->k2[e % plane], k4 = k2 * k2;
    return c->A2 * k2 + a8[e / plane] * (k4 * k4);
  };
  get_rhs(c, q, c->N[1].data());
#pragma omp parallel for
  for (long e = 0; e < spec; ++e) c->L[1][e] = -lam(e) * q[e];
  const double reject_factor = qgleith ? 0.5 : 0.75, safety = qgleith ? 0.2 : 0.75;
  while (true) {
    const double h = *dt;
    for (int i = 2; i <= 6; ++i) {
#pragma omp parallel for
      for (long e = 0; e < spec; ++e) {
        cd acc = 0.0;
        for (int j = 1; j < i; ++j) acc += kArk.ae[i][j] * c->N[j][e] + kArk.ai[i][j] * c->L[j][e];
        c->qtmp[e] = (q[e] + h * acc) / (1.0 + 0.25 * h * lam(e));
      }
      get_rhs(c, c->qtmp.data(), c->N[i].data());
#pragma omp parallel for
      for (long e = 0; e < spec; ++e) c->L[i][e] = -lam(e) * c->qtmp[e];
    }
    // error over all layers (ARK43_QGLeith.f90:121-124)
#pragma omp parallel for
    for (long e = 0; e < spec; ++e) {
      cd s = 0.0;
      for (int j = 1; j <= 6; ++j)
        if (kArk.be[j] != 0.0) s += kArk.be[j] * (c->N[j][e] + c->L[j][e]);
      c->qtmp[e] = s;
    }
    c2r(c, c->qtmp.data(), c->g1.data());
    double mx = 0.0;
#pragma omp parallel for reduction(max : mx)
    for (long e = 0; e < c->phys; ++e) mx = std::fmax(mx, std::fabs(c->g1[e]));
    const double err1 = h * mx / ((double)c->nx * c->ny);
    if (!std::isfinite(err1)) return -4;
    if (err1 > c->TOL) {
      *dt = reject_factor * h;
      if (rejects) ++*rejects;
      continue;
    }
#pragma omp parallel for
    for (long e = 0; e < spec; ++e) {
      cd s = 0.0;
      for (int j = 1; j <= 6; ++j)
        if (kArk.b[j] != 0.0) s += kArk.b[j] * (c->N[j][e] + c->L[j][e]);
      q[e] += h * s;
    }
    for (int z = 0; z < nz; ++z) q[z * plane] = 0.0;
    *t += h;
    *dt = std::fmin(h * std::pow(safety * c->TOL / err1, 0.075) * std::pow(c->err0 / err1, 0.1), c->dt_max);
    c->err0 = err1;
    dealias(c, q);   // TimeStep's DeAlias (QGN_Module.f90:307)
    return 0;
  }
}

}  // namespace

extern "C" {

const char* qcpu_last_error(void) { return g_err.c_str(); }

void qcpu_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

int qcpu_max_threads(void) { return omp_get_max_threads(); }

qcpu* qcpu_create(int nx, int ny, int nz, double Lx, double Ly, double Htot, double beta, double A2, double A8,
                  double r0, double Cd, double TOL, double dt_max, int is_explicit, const double* H,
                  const double* Modes, const double* EVals, const double* uBar, const double* qyBar) {
  if (!load_mkl()) return nullptr;
  qcpu* c = new qcpu();
  c->nx = nx; c->ny = ny; c->nz = nz; c->nxh = nx / 2 + 1;
  c->plane = (long)ny * c->nxh;
  c->spec = c->plane * nz;
  c->phys = (long)nx * ny * nz;
  c->Lx = Lx; c->Ly = Ly; c->Htot = Htot; c->beta = beta; c->A2 = A2; c->A8 = A8; c->r0 = r0; c->Cd = Cd;
  c->TOL = TOL; c->dt_max = dt_max; c->expl = is_explicit;
  c->H.assign(H, H + nz);
  c->Modes.assign(Modes, Modes + (long)nz * nz);
  c->EVals.assign(EVals, EVals + nz);
  c->uBar.assign(uBar, uBar + nz);
  c->qyBar.assign(qyBar, qyBar + nz);
  // wavenumbers (QGN_Module.f90:161-172)
  const double pi = 4.0 * std::atan(1.0);
  c->kx.resize(c->nxh);
  c->ky.resize(ny);
  for (int i = 0; i < c->nxh; ++i) c->kx[i] = (2.0 * pi / Lx) * i;
  for (int j = 0; j < ny; ++j) c->ky[j] = (2.0 * pi / Ly) * (j <= ny / 2 ? j : j - ny);
  c->k2.resize(c->plane);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < c->nxh; ++i) c->k2[(long)j * c->nxh + i] = c->kx[i] * c->kx[i] + c->ky[j] * c->ky[j];
  if (!make_desc(&c->fwd, nx, ny, nz, true) || !make_desc(&c->bwd, nx, ny, nz, false) ||
      !make_desc(&c->fwd1, nx, ny, 1, true)) {
    g_err = "DFTI descriptor set-up failed";
    for (DftiHandle* h : {&c->fwd, &c->bwd, &c->fwd1})
      if (*h) g_mkl.free_desc(h);
    delete c;
    return nullptr;
  }
  for (auto* v : {&c->psi, &c->uh, &c->vh, &c->w1, &c->w2, &c->qtmp}) v->assign(c->spec, 0.0);
  for (int i = 1; i <= 6; ++i) {
    c->N[i].assign(c->spec, 0.0);
    c->L[i].assign(c->spec, 0.0);
  }
  for (auto* v : {&c->qp, &c->up, &c->vp, &c->g1, &c->g2}) v->assign(c->phys, 0.0);
  return c;
}

void qcpu_destroy(qcpu* c) {
  if (!c) return;
  for (DftiHandle* h : {&c->fwd, &c->bwd, &c->fwd1})
    if (*h) g_mkl.free_desc(h);
  delete c;
}

// GetRHS of q_hat (complex interleaved, (nz, ny, nx/2+1)) into rhs
int qcpu_rhs(qcpu* c, const double* q_hat, double* rhs) {
  get_rhs(c, reinterpret_cast<const cd*>(q_hat), reinterpret_cast<cd*>(rhs));
  return 0;
}

// One accepted ARK43 (qgleith = 0) or ARK43_QGLeith step of q_hat in place;
// the stepper's saved err0 persists in the handle as the reference's `save`
int qcpu_ark43_step(qcpu* c, double* q_hat, double* t, double* dt, int qgleith, long* rejects) {
  return ark43_step(c, reinterpret_cast<cd*>(q_hat), t, dt, qgleith != 0, rejects);
}

}  // extern "C"
